"""GPU vs CPU oracle, through the C ABI, on identical inputs.

Bar (FP64): every slice field, every history slice written by the slice and
every plasma particle (matched by id) within 1e-10 norm-wise relative error
of the oracle, and the same iteration count, when both start the slice from
the same state.  The GPU reorders sums (per-cell partials, block scans), so
bit equality is not expected; 1e-10 is the physics tolerance.
"""
import numpy as np
import pytest

from conftest import rel_err
import twin

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _sim(deck):
    from qpad_amd import Simulation
    return Simulation(deck)


@pytest.mark.parametrize("max_mode,ions", [(1, False), (2, False), (4, True)])
def test_slice_parity_identical_inputs(oracle_mod, max_mode, ions):
    from qpad_amd import configs
    deck = configs.small(nr=64, nxi=40, max_mode=max_mode, ppc=(2, 4), total_num=1 << 13,
                         offset_x=0.03 if max_mode > 1 else 0.0, ions=ions)
    sim = _sim(deck)
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    # beam charge deposit (whole history) on identical beam particles
    qb_err = max(rel_err(sim.fields["q_beam"].hist(s), o.hist(twin.po.Q_BEAM)[s])
                 for s in range(1, sim.grid.nxi + 1))
    assert qb_err <= TOL, f"q_beam history rel err {qb_err}"
    norm, iters = twin.ErrorNorm(), []
    for j in range(1, sim.grid.nxi + 1):
        twin.sync_slice_inputs(sim, o, j)
        it_g = sim.slice(j)
        it_o = o.slice(j)
        iters.append((j, it_g, it_o))
        twin.compare_fields(sim, o, norm, j)
        twin.compare_species(sim, o, norm, j)
    sim.close()
    print("relative error per field (norm over the run):",
          {n: f"{e:.2e}" for n, e in sorted(norm.errors().items())})
    assert all(a == b for _, a, b in iters), f"iteration counts differ: {iters}"
    bad = norm.bad(TOL)
    assert not bad, f"fields off (rel err, worst slice): {bad}"


def test_injection_matches_oracle(oracle_mod):
    from qpad_amd import configs
    deck = configs.small(nr=48, nxi=8, max_mode=2, ppc=(3, 5), total_num=1 << 10)
    sim = _sim(deck)
    o = twin.make_twin(sim)
    g = twin.particles_by_id(sim.spe[0].get())
    r = twin.particles_by_id({k: np.array(v) for k, v in o.species(0).items()})
    assert np.array_equal(g["id"], r["id"])
    for k in ("x1", "x2", "q", "p1", "p2", "p3", "gamma", "psi"):
        assert rel_err(g[k], r[k]) <= 1e-14, k
    # background qn = -(charge deposit of the fresh lattice)
    _, qn = o.species_q_own(0)
    assert rel_err(sim.spe[0].qn(), qn) <= 1e-13
    sim.close()


def test_full_step_and_beam_push(oracle_mod):
    """One whole 3D step without re-synchronisation, then the beam push."""
    from qpad_amd import configs
    deck = configs.small(nr=48, nxi=32, max_mode=1, ppc=(2, 4), total_num=1 << 12)
    sim = _sim(deck)
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    its_g, its_o = [], []
    for j in range(1, sim.grid.nxi + 1):
        its_g.append(sim.slice(j))
        its_o.append(o.slice(j))
    assert its_g == its_o
    drift = max(rel_err(sim.fields[n].hist(sim.grid.nxi), o.hist(twin.ORA_ID[n])[sim.grid.nxi])
                for n in ("psi", "e", "b"))
    print(f"end-to-end drift over {sim.grid.nxi} slices: {drift:.2e}")
    assert drift <= 1e-8
    sim.step3d_end()
    o.step_end(1)
    gb, ob = sim.beam[0].get(), o.beam(0)
    assert len(gb["x1"]) == len(ob["x1"])
    for k in ("x1", "x2", "x3", "p1", "p2", "p3"):
        assert rel_err(np.sort(gb[k]), np.sort(ob[k])) <= 1e-9, k
    sim.close()


def test_deterministic_rerun():
    """Two runs of the same deck are bit-identical (no atomics in deposits)."""
    from qpad_amd import configs
    deck = configs.small(nr=48, nxi=16, max_mode=2, ppc=(2, 8), total_num=1 << 12)
    out = []
    for _ in range(2):
        sim = _sim(deck)
        sim.step3d_begin()
        for j in range(1, sim.grid.nxi + 1):
            sim.slice(j)
        out.append((sim.fields["e"].hist(sim.grid.nxi), sim.spe[0].get()["x1"]))
        sim.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
