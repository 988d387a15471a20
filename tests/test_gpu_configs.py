"""The benchmark workloads themselves on the GPU, at their full sizes.

C1 (Nr 256, Nxi 256, modes 0..1, 4,096 plasma particles per slice, 2^18 beam
particles): a whole 3D step compared with the oracle slice by slice on
identical inputs, then the beam push.

C4 (16 independent instances, Nr in {512 .. 4096}, modes 0..1 to 0..4, Nxi
512): one 3D step of the whole scan as an eager batch, every instance bit
for bit equal to the same deck stepped alone; and one slice inside the wake
of its Nr 4096, modes 0..4 instance compared with the oracle.

QPAD's loop: simulation_class.f03:294-512 (3D step), :342-469 (slice)."""
import numpy as np
import pytest

import twin
from conftest import rel_err
from qpad_amd import Batch, Simulation, configs

pytestmark = pytest.mark.gpu

TOL = 1e-10


def test_c1_step_matches_oracle_slice_by_slice():
    sim = Simulation(configs.c1())
    g = sim.grid
    assert (g.nr, g.nxi, g.max_mode) == (256, 256, 1)
    assert sim.spe[0].count() == 256 * 2 * 8 and sim.beam[0].count() == 1 << 18
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    qb = max(rel_err(sim.fields["q_beam"].hist(s), o.hist(twin.po.Q_BEAM)[s]) for s in range(1, g.nxi + 1))
    assert qb <= TOL, f"q_beam history {qb}"
    norm = twin.ErrorNorm()
    iters = []
    for j in range(1, g.nxi + 1):
        twin.sync_slice_inputs(sim, o, j)
        it_g, it_o = sim.slice(j), o.slice(j)
        iters.append((it_g, it_o))
        twin.compare_fields(sim, o, norm, j)
        if j % 16 == 0 or j == g.nxi:
            twin.compare_species(sim, o, norm, j)
    print({n: f"{e:.1e}" for n, e in sorted(norm.errors().items())})
    print("iterations per slice:", np.mean([a for a, _ in iters]))
    assert all(a == b for a, b in iters), [(j + 1, a, b) for j, (a, b) in enumerate(iters) if a != b]
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"
    # the beam push on each side's own field histories
    sim.step3d_end()
    o.step_end(1)
    gb, ob = sim.beam[0].get(), o.beam(0)
    assert len(gb["x1"]) == len(ob["x1"])
    for k in ("x1", "x2", "x3", "p1", "p2", "p3"):
        e = rel_err(np.sort(gb[k]), np.sort(ob[k]))
        assert e <= 1e-9, (k, e)
    sim.close()


def _state(sim, every=8):
    nxi = sim.grid.nxi
    out = {n: np.stack([sim.fields[n].hist(j) for j in range(1, nxi + 1, every)])
           for n in ("psi", "b_spe", "e", "cu")}
    for k, sp in enumerate(sim.spe):
        out.update({f"sp{k}.{n}": v for n, v in sp.get().items()})
    out.update({"beam." + n: v for n, v in sim.beam[0].get().items()})
    return out


def test_c4_scan_batch_matches_instances_alone():
    decks, params = configs.c4()
    assert len(decks) == 16
    sims = [Simulation(d) for d in decks]
    batch = Batch(sims, graph="eager")
    its = batch.step3d()
    batch.close()
    got = []
    for s_ in sims:
        got.append(_state(s_))
        s_.close()
    for i, d in enumerate(decks):
        alone = Simulation(d)
        it = alone.step3d()
        ref = _state(alone)
        alone.close()
        assert its[i] == it, f"instance {i} ({params[i]}): {its[i]} vs {it} iterations"
        for k in ref:
            assert np.array_equal(got[i][k], ref[k]), f"instance {i} ({params[i]}): {k} differs"
    print("iterations per slice:", [round(int(t) / 512, 2) for t in its])


def test_c4_largest_instance_slice_matches_oracle():
    decks, params = configs.c4()
    i = next(k for k, p in enumerate(params) if p["nr"] == 4096 and p["max_mode"] == 4)
    sim = Simulation(decks[i])
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    j0 = 150   # behind the beam centre (slice 125): plasma response of every mode
    for j in range(1, j0):
        sim.slice(j)
    norm = twin.ErrorNorm()
    for j in (j0, j0 + 1):
        twin.sync_slice_inputs(sim, o, j)
        it_g, it_o = sim.slice(j), o.slice(j)
        assert it_g == it_o, (j, it_g, it_o)
        twin.compare_fields(sim, o, norm, j)
        twin.compare_species(sim, o, norm, j)
    sim.close()
    print(params[i], {n: f"{e:.1e}" for n, e in sorted(norm.errors().items())})
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"


def test_tri_gaussian_deck_lattice_beam_and_step_match_oracle():
    """QPAD's blowout-regime example deck (lattice beam with quiet start,
    input_file/blowout_regime/qpinput_tri-gaussian.json): the library
    injects the same 4.4 million beam particles as the oracle's restatement
    of fdist3d_std_class.f03:439-613, and one 3D step agrees slice by slice,
    then through the beam push."""
    deck = configs.tri_gaussian()
    sim = Simulation(deck)
    ob = twin.oracle_from_deck(deck)
    gb, rb = sim.beam[0].get(), ob.beam(0)
    assert len(gb["x1"]) == len(rb["x1"]) == 4405856
    for k in ("x1", "x2", "x3", "p1", "p2", "p3", "q"):
        e = rel_err(np.sort(gb[k]), np.sort(rb[k]))
        assert e <= 1e-12, (k, e)
    ob.close()
    g = sim.grid
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    norm = twin.ErrorNorm()
    iters = []
    for j in range(1, g.nxi + 1):
        twin.sync_slice_inputs(sim, o, j)
        it_g, it_o = sim.slice(j), o.slice(j)
        iters.append((it_g, it_o))
        twin.compare_fields(sim, o, norm, j)
        if j % 50 == 0 or j == g.nxi:
            twin.compare_species(sim, o, norm, j)
    print({n: f"{e:.1e}" for n, e in sorted(norm.errors().items())})
    print("iterations per slice:", np.mean([a for a, _ in iters]))
    assert all(a == b for a, b in iters), [(j + 1, a, b) for j, (a, b) in enumerate(iters) if a != b]
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"
    sim.step3d_end()
    o.step_end(1)
    gb, ob = sim.beam[0].get(), o.beam(0)
    assert len(gb["x1"]) == len(ob["x1"])
    for k in ("x1", "x2", "x3", "p1", "p2", "p3"):
        e = rel_err(np.sort(gb[k]), np.sort(ob[k]))
        assert e <= 1e-9, (k, e)
    sim.close()


def test_file_beam_deck_matches_oracle(tmp_path):
    """A "file" beam (fdist3d_file_class.f03:154-343): particles written to
    an HDF5 file in other units (micrometres, charge in other units, x3
    growing toward the head, a few outside the box), read back by the
    library, placed and filtered as QPAD places them; the oracle reads the
    same file with its own reader and placement.  Same particles, then one
    3D step slice by slice and the beam push."""
    import h5read
    from qpad_amd.input import grid_from_deck
    src = Simulation(configs.small())
    b = src.beam[0].get()
    src.close()
    deck = configs.small()
    _, _, z0 = grid_from_deck(deck)
    bc, fc, lconv, qconv = (0.03, -0.02, z0 + 1.1), (5.0, 7.0, 300.0), 1e-2, 1e-3
    raw = {"x1": fc[0] + (b["x1"] - bc[0]) / lconv, "x2": fc[1] + (b["x2"] - bc[1]) / lconv,
           "x3": fc[2] - (b["x3"] - (bc[2] - z0)) / lconv, "p1": b["p1"], "p2": b["p2"], "p3": b["p3"],
           "q": b["q"] / qconv}
    n_in = len(raw["x1"])
    # box r < 4, 0 <= xi < 2.4 (xi of the file centre 1.1): r 4.53 dropped,
    # xi 3.1 dropped, xi -0.1 dropped, xi 2.3 kept
    extra = {"x1": [fc[0] + 4.5 / lconv, fc[0], fc[0], fc[0]], "x2": [fc[1]] * 4,
             "x3": [fc[2], fc[2] - 2.0 / lconv, fc[2] + 1.2 / lconv, fc[2] - 1.2 / lconv], "p1": [0.0] * 4,
             "p2": [0.0] * 4, "p3": [2000.0] * 4, "q": [-1.0] * 4}
    raw = {k: np.concatenate([raw[k], extra[k]]).reshape(-1, 1) for k in raw}
    path = str(tmp_path / "beam.h5")
    h5read.write_datasets(path, raw)
    deck["beam"][0] = {"profile_type": "file", "filename": path, "beam_center": list(bc), "file_center": list(fc),
                       "length_conv_fac": lconv, "charge_conv_fac": qconv, "q": -1.0, "m": 1.0,
                       "evolution": True}
    sim = Simulation(deck)
    ob = twin.oracle_from_deck(deck)
    gb, rb = sim.beam[0].get(), ob.beam(0)
    assert len(gb["x1"]) == len(rb["x1"]) == n_in + 1
    for k in ("x1", "x2", "x3", "p1", "p2", "p3", "q"):
        assert np.array_equal(np.sort(gb[k]), np.sort(rb[k])), k
    ob.close()
    g = sim.grid
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    norm = twin.ErrorNorm()
    for j in range(1, g.nxi + 1):
        twin.sync_slice_inputs(sim, o, j)
        assert sim.slice(j) == o.slice(j), j
        twin.compare_fields(sim, o, norm, j)
        twin.compare_species(sim, o, norm, j)
    print({n: f"{e:.1e}" for n, e in sorted(norm.errors().items())})
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"
    sim.step3d_end()
    o.step_end(1)
    gb, ob = sim.beam[0].get(), o.beam(0)
    assert len(gb["x1"]) == len(ob["x1"])
    for k in ("x1", "x2", "x3", "p1", "p2", "p3"):
        e = rel_err(np.sort(gb[k]), np.sort(ob[k]))
        assert e <= 1e-9, (k, e)
    sim.close()
