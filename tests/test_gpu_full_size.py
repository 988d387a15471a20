"""Parity at the benchmark sizes.

* C2 and C3 slices inside the beam, and C2 slices deep in the wake,
  compared with the oracle on identical inputs (a few slices: the serial
  oracle takes ~0.3 s per C2 slice), each against the oracle's own
  sensitivity to a round-off change of its input where that is larger;
* size-independent properties over whole C2 3D steps: particles stay sorted
  by radial cell with consistent cell starts, the particle count is
  conserved (inside + left the box), and a rerun is bit-identical.
"""
import numpy as np
import pytest

import twin
from qpad_amd import Simulation, configs

pytestmark = pytest.mark.gpu

TOL = 1e-10


def run_to(sim, j_end):
    for j in range(1, j_end):
        sim.slice(j)


# C2 slices 520 and 900 are deep in the wake, where the iteration runs to
# iter_max and does not contract: there a 1e-15 change of the slice's input
# moves the oracle's own result by up to 1e-4 (particles) and 1e-6 (fields).
# Each slice is therefore held to TOL or to SENS_FACTOR times the oracle's
# own sensitivity to such a change, whichever is larger; in the beam and
# ahead of the wake that is TOL.
SENS_FACTOR = 10.0


@pytest.mark.parametrize("name,j0,n", [("c2", 100, 3), ("c2", 520, 1), ("c2", 900, 1), ("c3", 100, 2)])
def test_full_size_slices_match_oracle(name, j0, n):
    deck = getattr(configs, name)()
    sim = Simulation(deck)
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    run_to(sim, j0)
    norm = twin.ErrorNorm()
    bound = {}
    for j in range(j0, j0 + n):
        twin.sync_slice_inputs(sim, o, j)
        it_o, sens = twin.oracle_self_sensitivity(o, j, len(sim.spe))
        it_g = sim.slice(j)
        assert it_g == it_o, f"{name} slice {j}: {it_g} iterations (GPU) vs {it_o} (oracle)"
        twin.compare_fields(sim, o, norm, j)
        twin.compare_species(sim, o, norm, j)
        for q, e in sens.errors().items():
            bound[q] = max(bound.get(q, TOL), SENS_FACTOR * e)
    sim.close()
    errs = norm.errors()
    print(name, j0, {q: f"{e:.1e} (bound {bound[q]:.1e})" for q, e in sorted(errs.items())})
    bad = {q: (e, bound[q]) for q, e in errs.items() if not e <= bound[q]}
    assert not bad, f"{name}: beyond the bound: {bad}"


def cell_keys(p, g):
    r = np.hypot(p["x1"], p["x2"])
    key = np.floor(r / g.dr).astype(np.int64)
    key = np.minimum(key, g.nr - 1)
    key[~(r < g.nr * g.dr)] = g.nr
    return key


def test_c2_step_keeps_cell_order_and_particles():
    """Live particles stay sorted by radial cell, inside the box, each once,
    and the only ones that go are those that crossed the outer boundary
    (update_bound): a particle missing at slice j+1 was, at slice j, within
    one slice's largest radial step of r_max."""
    sim = Simulation(configs.c2())
    g = sim.grid
    rmax = g.nr * g.dr
    ids0 = np.sort(sim.spe[0].get()["id"])
    sim.step3d_begin()
    prev = None
    for j in range(1, g.nxi + 1):
        sim.slice(j)
        if j in (1, 299, 300, 699, 700, g.nxi):
            p = sim.spe[0].get()
            key = cell_keys(p, g)
            assert np.all(np.diff(key) >= 0), f"slice {j}: particles not sorted by cell"
            assert np.all(key < g.nr), f"slice {j}: a live particle is outside the box"
            ids = np.sort(p["id"])
            assert np.all(np.diff(ids) > 0), f"slice {j}: duplicated particle"
            assert np.all(np.isin(ids, ids0)), f"slice {j}: unknown particle"
            if prev is not None and prev[0] == j - 1:
                a, b = prev[1], twin.particles_by_id(p)
                ra = np.hypot(a["x1"], a["x2"])
                kept = np.isin(a["id"], b["id"])
                rb = np.hypot(b["x1"], b["x2"])
                step = np.max(np.abs(rb - ra[kept]))
                lost = ra[~kept]
                assert np.all(lost >= rmax - 2.0 * step - g.dr), \
                    f"slice {j}: a particle vanished away from the boundary"
            prev = (j, twin.particles_by_id(p))
    sim.step3d_end()
    sim.close()


def test_c2_rerun_is_bit_identical():
    def run():
        sim = Simulation(configs.c2())
        sim.step3d()
        sim.step3d_begin()
        for j in range(1, 201):
            sim.slice(j)
        out = {n: sim.fields[n].get() for n in ("psi", "e", "b", "b_spe", "cu", "q_spe")}
        out.update({"sp." + k: v for k, v in sim.spe[0].get().items()})
        out.update({"beam." + k: v for k, v in sim.beam[0].get().items()})
        sim.close()
        return out

    a = run()
    b = run()
    for k in a:
        assert np.array_equal(a[k], b[k]), f"{k} differs between identical runs"
