"""qpad_species_qdp_noalloc against qpad_species_qdp, through the C ABI.

The no-allocation variant keeps the charge deposit's per-cell partial rows in
a static device array shared by every call on the GPU.  Its result must be
bit for bit that of qpad_species_qdp (the same kernels in the same order),
on two calls in a row where the second deposits fewer particles: rows the
first call wrote and the second does not (overflow rows of crowded cells)
must not leak into the second result.  Each call is also checked against
the oracle's qdeposit (part2d_class.f03:231-359) on the same particles."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
import twin
from conftest import rel_err
from qpad_amd import Simulation, configs
from qpad_amd import _lib as L

pytestmark = pytest.mark.gpu


def _q_spe(sim, entry):
    """q_spe after zeroing it and depositing species 0 with `entry`."""
    q = sim.fields["q_spe"]
    q.set(np.zeros(q.shape()))
    L.check(getattr(sim.lib, entry)(sim.ctx, C.c_int32(0)))
    return q.get()


def _oracle_q(sim, o, arrs):
    o.set_species(0, arrs)
    q = o.new_field(1)
    po.lib().ora_qdeposit(o.h, o.species_raw(0), po.ptr(q))
    return q + sim.spe[0].qn()


def _crowded(g, per_cell, cells, rng):
    """per_cell particles in each of `cells` radial cells (uniform in r
    within the cell and in phi), at rest, unit-ish charges."""
    r = np.concatenate([(c + rng.uniform(0.02, 0.98, per_cell)) * g.dr for c in cells])
    ph = rng.uniform(0.0, 2 * np.pi, r.size)
    n = r.size
    return {"x1": r * np.cos(ph), "x2": r * np.sin(ph), "p1": np.zeros(n), "p2": np.zeros(n),
            "p3": np.zeros(n), "q": -rng.uniform(0.5, 1.5, n) * 1e-3, "gamma": np.ones(n),
            "psi": np.zeros(n), "id": np.arange(n, dtype=np.int32)}


def test_noalloc_qdp_matches_qdp_on_two_calls():
    # ppc 10 x 30: room (npmax = 24,000) for 2600 particles in each of three cells
    sim = Simulation(configs.small(nr=40, nxi=4, max_mode=2, ppc=(10, 30), total_num=1 << 10))
    o = twin.make_twin(sim)
    g = sim.grid
    rng = np.random.default_rng(11)
    # call 1: three cells crowded past two deposit chunks (overflow rows)
    # among lightly filled ones
    big = _crowded(g, 2600, [3, 17, 31], rng)
    light = _crowded(g, 9, range(g.nr), rng)
    first = {k: np.concatenate([big[k], light[k]]) for k in big}
    first["id"] = np.arange(first["x1"].size, dtype=np.int32)
    # call 2: fewer particles, no crowded cell
    second = _crowded(g, 5, range(0, g.nr, 2), rng)
    for arrs in (first, second):
        sim.spe[0].set(arrs)
        a = _q_spe(sim, "qpad_species_qdp")
        b = _q_spe(sim, "qpad_species_qdp_noalloc")
        assert np.array_equal(a, b), f"{np.max(np.abs(a - b))} at n = {arrs['x1'].size}"
        # the same deposit restated on the CPU
        ref = _oracle_q(sim, o, arrs)
        assert rel_err(b[0, :, 1:g.nr + 1], ref[0, :, 1:g.nr + 1]) <= 1e-12
    sim.close()


def test_noalloc_qdp_refuses_bad_species():
    sim = Simulation(configs.small(nr=16, nxi=4, max_mode=1, ppc=(1, 2), total_num=1 << 8))
    assert sim.lib.qpad_species_qdp_noalloc(sim.ctx, C.c_int32(3)) == L.E_INVALID
    assert b"bad species" in sim.lib.qpad_last_error()
    # the context is intact: the variant still matches the plain entry point
    assert np.array_equal(_q_spe(sim, "qpad_species_qdp"), _q_spe(sim, "qpad_species_qdp_noalloc"))
    sim.close()
