"""The laser path on the GPU against the oracle: the PGC pushers
(species push_type "robust_pgc" / "standard_pgc", part2d_class.f03:1012-1602
deposits, :1967-2219 push), the slice's envelope and its gradients
(laser_all, simulation_class.f03:362-367, field_laser_class.f03:637-750)
and the plasma susceptibility chi (sim_lasers%deposit_chi,
sim_lasers_class.f03:176-200, part2d_class.f03:361-476).  The envelope is a
synthetic Gaussian x sin^2 pulse with a wavefront phase and m = 1 content
(the laser solver that would evolve it is not on the path); both sides get
the same history."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
import twin
from conftest import rel_err
from qpad_amd import Simulation, configs
from qpad_amd import _lib as L

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _deck(push_type, ion_push_type=None, max_mode=1, ions=False, nr=48, nxi=40):
    return configs.small(nr=nr, nxi=nxi, max_mode=max_mode, ppc=(2, 4), total_num=1 << 12, ions=ions,
                         ion_ppc=(1, 2), offset_x=0.03, push_type=push_type, ion_push_type=ion_push_type)


@pytest.mark.parametrize("push_type,ion_push_type,max_mode,a0", [
    ("robust_pgc", None, 1, 1.0),
    ("standard_pgc", "robust", 2, 0.8),
    ("robust_pgc", "standard_pgc", 2, 1.2),
])
def test_pgc_slices_match_oracle(push_type, ion_push_type, max_mode, a0):
    ions = ion_push_type is not None
    deck = _deck(push_type, ion_push_type, max_mode, ions)
    print("pgc slices:", push_type, ion_push_type, "max_mode", max_mode, "a0", a0, flush=True)
    sim = Simulation(deck)
    o = twin.make_twin(sim)
    ar, ai = twin.laser_envelope(sim.grid, sim.nm, a0=a0)
    sim.set_laser(ar, ai)
    o.set_laser(ar, ai)
    sim.step3d_begin()
    o.step_begin()
    norm = twin.ErrorNorm()
    for j in range(1, sim.grid.nxi + 1):
        twin.sync_slice_inputs(sim, o, j)
        it_g, it_o = sim.slice(j), o.slice(j)
        assert it_g == it_o, (j, it_g, it_o)
        twin.compare_fields(sim, o, norm, j)
        twin.compare_species(sim, o, norm, j)
        twin.compare_laser(sim, o, norm, j)
    print({n: f"{e:.1e}" for n, e in sorted(norm.errors().items())})
    # the laser drove the plasma: the susceptibility and the transverse
    # current are far from the laser-free state
    assert np.max(np.abs(o.laser(po.LASER_H_CHI))) > 0.5
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"
    sim.close()
    o.close()


def test_pgc_step_and_beam_push_match_oracle():
    """a whole 3D step from the deck alone (each side injects its own plasma
    and beam), histories and chi compared norm-wise, then the beam push"""
    deck = _deck("robust_pgc", "standard_pgc", max_mode=1, ions=True, nr=40, nxi=32)
    sim = Simulation(deck)
    o = twin.oracle_from_deck(deck)
    ar, ai = twin.laser_envelope(sim.grid, sim.nm, a0=1.0, length=0.4)
    sim.set_laser(ar, ai)
    o.set_laser(ar, ai)
    sim.step3d()
    o.step_begin()
    for j in range(1, sim.grid.nxi + 1):
        o.slice(j)
    o.step_end(1)
    nxi = sim.grid.nxi
    for name in ("psi", "e", "b", "b_spe", "cu", "q_spe"):
        g = np.stack([sim.fields[name].hist(j) for j in range(1, nxi + 1)])
        r = o.hist(twin.ORA_ID[name])[1:nxi + 1]
        e = rel_err(g, r)
        assert e <= TOL, (name, e)
    g = np.stack([sim.fields["chi"].hist(j) for j in range(1, nxi + 1)])
    e = rel_err(g, o.laser(po.LASER_H_CHI)[1:nxi + 1])
    assert e <= TOL, ("chi", e)
    gb, ob = sim.beam[0].get(), o.beam(0)
    assert len(gb["x1"]) == len(ob["x1"])
    for k in ("x1", "x2", "x3", "p1", "p2", "p3"):
        assert rel_err(np.sort(gb[k]), np.sort(ob[k])) <= 1e-9, k
    sim.close()
    o.close()


def test_laser_entry_points_match_oracle():
    """qpad_laser_slice and qpad_deposit_chi alone (the fine-grained API a
    QPAD driver calls): the envelope and gradients of several slices, and
    chi of two species on identical particles"""
    deck = _deck("robust_pgc", "robust", max_mode=2, ions=True)
    sim = Simulation(deck)
    o = twin.make_twin(sim)
    ar, ai = twin.laser_envelope(sim.grid, sim.nm, a0=0.9, curv=0.5)
    sim.set_laser(ar, ai)
    o.set_laser(ar, ai)
    for j in (1, 2, 17, sim.grid.nxi):
        sim.laser_slice(j)
        o.laser_slice(j)
        for name, which in twin.LASER_FIELDS[:4]:
            e = rel_err(sim.fields[name].get(), o.laser(which))
            assert e <= 1e-13, (j, name, e)
    # chi on the particles as injected (psi from the injection), slice 5
    for k, sp in enumerate(sim.spe):
        o.set_species(k, sp.get())
    sim.deposit_chi(5)
    chi = o.laser(po.LASER_CHI)
    chi[...] = 0.0
    for k in range(len(sim.spe)):
        o.L.ora_deposit_chi(o.h, o.species_raw(k), po.ptr(chi[0]))
    e = rel_err(sim.fields["chi"].get(), chi)
    assert e <= 1e-13, ("chi", e)
    assert rel_err(sim.fields["chi"].hist(5), chi) <= 1e-13
    sim.close()
    o.close()


def test_pgc_species_without_laser_pushes_in_zero_envelope():
    """a PGC species before any qpad_set_laser: a zero envelope (QPAD with
    nlasers = 0), chi still deposited"""
    deck = _deck("robust_pgc", max_mode=1, nr=40, nxi=24)
    sim = Simulation(deck)
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    norm = twin.ErrorNorm()
    for j in range(1, sim.grid.nxi + 1):
        twin.sync_slice_inputs(sim, o, j)
        assert sim.slice(j) == o.slice(j)
        twin.compare_fields(sim, o, norm, j)
        twin.compare_species(sim, o, norm, j)
        twin.compare_laser(sim, o, norm, j)
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"
    sim.close()
    o.close()


def test_laser_entry_points_refuse_without_laser():
    sim = Simulation(_deck("robust", nr=40, nxi=16))
    assert sim.lib.qpad_laser_slice(sim.ctx, 1) == L.E_STATE
    assert sim.lib.qpad_deposit_chi(sim.ctx, 1) == L.E_STATE
    with pytest.raises(L.QpadError):
        sim.set_laser(np.zeros((3, 3, 3)), np.zeros((3, 3, 3)))
    sim.close()
