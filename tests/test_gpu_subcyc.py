"""QPAD's sub-cycling algorithm on the GPU against the oracle (simulation
algorithm "subcycling", proj_subcyc/simulation_subcyc_class.f03:230-449):
per slice the largest gamma / (gamma - p_z) sets n sub-cycles of dxi / n,
each a charge deposit, the solves, the iteration and the clamped push
(part2d_subcyc_class.f03:27-66); the neutrals ionise in every sub-cycle."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
import twin
from conftest import rel_err
from qpad_amd import QpadError, Simulation, configs
from qpad_amd.input import subcycling_from_deck

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.mark.parametrize("kw", [
    dict(),                                                          # 1-2 sub-cycles, the clamp active
    dict(max_mode=2, push_type="standard", offset_x=0.05),
    dict(exp_fac_max=0.3, dt_2d_min=0.011, neutral=(3, 3)),           # dt_2d_min binds: 4 per slice
])
def test_subcyc_slices_match_oracle(kw):
    deck = configs.subcycling(**kw)
    print("sub-cycled slices:", kw, flush=True)
    sim = Simulation(deck)
    o = twin.make_twin(sim)
    o.set_subcycling(*subcycling_from_deck(deck))
    sim.step3d_begin()
    o.step_begin()
    norm = twin.ErrorNorm()
    subs = []
    for j in range(1, sim.grid.nxi + 1):
        twin.sync_slice_inputs(sim, o, j)
        it_g, it_o = sim.slice(j), o.slice(j)
        assert it_g == it_o, (j, it_g, it_o)
        n_g, d_g, f_g = sim.last_subcycles()
        n_o, d_o, f_o = o.last_subcycles()
        assert (n_g, d_g) == (n_o, d_o) and abs(f_g - f_o) <= 1e-12 * f_o, (j, (n_g, d_g, f_g), (n_o, d_o, f_o))
        subs.append(n_g)
        twin.compare_fields(sim, o, norm, j)
        twin.compare_species(sim, o, norm, j)
        for nt in sim.neut:
            norm.add("levels", nt.levels(), o.neutral_levels(nt.id), j)
            norm.add("rho_ion", nt.ion_density(j), o.neutral_rho_hist(nt.id)[j], j)
    print("sub-cycles per slice", subs, {n: f"{e:.1e}" for n, e in sorted(norm.errors().items())})
    assert max(subs) >= 2, "the deck must sub-cycle"
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"
    sim.close()
    o.close()


def test_subcyc_step_from_deck():
    """one 3D step from the deck alone on each side: the histories, the
    plasma, the beam after its push, the iteration total; the clamp holds
    every electron's gamma / (gamma - p_z) at or below its value"""
    deck = configs.subcycling(nxi=48)
    sim = Simulation(deck)
    o = twin.oracle_from_deck(deck)
    nxi = sim.grid.nxi
    sim.step3d_begin()
    o.step_begin()
    tot_g = tot_o = 0
    for j in range(1, nxi + 1):
        tot_g += sim.slice(j)
        tot_o += o.slice(j)
    assert tot_g == tot_o
    assert sim.exp_fac_max() <= 1.5 * (1 + 1e-12)
    for name in ("psi", "e", "b", "b_spe", "cu", "q_spe"):
        g = np.stack([sim.fields[name].hist(j) for j in range(1, nxi + 1)])
        r = o.hist(twin.ORA_ID[name])[1:nxi + 1]
        assert rel_err(g, r) <= TOL, (name, rel_err(g, r))
    g = twin.particles_by_id(sim.spe[0].get())
    r = twin.particles_by_id({k: np.array(v) for k, v in o.species(0).items()})
    assert np.array_equal(g["id"], r["id"])
    for k in ("x1", "x2", "p1", "p2", "p3", "gamma"):
        assert rel_err(g[k], r[k]) <= TOL, k
    sim.step3d_end()
    o.step_end(1)
    gb, ob = sim.beam[0].get(), o.beam(0)
    assert len(gb["x1"]) == len(ob["x1"])
    for k in ("x1", "x2", "x3", "p1", "p2", "p3"):      # the two sides store the beam in different orders
        assert rel_err(np.sort(gb[k]), np.sort(ob[k])) <= 1e-9, k
    sim.close()
    o.close()


def test_subcyc_entry_points():
    """qpad_exp_fac_max and qpad_species_push_clamped against the oracle on
    the same particles and fields; the settings sub-cycling refuses"""
    deck = configs.subcycling()
    sim = Simulation(deck)
    o = twin.make_twin(sim)
    o.set_subcycling(*subcycling_from_deck(deck))
    sim.step3d_begin()
    o.step_begin()
    for j in range(1, 30):
        twin.sync_slice_inputs(sim, o, j)
        sim.slice(j)
        o.slice(j)
    twin.sync_slice_inputs(sim, o, 30)
    assert abs(sim.exp_fac_max() - o.exp_fac_max()) <= 1e-13 * o.exp_fac_max()
    assert sim.exp_fac_max() > 1.1
    e = sim.fields["e"].get() * 3.0            # a stronger field than the slice's
    b = sim.fields["b"].get()
    sim.fields["e"].set(e)
    o.field(po.E)[...] = e
    o.field(po.B)[...] = b
    # a clamp halfway between 1 and the largest factor: the electrons near
    # the top are above it after push_u
    dsub, F = 0.0125, 1.0 + 0.5 * (sim.exp_fac_max() - 1.0)
    sp = o.species_raw(0)
    sim.spe[0].push(dsub, exp_fac_clamped=F)
    o.L.ora_push_u_robust(o.h, sp, po.ptr(o.field(po.E)), po.ptr(o.field(po.B)), dsub)
    o.L.ora_clamp_exp_fac(sp, F)
    o.L.ora_push_x(sp, dsub)
    o.L.ora_update_bound(o.h, sp)
    g = twin.particles_by_id(sim.spe[0].get())
    r = twin.particles_by_id({k: np.array(v) for k, v in o.species(0).items()})
    assert np.array_equal(g["id"], r["id"])
    for k in ("x1", "x2", "p1", "p2", "p3", "gamma"):
        assert rel_err(g[k], r[k]) <= 1e-12, k
    ef = g["gamma"] / (g["gamma"] - g["p3"])
    assert abs(ef.max() - F) <= 1e-12 * F and np.sum(ef > F * (1 - 1e-12)) > 1, (ef.max(), F)
    for args, match in [((1, 1.2, 1.0, 0.01), "clamped"), ((1, 0.0, 1.5, 0.01), "expansion_fac_max"),
                        ((1, 1.2, 1.5, 0.06), "dt_2d_min"), ((2, 1.2, 1.5, 0.01), "on")]:
        with pytest.raises(QpadError, match=match):
            sim.set_subcycling(*args[1:], on=args[0])
    with pytest.raises(QpadError, match="clamped"):
        sim.spe[0].push(dsub, exp_fac_clamped=0.5)
    sim.close()
    o.close()
