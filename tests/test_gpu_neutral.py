"""Field ionisation on the GPU against the oracle (row f4): a neutral gas
(neutral_class.f03) whose level fractions advance with the ADK rates of each
slice's final E (ionize_neutral, :600-753), whose released electrons join
the plasma at rest on the release lattice (add_particles, :755-837) and are
deposited and pushed like a species, and whose ion charge builds up slice by
slice (ion_deposit_neutral, :904-930).  Simulation order:
simulation_class.f03:351-353, 386-387, 406-407, 443-447, 503-509."""
import numpy as np
import pytest

import pyoracle as po
import twin
from conftest import rel_err
from qpad_amd import Simulation, configs

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _compare_neutral(sim, o, norm, j):
    for nt in sim.neut:
        k = nt.id
        norm.add(f"levels{k}", nt.levels(), o.neutral_levels(k), j)
        norm.add(f"rho_ion{k}", nt.ion_density(j), o.neutral_rho_hist(k)[j], j)


@pytest.mark.parametrize("element,ion_max,plasma,max_mode,n0,n_peak,push", [
    (3, 3, False, 1, 1e17, 6.0, "robust"),
    (7, 5, True, 2, 1e18, 8.0, "robust"),
    (18, 8, True, 1, 1e18, 8.0, "standard"),
])
def test_neutral_slices_match_oracle(element, ion_max, plasma, max_mode, n0, n_peak, push):
    deck = configs.ionization(element=element, ion_max=ion_max, plasma=plasma, max_mode=max_mode, n0=n0,
                              n_peak=n_peak, push_type=push, offset_x=0.03)
    print("neutral slices: element", element, "ion_max", ion_max, "plasma", plasma, "M", max_mode, flush=True)
    sim = Simulation(deck)
    o = twin.make_twin(sim)
    sim.step3d_begin()
    o.step_begin()
    norm = twin.ErrorNorm()
    for j in range(1, sim.grid.nxi + 1):
        twin.sync_slice_inputs(sim, o, j)
        it_g, it_o = sim.slice(j), o.slice(j)
        assert it_g == it_o, (j, it_g, it_o)
        twin.compare_fields(sim, o, norm, j)
        twin.compare_species(sim, o, norm, j)
        _compare_neutral(sim, o, norm, j)
    released = sim.neut[0].count()
    lv = sim.neut[0].levels()
    print({n: f"{e:.1e}" for n, e in sorted(norm.errors().items())}, "released", released)
    assert released > 0 and lv[:, :, ion_max + 1].max() > 0, "the driver must ionise the gas"
    # bookkeeping: every unit of ion charge released ppc / ion_max electrons
    ppc = sim.neut[0].cfg.ppc1 * sim.neut[0].cfg.ppc2
    assert released == int(round(lv[:, :, ion_max + 1].sum() * ppc / ion_max))
    assert np.allclose(lv[:, :, :ion_max + 1].sum(axis=2), 1.0, atol=1e-12)
    bad = norm.bad(TOL)
    assert not bad, f"beyond {TOL}: {bad}"
    sim.close()
    o.close()


def test_neutral_step_from_deck_and_renew():
    """two 3D steps from the deck alone (each side injects its own beam and
    gas): histories, levels and the ion charge after each step's last slice;
    the renew returns the gas to neutral between steps"""
    deck = configs.ionization(element=18, ion_max=8, plasma=True, n0=1e18, n_peak=8.0, nr=40, nxi=32)
    sim = Simulation(deck)
    o = twin.oracle_from_deck(deck)
    nxi = sim.grid.nxi
    isp = len(sim.spe) - 1
    for step in (1, 2):
        sim.step3d_begin()
        o.step_begin()
        for j in range(1, nxi + 1):
            sim.slice(j)
            o.slice(j)
        assert sim.neut[0].count() == len(o.species(isp)["x1"])
        e = rel_err(sim.neut[0].levels(), o.neutral_levels(isp))
        assert e <= TOL, ("levels", step, e)
        for name in ("psi", "e", "b", "b_spe", "cu", "q_spe"):
            g = np.stack([sim.fields[name].hist(j) for j in range(1, nxi + 1)])
            r = o.hist(twin.ORA_ID[name])[1:nxi + 1]
            assert rel_err(g, r) <= TOL, (name, step, rel_err(g, r))
        g = np.stack([sim.neut[0].ion_density(j) for j in range(1, nxi + 1)])
        assert rel_err(g, o.neutral_rho_hist(isp)[1:nxi + 1]) <= TOL, "rho_ion"
        sim.step3d_end()
        o.step_end(step)
        assert sim.neut[0].count() == 0 and np.all(sim.neut[0].levels()[:, :, 8] == 1.0)
    sim.close()
    o.close()


def test_neutral_entry_points_match_oracle():
    """qpad_neutral_update and qpad_neutral_ion_deposit alone on a given E:
    the levels, the released electrons (by id) and the ion charge"""
    deck = configs.ionization(element=7, ion_max=5, n0=1e18, max_mode=2)
    sim = Simulation(deck)
    o = twin.make_twin(sim)
    isp = sim.neut[0].id
    rng = np.random.default_rng(5)
    e = np.zeros((3, sim.nm, sim.grid.nr + 2))
    r = (np.arange(sim.grid.nr + 2) - 1.0) * sim.grid.dr
    e[2, 0] = 0.9 * np.exp(-r ** 2)          # O(100 GV/m) on the axis at n0 = 1e18
    e[0, 1] = 0.6 * r * np.exp(-r ** 2)
    e[1, 2] = 0.2 * rng.standard_normal(sim.grid.nr + 2)
    sim.fields["e"].set(e)
    o.field(po.E)[...] = e
    for rep in range(2):
        sim.neut[0].update()
        sp = o.species_raw(isp)
        L = sp.contents.multi_max
        # neutral%update: the discrete ion charge before, ionize, add
        ion_old = np.ctypeslib.as_array(sp.contents.ion_old, shape=(sp.contents.cfg.num_theta * sim.grid.nr,))
        ion_old[...] = o.neutral_levels(isp)[:, :, L + 1].reshape(-1)
        o.L.ora_neutral_ionize(o.h, sp, po.ptr(o.field(po.E)))
        o.L.ora_neutral_add(o.h, sp)
        assert rel_err(sim.neut[0].levels(), o.neutral_levels(isp)) <= 1e-13
        g = twin.particles_by_id(sim.neut[0].get())
        ref = twin.particles_by_id({n: np.array(v) for n, v in o.species(isp).items()})
        assert np.array_equal(g["id"], ref["id"]) and len(g["id"]) > 0
        for k in ("x1", "x2", "q"):
            assert rel_err(g[k], ref[k]) <= 1e-13, k
        # the ions of this release into rho_ion and q_spe
        sim.fields["q_spe"].set(np.zeros((1, sim.nm, sim.grid.nr + 2)))
        sim.neut[0].ion_deposit(sim.fields["q_spe"])
        o.L.ora_neutral_ion_deposit(o.h, sp)
        _, qn = o.species_q_own(isp)
        assert rel_err(sim.neut[0].qn(), qn) <= 1e-13
        assert rel_err(sim.fields["q_spe"].get(), qn) <= 1e-13
        e *= 1.3
        sim.fields["e"].set(e)
        o.field(po.E)[...] = e
    sim.close()
    o.close()


def test_ionization_example_deck_runs():
    """QPAD's own ionisation deck (input_file/ionization/qpinput.json) at its
    size: one 3D step; the gas ionises behind the driver, the level
    fractions stay a partition of one, every unit of ion charge released
    ppc / ion_max electrons, and a rerun is identical"""
    out = []
    for rep in range(2):
        sim = Simulation(configs.ionization_example())
        assert (sim.grid.nr, sim.grid.nxi) == (500, 1000) and len(sim.spe) == 1
        its = sim.step3d_begin() or 0
        for j in range(1, sim.grid.nxi + 1):
            its += sim.slice(j)
        nt = sim.neut[0]
        lv = nt.levels()
        n = nt.count()
        out.append((its, n, lv, nt.ion_density(sim.grid.nxi), sim.fields["psi"].hist(600)))
        sim.close()
    its, n, lv, rho, psi = out[0]
    print("iterations", its, "released", n, "max level-1 fraction", lv[:, :, 0].max())
    assert n > 1000 and lv[:, :, 0].max() > 0.5
    assert np.all(lv >= 0) and np.allclose(lv[:, :, :4].sum(axis=2), 1.0, atol=1e-12)
    assert n == int(round(lv[:, :, 4].sum() * 64 / 3))
    assert np.max(np.abs(rho)) > 0
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(np.asarray(a), np.asarray(b)), "rerun differs"
