"""Non-uniform plasmas on the GPU against the oracle: the profiled injection
(inject_fdist2d, fdist2d_class.f03:289-357, with the densities of
fdist2d_lib.f03 evaluated on the host at the lattice), the den_min cut that
compacts the lattice, and the renew at each 3D step's propagation distance
(simulation_class.f03:497-501)."""
import numpy as np
import pytest

import twin
from conftest import rel_err
from qpad_amd import Simulation, configs

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _same_particles(sim, o, what):
    g = twin.particles_by_id(sim.spe[0].get())
    r = twin.particles_by_id({k: np.array(v) for k, v in o.species(0).items()})
    assert np.array_equal(g["id"], r["id"]), (what, len(g["id"]), len(r["id"]))
    for k in ("x1", "x2", "p1", "p2", "p3", "q", "gamma", "psi"):
        assert rel_err(g[k], r[k]) <= 1e-14, (what, k, rel_err(g[k], r[k]))
    return len(g["id"])


@pytest.mark.parametrize("kw", [
    dict(),                                              # parabolic channel cut at r = 3, rising density
    dict(den_min=0.45, uth=(0.05, 0.05, 0.02)),          # the cut drops the first step's inner lattice
    dict(profile={"profile": ["hollow-channel", "sine-series"], "channel_rmin": 0.5, "channel_rmax": 2.5,
                  "channel_depth": 1.0, "start": -1.0, "length": 40.0, "sine_coefs": [1.0, 0.3]}, max_mode=2),
])
def test_profiled_steps_match_oracle(kw):
    """two 3D steps from the deck alone on each side: the injected plasma,
    every history of each step, and the plasma renewed at s = dt"""
    deck = configs.profiled(**kw)
    print("profiled steps:", kw, flush=True)
    sim = Simulation(deck)
    o = twin.oracle_from_deck(deck)
    nxi = sim.grid.nxi
    counts = [_same_particles(sim, o, "injected")]
    for step in (1, 2):
        sim.step3d_begin()
        o.step_begin()
        its_g = [sim.slice(j) for j in range(1, nxi + 1)]
        its_o = [o.slice(j) for j in range(1, nxi + 1)]
        assert its_g == its_o, step
        for name in ("psi", "e", "b", "b_spe", "cu", "q_spe"):
            g = np.stack([sim.fields[name].hist(j) for j in range(1, nxi + 1)])
            r = o.hist(twin.ORA_ID[name])[1:nxi + 1]
            assert rel_err(g, r) <= TOL, (step, name, rel_err(g, r))
        sim.step3d_end()
        twin.oracle_profiles_at(o, deck, step * sim.grid.dt)
        o.step_end(step)
        counts.append(_same_particles(sim, o, f"renewed after step {step}"))
    print("particles per injection", counts)
    sim.close()
    o.close()


def test_analytic_profile_deck():
    """QPAD's analytic density z^2/(z + 10)^2 exp(-(x^2 + y^2)/50): no plasma
    at z = 0 (den_min drops the whole lattice), a thin one at z = dt; the
    slices match the oracle on both steps"""
    deck = configs.analytic_profile()
    sim = Simulation(deck)
    o = twin.oracle_from_deck(deck)
    assert sim.spe[0].count() == 0 and len(o.species(0)["x1"]) == 0
    nxi = sim.grid.nxi
    for step in (1, 2):
        sim.step3d_begin()
        o.step_begin()
        assert [sim.slice(j) for j in range(1, nxi + 1)] == [o.slice(j) for j in range(1, nxi + 1)]
        g = np.stack([sim.fields["psi"].hist(j) for j in range(1, nxi + 1)])
        assert rel_err(g, o.hist(twin.ORA_ID["psi"])[1:nxi + 1]) <= TOL
        sim.step3d_end()
        twin.oracle_profiles_at(o, deck, step * sim.grid.dt)
        o.step_end(step)
        n = _same_particles(sim, o, f"renewed after step {step}")
        assert n == 384
    q = sim.spe[0].get()["q"]
    x1, x2 = sim.spe[0].get()["x1"], sim.spe[0].get()["x2"]
    z = 2 * sim.grid.dt
    rn = np.sqrt(x1 ** 2 + x2 ** 2) / sim.grid.dr
    want = -rn * z ** 2 / (z + 10.0) ** 2 * np.exp(-(x1 ** 2 + x2 ** 2) / 50.0) / 8.0
    assert np.allclose(q, want, rtol=1e-12, atol=0)
    sim.close()
    o.close()


def _shaped_beam_deck():
    d = configs.tri_gaussian()
    d["simulation"]["grid"] = [128, 256]
    d["beam"][0].update({"profile": ["super-gaussian", "parabolic", "rational"],
                         "super_gauss_center": [0.0, 0.0, 0.0], "super_gauss_sigma": [0.3, 1.0, 1.0],
                         "super_gauss_order": [2.0, 1.0, 1.0], "parabolic_center": [0.0, 0.05, 0.0],
                         "parabolic_radius": [1.0, 0.6, 1.0], "poly_x0": [0.0, 0.0, -2.5],
                         "poly_numerator3": [1.0], "poly_denominator3": [1.0, 0.0, 4.0]})
    return d


@pytest.mark.parametrize("which", ["shaped", "analytic"])
def test_tabulated_beam_matches_oracle(which):
    """lattice beams whose profiles the host evaluates (super-gaussian x
    parabolic x rational; an analytic math_func): the library's lattice box
    is the host's, the injected beam is the oracle's, and one 3D step
    matches slice by slice"""
    from qpad_amd.input import beam_from_deck, grid_from_deck
    from qpad_amd.profiles import beam_lattice_box
    import ctypes as C
    if which == "shaped":
        deck = _shaped_beam_deck()
    else:
        deck = _shaped_beam_deck()
        deck["beam"][0].update({"profile": ["analytic"] * 3,
                                "math_func": "exp(-z^2/0.5)*exp(-(x^2+y^2)/0.18)*(1+0.5*step(x))"})
    print("tabulated beam:", which, flush=True)
    sim = Simulation(deck)
    o = twin.oracle_from_deck(deck)
    g, _, z0 = grid_from_deck(deck)
    tb = beam_from_deck(deck["beam"][0], z0, seed=10 + 7)
    k0, nk, ni = C.c_int32(0), C.c_int32(0), C.c_int32(0)
    assert sim.lib.qpad_beam_lattice_box(sim.ctx, C.byref(tb.cfg), C.byref(k0), C.byref(nk), C.byref(ni)) == 0
    assert (k0.value, nk.value, ni.value) == beam_lattice_box(g.nr, g.nxi, g.dr, g.dxi, tb.cfg)
    gb, ob = sim.beam[0].get(), o.beam(0)
    assert len(gb["x1"]) == len(ob["x1"]) > 1000
    for k in ("x1", "x2", "x3", "p1", "p2", "p3", "q"):
        assert rel_err(np.sort(gb[k]), np.sort(ob[k])) <= 1e-14, k
    nxi = sim.grid.nxi
    sim.step3d_begin()
    o.step_begin()
    assert [sim.slice(j) for j in range(1, nxi + 1)] == [o.slice(j) for j in range(1, nxi + 1)]
    for name in ("psi", "e", "b", "q_spe"):
        gh = np.stack([sim.fields[name].hist(j) for j in range(1, nxi + 1)])
        assert rel_err(gh, o.hist(twin.ORA_ID[name])[1:nxi + 1]) <= TOL, name
    sim.close()
    o.close()


def test_analytic_example_deck_runs():
    """QPAD's analytic-profile example deck at its size (Nr 1024, Nxi 1024,
    an analytic lattice beam of about 4 million points with quiet start, an
    analytic plasma): one 3D step; the beam keeps its charge, the plasma of
    the second step follows the expression"""
    deck = configs.analytic_example()
    sim = Simulation(deck)
    nb = sim.beam[0].count()
    assert nb > 4_000_000 and sim.spe[0].count() == 0
    its = sim.step3d()
    n = sim.spe[0].count()
    b = sim.fields["b"].hist(146)              # the beam's centre (absolute z = 0)
    print("beam particles", nb, "iterations", its, "plasma particles after the renew", n)
    assert n > 0 and np.all(np.isfinite(b)) and np.max(np.abs(b)) > 0
    q = sim.spe[0].get()
    z = sim.grid.dt
    r2 = q["x1"] ** 2 + q["x2"] ** 2
    want = -np.sqrt(r2) / sim.grid.dr * z ** 2 / (z + 10.0) ** 2 * np.exp(-r2 / 50.0) / 32.0
    assert np.allclose(q["q"], want, rtol=1e-12, atol=0)
    sim.close()


def test_profile_on_off_on():
    """a species returned to the uniform plasma and profiled again: every
    injection has the lattice the current setting gives"""
    from qpad_amd.profiles import PlasmaProfile
    deck = configs.profiled()
    sim = Simulation(deck)
    sp = sim.spe[0]
    prof = sp.profile
    n_prof = sp.count()
    sp.set_profile(None, 0.0)
    sp.renew()
    assert sp.count() == sim.grid.nr * 2 * 4 and n_prof < sp.count()
    hollow = PlasmaProfile({"profile": ["hollow-channel", "uniform"], "channel_rmin": 1.0, "channel_rmax": 2.0,
                            "channel_depth": 1.0})
    sp.set_profile(hollow, 0.0)
    sp.renew()
    r = np.hypot(sp.get()["x1"], sp.get()["x2"])
    assert sp.count() > 0 and np.all((r > 1.0) & (r < 2.0))
    sp.set_profile(prof, 0.0)
    sp.renew()
    assert sp.count() == n_prof
    sim.close()
