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
