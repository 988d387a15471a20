"""The learned seed's protocol on C3-style physics: train on one deck,
evaluate frozen on another.

The network is trained online on the GPU (its own forward, backward and
Adam kernels) over several 3D steps of a reduced C3-style deck (modes 0..4,
electrons and mobile ions, neither neutralised, a beam offset from the
axis), then frozen and applied to a deck it was not trained on (other beam
charge and offset).  On that deck three legs run one 3D step each at the
same tolerance: QPAD's seed (the previous slice's B_perp), the untrained
network (its zero output layer leaves the first-order extrapolation in xi)
and the trained network.  Each leg's converged fields are compared with the
FP64 oracle's plain run of the same deck (simulation_class.f03:342-469), so
the cost of seeding is stated as a field error next to its gain in
iterations."""
import numpy as np
import pytest

import twin
from qpad_amd import SeedNet, Simulation, configs

pytestmark = pytest.mark.gpu

FIELDS = ("psi", "e", "b_spe")


def c3_style(n_b, offset, nr=256, nxi=128):
    """C3 physics (modes 0..4, ions of mass 1836, neither species
    neutralised, beam offset) on a reduced grid: r_max 6, dxi 0.04, ppc 4x16,
    sigma_r 0.5, sigma_xi 0.3 (so that the box holds the beam and two plasma
    periods behind it)."""
    return configs.small(nr=nr, nxi=nxi, max_mode=4, ppc=(4, 16), total_num=1 << 16, rmax=6.0, dxi=0.04,
                         n_peak=n_b, sigma_r=0.5, sigma_xi=0.3, offset_x=offset, ions=True)


def run_leg(deck, params=None, seeded=True):
    """one 3D step; params: trained weights (None: a fresh network)"""
    sim = Simulation(deck)
    net = SeedNet(sim, seed=3)
    if params is not None:
        net.set_params(params)
    net.mode(seed=seeded, train=False)
    sim.step3d_begin()
    its = [sim.slice(j) for j in range(1, sim.grid.nxi + 1)]
    hist = {n: np.stack([sim.fields[n].hist(j) for j in range(1, sim.grid.nxi + 1)]) for n in FIELDS}
    sim.close()
    return np.array(its), hist


def field_error(hist, ref):
    """max over slices of |leg - oracle| / max |oracle| per field"""
    return {n: float(np.max(np.abs(hist[n] - ref[n])) / np.max(np.abs(ref[n]))) for n in FIELDS}


def test_trained_seed_on_an_unseen_deck():
    # training pairs come from the plain solver (seed off), so that the
    # converged labels do not carry the network's own seed; lr 1e-5: one Adam
    # step per slice on the xi-ordered stream of slices
    train_deck = c3_style(n_b=8.0, offset=0.05)
    sim = Simulation(train_deck)
    net = SeedNet(sim, seed=3, lr=1e-5)
    net.mode(seed=False, train=True)
    nsteps = 8
    for _ in range(nsteps):
        sim.step3d()
    losses = net.losses()
    params = net.get_params()
    assert int(net.info().steps) == nsteps * sim.grid.nxi
    sim.close()
    nx = train_deck["simulation"]["grid"][1]
    first, last = float(np.median(losses[:nx])), float(np.median(losses[-nx:]))
    assert np.all(np.isfinite(losses)) and last <= first, ("training loss", first, last)

    eval_deck = c3_style(n_b=5.0, offset=0.02)
    o = twin.oracle_from_deck(eval_deck)
    o.step_begin()
    oits = np.array([o.slice(j) for j in range(1, o.grid.nxi + 1)])
    ref = {"psi": o.hist(twin.po.PSI)[1:o.grid.nxi + 1].copy(), "e": o.hist(twin.po.E)[1:o.grid.nxi + 1].copy(),
           "b_spe": o.hist(twin.po.B_SPE)[1:o.grid.nxi + 1].copy()}
    o.close()
    legs = {"plain": run_leg(eval_deck, seeded=False), "untrained": run_leg(eval_deck),
            "trained": run_leg(eval_deck, params)}
    iter_max = eval_deck["simulation"]["iter_max"]
    report = {}
    for name, (its, hist) in legs.items():
        report[name] = {"iters_per_slice": float(its.mean()), "slices_at_iter_max": int(np.sum(its == iter_max)),
                        "field_error_vs_oracle": field_error(hist, ref)}
    print("training loss median first/last step", first, last)
    print("oracle iterations per slice", float(oits.mean()))
    for name, r in report.items():
        print(name, r)
    # the plain leg is the oracle's algorithm on the GPU
    assert np.array_equal(legs["plain"][0], oits)
    assert max(report["plain"]["field_error_vs_oracle"].values()) <= 1e-9
    # seeding saves iterations; the trained network's correction keeps them
    # (within 2 %) on a deck it was not trained on
    assert report["untrained"]["iters_per_slice"] < report["plain"]["iters_per_slice"]
    assert report["trained"]["iters_per_slice"] <= 1.02 * report["untrained"]["iters_per_slice"]
    # and the converged fields stay at the iteration tolerance's distance from
    # the oracle's (tolerance 1e-3 on B_phi)
    for name in ("untrained", "trained"):
        assert max(report[name]["field_error_vs_oracle"].values()) <= 2e-2, (name, report[name])
