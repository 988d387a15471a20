"""GPU tests of the learned B_perp seed network (BF16 matrix-core kernels)
against the numpy oracle (oracle/seednet.py), through the C ABI.

Tolerances: the device and the oracle round the same quantities to BF16 at
the same points; they differ only by the FP32 summation order (the device
sums on the matrix cores), which can move a value across a BF16 rounding
boundary.  Outputs agree to 5e-3 of their largest magnitude and gradients to
2e-2 in norm per layer; the Adam update, given the same gradient, to 1e-5.
"""
import numpy as np
import pytest

import seednet as SN
from qpad_amd import SeedNet, Simulation, configs

pytestmark = pytest.mark.gpu


def small_sim(max_mode=2, nr=200, nxi=32):
    return Simulation(configs.small(nr=nr, nxi=nxi, max_mode=max_mode, ppc=(2, 4), total_num=1 << 12,
                                    offset_x=0.03))


def shape_of(net):
    inf = net.info()
    shape = SN.NetShape(inf.cin, inf.cout, net.cfg.width, inf.layers)
    assert shape.nparams == inf.nparams
    assert list(shape.cin_p) == list(inf.cin_p)[:inf.layers]
    assert list(shape.cout_p) == list(inf.cout_p)[:inf.layers]
    return shape


def layer_slices(shape):
    return [slice(shape.w_off[l], shape.b_off[l] + shape.cout_p[l]) for l in range(shape.layers)]


@pytest.mark.parametrize("max_mode,npos", [(2, 200), (4, 77), (1, 16)])
def test_forward_and_gradient_match_oracle(max_mode, npos):
    sim = small_sim(max_mode=max_mode)
    net = SeedNet(sim, seed=3)
    shape = shape_of(net)
    rng = np.random.default_rng(max_mode)
    theta = shape.random_params(rng)
    net.set_params(theta)
    x = rng.normal(size=(npos, shape.cin)).astype(np.float32)
    t = rng.normal(size=(npos, shape.cout)).astype(np.float32) * 0.5
    y, loss, g = net.eval(x, t, grad=True)
    yo, cache = SN.forward(shape, theta, x)
    lo, go = SN.backward(shape, cache, t)
    sim.close()
    assert np.max(np.abs(y - yo)) <= 5e-3 * np.max(np.abs(yo)), np.max(np.abs(y - yo))
    assert abs(loss - lo) <= 1e-3 * lo
    for l, sl in enumerate(layer_slices(shape)):
        err = np.linalg.norm(g[sl] - go[sl]) / np.linalg.norm(go[sl])
        assert err <= 2e-2, (l, err)
    # padded parameters get exactly zero gradient
    pad = theta == 0.0
    pad &= go == 0.0
    assert np.all(g[pad] == 0.0)


def test_adam_step_matches_oracle():
    sim = small_sim()
    net = SeedNet(sim, lr=2e-3, seed=4)   # lr also in the oracle's steps below
    shape = shape_of(net)
    rng = np.random.default_rng(11)
    theta = shape.random_params(rng)
    net.set_params(theta)
    x = rng.normal(size=(150, shape.cin)).astype(np.float32)
    t = rng.normal(size=(150, shape.cout)).astype(np.float32)
    th = theta.copy()
    m = np.zeros_like(th)
    v = np.zeros_like(th)
    for step in (1, 2, 3):
        _, _, g = net.eval(x, t, grad=True, update=True)
        th, m, v = SN.adam_step(th, m, v, g, step, lr=2e-3)
        got = net.get_params()
        assert np.allclose(got, th, rtol=1e-5, atol=1e-7), (step, np.max(np.abs(got - th)))
        assert net.info().steps == step
    sim.close()


def test_training_reduces_loss_and_is_deterministic():
    def run():
        sim = small_sim(max_mode=1, nr=128)
        net = SeedNet(sim, lr=3e-3, seed=5)
        shape = shape_of(net)
        rng = np.random.default_rng(9)
        x = rng.normal(size=(128, shape.cin)).astype(np.float32)
        # a target the convolution can represent: a smoothed mix of inputs
        t = 0.3 * x[:, :shape.cout] + 0.2 * np.roll(x[:, shape.cout:2 * shape.cout], 1, axis=0)
        first = None
        for _ in range(300):
            _, loss, _ = net.eval(x, t, update=True)
            first = loss if first is None else first
        losses = net.losses()
        th = net.get_params()
        sim.close()
        return first, loss, losses, th

    first, last, losses, th = run()
    assert last < 0.2 * first, (first, last)
    assert len(losses) == 300 and np.isclose(losses[0], first, rtol=1e-6)
    _, _, losses2, th2 = run()
    assert np.array_equal(th, th2) and np.array_equal(losses, losses2)


def test_slice_seed_matches_oracle():
    """Fine-grained seed at slice j: scales, packed inputs, prediction and its
    addition to b_spe, against the oracle on the same fields."""
    sim = small_sim(max_mode=2, nr=96, nxi=40)
    net = SeedNet(sim, seed=6)
    shape = shape_of(net)
    rng = np.random.default_rng(12)
    theta = shape.random_params(rng)
    net.set_params(theta)
    g = sim.grid
    f = sim.fields
    net.mode(seed=False, train=False)
    sim.step3d_begin()
    j0 = 30
    b_old = None
    for j in range(1, j0):
        if j == j0 - 1:
            b_old = f["b_spe"].get()        # b_spe at the start of slice j0 - 1
        sim.slice(j)
    # the seed of slice j0 comes first in the slice, from the fields slice
    # j0 - 1 left: b_spe, the J predictor cu, q_spe of slice j0 - 1
    b_spe, cu, q = f["b_spe"].get(), f["cu"].get(), f["q_spe"].get()
    qb = f["q_beam"].hist(j0)
    # the library keeps b_spe of the previous slice start (for dB(xi-1));
    # the plain slices above did not pack, so pack slice j0 - 1 once with
    # b_spe at its start value (modes off: nothing is added)
    f["b_spe"].set(b_old)
    net.seed(j0 - 1)
    f["b_spe"].set(b_spe)
    net.mode(seed=True, train=False)
    net.seed(j0)
    got = f["b_spe"].get()
    sim.close()
    assert np.max(np.abs(b_spe)) > 0.0
    x, sB = SN.slice_inputs(b_spe, b_old, cu, q, qb, g.nr)
    y, _ = SN.forward(shape, theta, x)
    want = SN.apply_prediction(b_spe, b_old, y, sB, net.cfg.out_scale)
    tol = 5e-3 * np.max(np.abs(y)) * sB * net.cfg.out_scale + 1e-15 * np.max(np.abs(b_spe))
    assert np.max(np.abs(got - want)) <= tol, (np.max(np.abs(got - want)), tol)
    assert np.array_equal(got[2], b_spe[2])


def test_seeded_iteration_converges_like_plain():
    """With the seed on (freshly initialised network: the first-order
    extrapolation), every slice still converges within iter_max, in fewer
    iterations than the plain seed, and the converged fields stay within the
    iteration tolerance's reach of the plain run's."""
    deck = configs.small(nr=64, nxi=48, max_mode=1, ppc=(2, 4), total_num=1 << 12)
    out = {}
    for seeded in (False, True):
        sim = Simulation(deck)
        net = SeedNet(sim, seed=7)
        net.mode(seed=seeded, train=False)
        sim.step3d_begin()
        its = [sim.slice(j) for j in range(1, sim.grid.nxi + 1)]
        out[seeded] = (its, sim.fields["b_spe"].get(), sim.fields["psi"].get())
        sim.close()
    its_p, bp, pp = out[False]
    its_s, bs, ps = out[True]
    iter_max = deck["simulation"]["iter_max"]
    assert max(its_s) <= iter_max
    assert sum(its_s) < sum(its_p)
    scale = np.max(np.abs(bp[0:2]))
    db = np.max(np.abs(bs[0:2] - bp[0:2])) / scale
    dpsi = np.max(np.abs(ps - pp)) / np.max(np.abs(pp))
    print(f"iterations plain {sum(its_p)} seeded {sum(its_s)}; B_perp diff {db:.2e}, psi diff {dpsi:.2e}")
    assert db <= 5e-2 and dpsi <= 5e-2


def test_slice_training_step_matches_oracle():
    """Seed and one Adam step at slice j (fused forward keeping the
    activations, then the layer-by-layer backward): the logged loss and the
    direction of the update against the oracle on the same fields."""
    sim = small_sim(max_mode=2, nr=96, nxi=40)
    lr = 1e-4
    net = SeedNet(sim, lr=lr, seed=8)
    shape = shape_of(net)
    rng = np.random.default_rng(13)
    theta = shape.random_params(rng, scale=0.5)
    net.set_params(theta)
    g = sim.grid
    f = sim.fields
    net.mode(seed=False, train=False)
    sim.step3d_begin()
    j0 = 28
    for j in range(1, j0):
        if j == j0 - 1:
            b_old = f["b_spe"].get()
        sim.slice(j)
    b1, cu, q = f["b_spe"].get(), f["cu"].get(), f["q_spe"].get()
    qb = f["q_beam"].hist(j0)
    f["b_spe"].set(b_old)
    net.seed(j0 - 1)
    f["b_spe"].set(b1)
    net.mode(seed=True, train=True)
    net.seed(j0)
    x, sB = SN.slice_inputs(b1, b_old, cu, q, qb, g.nr)
    osc = net.cfg.out_scale
    # a "converged" field: the extrapolation plus a smooth correction
    b_conv = b1.copy()
    r = np.arange(g.nr + 2)
    for c in range(2):
        b_conv[c, :, :] = 2 * b1[c] - b_old[c] + 0.3 * osc * sB * np.sin(0.1 * r + c)[None, :]
    f["b_spe"].set(b_conv)
    net.learn(j0)
    loss = net.losses()[-1]
    th1 = net.get_params()
    sim.close()
    y, cache = SN.forward(shape, theta, x)
    t = SN.slice_target(b_conv, b1, b_old, sB, osc, g.nr)
    lo, go = SN.backward(shape, cache, t)
    assert abs(loss - lo) <= 2e-3 * lo, (loss, lo)
    step = th1 - theta
    big = np.abs(go) > 1e-2 * np.max(np.abs(go))
    agree = np.mean(np.sign(step[big]) == -np.sign(go[big]))
    assert agree >= 0.98, agree
    assert np.allclose(np.abs(step[big]), lr, rtol=0.05)
