"""Pins of the seed-network oracle (oracle/seednet.py) on the CPU.

Each check is independent of the restatement it pins:
* BF16 rounding against torch's float32 -> bfloat16 conversion;
* the convolution against a brute-force loop over (node, tap, channel);
* the backward pass against central finite differences of the exact
  (unrounded, float64) network -- a dropped term, a wrong tap direction or a
  transposed weight fails it;
* the first Adam step against its closed form theta - lr g / (|g| + eps);
* the slice packing against a direct per-channel construction.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import seednet as SN  # noqa: E402


def test_bf16_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e-20, 100), rng.normal(0, 1e20, 100),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0 + 2 ** -7, 3.0 + 3 * 2 ** -8]])
    x = x.astype(np.float32)
    # exact ties of the 8-bit mantissa, both parities
    u = (rng.integers(0, 2 ** 15, 200).astype(np.uint32) << 16) | 0x8000
    x = np.concatenate([x, u.view(np.float32)])
    x = x[np.isfinite(x)]
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(SN.bf16(x), ref)


def brute_conv(h, W, b):
    npos, cout = h.shape[0], W.shape[1]
    z = np.zeros((npos, cout))
    for p in range(npos):
        for co in range(cout):
            s = b[co]
            for tap in range(3):
                q = p + tap - 1
                if 0 <= q < npos:
                    for ci in range(h.shape[1]):
                        s += W[tap, co, ci] * h[q, ci]
            z[p, co] = s
    return z


def test_conv_matches_brute_force():
    rng = np.random.default_rng(1)
    h = rng.normal(size=(7, 5))
    W = rng.normal(size=(3, 4, 5))
    b = rng.normal(size=4)
    assert np.allclose(SN.conv(h, W, b), brute_conv(h, W, b), rtol=1e-13, atol=1e-13)
    # transpose and weight gradient are the adjoints of conv:
    # <conv(h) - b, dz> = <h, conv_transpose(dz)> = <W, dW>
    dz = rng.normal(size=(7, 4))
    lhs = np.sum((SN.conv(h, W, b) - b) * dz)
    assert np.isclose(lhs, np.sum(h * SN.conv_transpose(dz, W)), rtol=1e-12)
    dW, db = SN.conv_weight_grad(dz, h)
    assert np.isclose(lhs, np.sum(W * dW), rtol=1e-12)
    assert np.allclose(db, dz.sum(axis=0))


@pytest.mark.parametrize("layers", [2, 4])
def test_gradient_matches_finite_differences(layers):
    rng = np.random.default_rng(2 + layers)
    shape = SN.NetShape(cin=5, cout=3, width=32, layers=layers)
    theta = shape.random_params(rng).astype(np.float64)
    x = rng.normal(size=(9, 5))
    t = rng.normal(size=(9, 3))

    def loss(th):
        y, cache = SN.forward(shape, th, x, exact=True)
        return SN.backward(shape, cache, t, exact=True)[0]

    y, cache = SN.forward(shape, theta, x, exact=True)
    _, g = SN.backward(shape, cache, t, exact=True)
    # every layer's weights (all three taps) and biases
    idx = []
    for l in range(layers):
        for tap in range(3):
            co, ci = rng.integers(0, shape.cout_u[l]), rng.integers(0, shape.cin_u[l])
            idx.append(shape.w_off[l] + (tap * shape.cout_p[l] + co) * shape.cin_p[l] + ci)
        idx.append(shape.b_off[l] + rng.integers(0, shape.cout_u[l]))
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (loss(tp) - loss(tm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # padding channels carry no gradient
    pad = np.ones(shape.nparams, bool)
    for l in range(layers):
        W = np.zeros((3, shape.cout_p[l], shape.cin_p[l]), bool)
        W[:, :shape.cout_u[l], :shape.cin_u[l]] = True
        pad[shape.w_off[l]:shape.w_off[l] + W.size] = ~W.ravel()
        pad[shape.b_off[l]:shape.b_off[l] + shape.cout_u[l]] = False
        pad[shape.b_off[l] + shape.cout_u[l]:shape.b_off[l] + shape.cout_p[l]] = True
    assert np.all(g[pad] == 0.0)


def test_first_adam_step_closed_form():
    rng = np.random.default_rng(5)
    th = rng.normal(size=100).astype(np.float32)
    g = rng.normal(size=100).astype(np.float32) * 1e-3
    z = np.zeros(100, np.float32)
    th1, m, v = SN.adam_step(th, z, z, g, 1, lr=1e-3, eps=1e-8)
    assert np.allclose(th1, th - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-6, atol=1e-9)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_slice_inputs_layout():
    rng = np.random.default_rng(6)
    nm, nr = 5, 12
    ld = nr + 2
    b = rng.normal(size=(3, nm, ld))
    bo = rng.normal(size=(3, nm, ld))
    cu = rng.normal(size=(3, nm, ld)) * 3
    q = rng.normal(size=(1, nm, ld)) * 2
    qb = np.zeros((1, nm, ld))
    x, sB = SN.slice_inputs(b, bo, cu, q, qb, nr)
    assert x.shape == (nr, 8 * nm + 1)
    assert sB == np.abs(b[0:2, :, 1:nr + 1]).max()
    # node i = 3, mode array k = 2
    i, k = 3, 2
    p = i - 1
    sJ = np.abs(cu[0:2, :, 1:nr + 1]).max()
    assert x[p, 0 * nm + k] == b[0, k, i] / sB
    assert x[p, 1 * nm + k] == b[1, k, i] / sB
    assert x[p, 2 * nm + k] == (b[0, k, i] - bo[0, k, i]) / sB
    assert x[p, 5 * nm + k] == cu[1, k, i] / sJ
    assert np.all(x[:, 7 * nm:8 * nm] == 0.0)       # zero q_beam: scale 1, inputs 0
    assert x[p, 8 * nm] == p / nr
    y = rng.normal(size=(nr, 2 * nm))
    out = SN.apply_prediction(b, bo, y, sB, 0.02)
    assert np.isclose(out[1, k, i], 2 * b[1, k, i] - bo[1, k, i] + y[p, nm + k] * 0.02 * sB)
    assert np.array_equal(out[2], b[2]) and np.array_equal(out[:, :, 0], b[:, :, 0])
    t = SN.slice_target(out, b, bo, sB, 0.02, nr)
    assert np.allclose(t, np.clip(y, -1, 1))
