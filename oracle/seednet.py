"""numpy restatement of the learned B_perp seed network (forward, backward,
Adam, input packing) -- the checker of libqpad_hip's seed-network kernels.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py as the checker.
The product (qpad_amd) never imports it.

The network has no counterpart in QPAD.  It only replaces the seed of the
B_perp iteration, which in QPAD is b_spe of slice xi-1 (simulation_class.f03:
326, 370-375); the iteration then runs to QPAD's tolerance.  What is
restated here is the library's own definition of the network:

* conv1d layers over the radial nodes 1..nr, kernel 3, zero padding:
      z[p][co] = b[co] + sum_{tap=0..2} sum_ci W[tap][co][ci] h[p + tap - 1][ci]
  SiLU between layers, the last layer linear;
* BF16 (round to nearest even) weights, inputs and hidden activations; the
  products accumulate in FP32 (here: float64 sums rounded to float32);
* loss = mean over (node, used output channel) of (y - t)^2, its gradient
  rounded to BF16 where it feeds a matrix product, as on the device;
* Adam with bias correction on FP32 master weights;
* the seed it gives: 2 B(xi-1) - B(xi-2) + y * out_scale * scale
  (first-order extrapolation in xi plus the correction), trained on the
  residual of the converged B_perp in the same units, clipped to [-1, 1].

Pins (tests/test_seednet_oracle.py): the BF16 rounding against torch's
bfloat16 conversion; the convolution against a brute-force loop; the
gradient against central finite differences of the exact (unrounded,
float64) network; the first Adam step against its closed form.
``exact=True`` switches every rounding off for those checks.
"""
import numpy as np


def bf16(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


class NetShape:
    """Layer dimensions and the flat parameter layout of the library:
    layer after layer, W[3][cout_p][cin_p] then bias[cout_p]."""

    def __init__(self, cin, cout, width=64, layers=4):
        self.cin, self.cout, self.width, self.layers = cin, cout, width, layers
        r32 = lambda v: (v + 31) // 32 * 32  # noqa: E731
        self.cin_p = [r32(cin) if l == 0 else width for l in range(layers)]
        self.cout_p = [r32(cout) if l == layers - 1 else width for l in range(layers)]
        self.cin_u = [cin if l == 0 else width for l in range(layers)]
        self.cout_u = [cout if l == layers - 1 else width for l in range(layers)]
        self.w_off, self.b_off = [], []
        off = 0
        for l in range(layers):
            self.w_off.append(off)
            off += 3 * self.cout_p[l] * self.cin_p[l]
            self.b_off.append(off)
            off += self.cout_p[l]
        self.nparams = off

    @classmethod
    def for_modes(cls, max_mode, width=64, layers=4):
        nm = 2 * max_mode + 1
        return cls(8 * nm + 1, 2 * nm, width, layers)

    def unpack(self, theta):
        Ws, bs = [], []
        for l in range(self.layers):
            n = 3 * self.cout_p[l] * self.cin_p[l]
            Ws.append(theta[self.w_off[l]:self.w_off[l] + n].reshape(3, self.cout_p[l], self.cin_p[l]))
            bs.append(theta[self.b_off[l]:self.b_off[l] + self.cout_p[l]])
        return Ws, bs

    def pack(self, Ws, bs):
        th = np.zeros(self.nparams, np.float64)
        for l in range(self.layers):
            n = 3 * self.cout_p[l] * self.cin_p[l]
            th[self.w_off[l]:self.w_off[l] + n] = np.asarray(Ws[l]).ravel()
            th[self.b_off[l]:self.b_off[l] + self.cout_p[l]] = bs[l]
        return th

    def random_params(self, rng, scale=1.0):
        """He-uniform over the used channels, zero padding (as the library's
        initialisation, with its own generator)."""
        th = np.zeros(self.nparams, np.float32)
        for l in range(self.layers):
            W = np.zeros((3, self.cout_p[l], self.cin_p[l]), np.float32)
            a = np.sqrt(6.0 / (3 * self.cin_u[l])) * scale
            W[:, :self.cout_u[l], :self.cin_u[l]] = rng.uniform(-a, a, (3, self.cout_u[l], self.cin_u[l]))
            b = np.zeros(self.cout_p[l], np.float32)
            b[:self.cout_u[l]] = rng.uniform(-0.1, 0.1, self.cout_u[l]) * scale
            th[self.w_off[l]:self.w_off[l] + W.size] = W.ravel()
            th[self.b_off[l]:self.b_off[l] + b.size] = b
        return th


def shifted(h, tap):
    """rows p -> h[p + tap - 1], zero outside 0..npos-1"""
    out = np.zeros_like(h)
    npos = h.shape[0]
    s = tap - 1
    if s == 0:
        out[:] = h
    elif s > 0:
        out[:npos - s] = h[s:]
    else:
        out[-s:] = h[:npos + s]
    return out


def conv(h, W, b):
    """z[p][co] = b[co] + sum_tap sum_ci W[tap][co][ci] h[p + tap - 1][ci] (float64 sums)"""
    z = np.zeros((h.shape[0], W.shape[1]), np.float64) + np.asarray(b, np.float64)
    for tap in range(3):
        z += shifted(np.asarray(h, np.float64), tap) @ np.asarray(W[tap], np.float64).T
    return z


def conv_transpose(dz, W):
    """dh[p][ci] = sum_tap sum_co W[tap][co][ci] dz[p - tap + 1][co]"""
    dh = np.zeros((dz.shape[0], W.shape[2]), np.float64)
    for tap in range(3):
        dh += shifted(np.asarray(dz, np.float64), 2 - tap) @ np.asarray(W[tap], np.float64)
    return dh


def conv_weight_grad(dz, h):
    """dW[tap][co][ci] = sum_p dz[p][co] h[p + tap - 1][ci];  db[co] = sum_p dz[p][co]"""
    dW = np.stack([np.asarray(dz, np.float64).T @ shifted(np.asarray(h, np.float64), tap) for tap in range(3)])
    return dW, np.asarray(dz, np.float64).sum(axis=0)


def forward(shape, theta, x, exact=False):
    """y [npos][cout] and the cache of every layer; x [npos][cin]"""
    rnd = (lambda a: np.asarray(a, np.float64)) if exact else (lambda a: bf16(a))
    f32 = (lambda a: np.asarray(a, np.float64)) if exact else (lambda a: np.asarray(a, np.float32))
    Ws, bs = shape.unpack(np.asarray(theta, np.float32 if not exact else np.float64))
    npos = x.shape[0]
    h = np.zeros((npos, shape.cin_p[0]), np.float64)
    h[:, :shape.cin] = x
    h = rnd(h)
    cache = []
    for l in range(shape.layers):
        Wl = rnd(Ws[l])
        z = f32(conv(h, Wl, f32(bs[l])))
        cache.append((h, Wl, z))
        if l < shape.layers - 1:
            h = rnd(f32(silu(f32(z))))
    y = cache[-1][2][:, :shape.cout]
    return np.asarray(y, np.float64), cache


def backward(shape, cache, t, exact=False):
    """loss and the gradient of every parameter (flat layout) for target t"""
    rnd = (lambda a: np.asarray(a, np.float64)) if exact else (lambda a: bf16(a))
    f32 = (lambda a: np.asarray(a, np.float64)) if exact else (lambda a: np.asarray(a, np.float32))
    npos = t.shape[0]
    zL = cache[-1][2]
    d = np.zeros_like(zL, dtype=np.float64)
    d[:, :shape.cout] = np.asarray(zL[:, :shape.cout], np.float64) - t
    loss = float(np.mean(d[:, :shape.cout] ** 2))
    dz = rnd(f32(2.0 * d / (npos * shape.cout)))
    dWs, dbs = [None] * shape.layers, [None] * shape.layers
    for l in range(shape.layers - 1, -1, -1):
        h, Wl, _ = cache[l]
        dWs[l], dbs[l] = conv_weight_grad(dz, h)
        if l > 0:
            zp = cache[l - 1][2]
            dz = rnd(f32(f32(conv_transpose(dz, Wl)) * silu_grad(np.asarray(zp, np.float64))))
    return loss, shape.pack(dWs, dbs)


def adam_step(theta, m, v, g, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """one Adam step (t = 1, 2, ...): returns theta, m, v (float32)"""
    m = (b1 * m + (1.0 - b1) * g).astype(np.float32)
    v = (b2 * v + (1.0 - b2) * g * g).astype(np.float32)
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return (theta - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32), m, v


# ---------------------------------------------------------------------------
# slice inputs and target, from slice fields laid out [ncomp, nm, ld]
# ---------------------------------------------------------------------------
def slice_scales(b_spe, cu, q_spe, q_beam, nr):
    """maxima over modes and nodes 1..nr of |B_r|,|B_phi|; |J_r|,|J_phi|;
    |q_spe|; |q_beam| (1 where a quantity is identically zero)"""
    sl = slice(1, nr + 1)
    s = [np.abs(b_spe[0:2, :, sl]).max(), np.abs(cu[0:2, :, sl]).max(),
         np.abs(q_spe[0, :, sl]).max(), np.abs(q_beam[0, :, sl]).max()]
    return np.array([v if v > 0 else 1.0 for v in s])


def slice_inputs(b_spe, b_old, cu, q_spe, q_beam, nr):
    """x [nr][8 nm + 1]: channel qty * nm + k, then r / r_max.  The fields
    as slice xi-1 left them (b_spe, cu = the J predictor, q_spe), b_old =
    b_spe at the previous slice start, q_beam of slice xi."""
    nm = b_spe.shape[1]
    sB, sJ, sq, sqb = slice_scales(b_spe, cu, q_spe, q_beam, nr)
    sl = slice(1, nr + 1)
    qty = [b_spe[0, :, sl] / sB, b_spe[1, :, sl] / sB,
           (b_spe[0, :, sl] - b_old[0, :, sl]) / sB, (b_spe[1, :, sl] - b_old[1, :, sl]) / sB,
           cu[0, :, sl] / sJ, cu[1, :, sl] / sJ, q_spe[0, :, sl] / sq, q_beam[0, :, sl] / sqb]
    x = np.zeros((nr, 8 * nm + 1))
    for q, a in enumerate(qty):
        x[:, q * nm:(q + 1) * nm] = a.T
    x[:, 8 * nm] = np.arange(nr) / nr
    return x, sB


def apply_prediction(b_spe, b_old, y, sB, out_scale):
    """seed: b_spe components 0, 1 (nodes 1..nr) -> 2 b_spe - b_old + y
    out_scale sB (first-order extrapolation in xi plus the correction)"""
    nm = b_spe.shape[1]
    out = b_spe.copy()
    nr = y.shape[0]
    for c in range(2):
        sl = slice(1, nr + 1)
        out[c, :, sl] = 2.0 * b_spe[c, :, sl] - b_old[c, :, sl] + (y[:, c * nm:(c + 1) * nm] * (out_scale * sB)).T
    return out


def slice_target(b_conv, b_prev, b_old, sB, out_scale, nr):
    """(converged - (2 b_prev - b_old)) / (out_scale sB), clipped to [-1, 1]:
    t [nr][2 nm]"""
    nm = b_conv.shape[1]
    t = np.zeros((nr, 2 * nm))
    sl = slice(1, nr + 1)
    for c in range(2):
        t[:, c * nm:(c + 1) * nm] = ((b_conv[c, :, sl] - 2.0 * b_prev[c, :, sl] + b_old[c, :, sl]) /
                                     (out_scale * sB)).T
    return np.clip(t, -1.0, 1.0)
