"""Pins of the CPU oracle against mathematics and published definitions.

The reference ships no golden vectors for this path (its test/ programs are
stale and assertion-free), so each oracle routine is checked against
something that fixes it independently of the restatement:
  * Philox4x32-10: the Random123 known-answer vectors;
  * tridiagonal solves: dense LU of the assembled matrices;
  * field stencils: manufactured solutions with second-order convergence;
  * open boundaries: the vacuum fall-off each boundary row encodes;
  * mode convention: f = f0 + 2 sum_m (Re f_m cos m phi - Im f_m sin m phi)
    (tools/merge_modes_tools/helper.py:85-91) and the e^{-im phi} deposit;
  * beam charge normalisation: total_charge (session_beam.rst:145-150);
  * pushers: invariants (zero field, pure magnetic rotation);
  * the coupled slice loop: linear wakefield theory for a weak wide beam.
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

import pyoracle as po


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert po.philox4x32([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert po.philox4x32([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                                 0x6D5451FD]
    assert po.philox4x32([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                         [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_normal_draws_are_standard_normal():
    x = np.concatenate([po.normal4(7, s, 0) for s in range(20000)])
    assert abs(x.mean()) < 0.02 and abs(x.std() - 1.0) < 0.02
    # fourth moment of N(0,1) is 3
    assert abs(np.mean(x ** 4) - 3.0) < 0.15


@pytest.mark.parametrize("kind", range(6))
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_thomas_matches_dense_lu(kind, mode):
    nr, dr = 96, 0.05
    a, b, c = po.matrix(kind, mode, nr, dr, 1e-3 * (dr / 0.02) ** 2)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    rng = np.random.default_rng(kind * 10 + mode)
    d = rng.standard_normal(nr)
    x = po.thomas(a, b, c, d)
    ref = np.linalg.solve(A, d)
    assert np.max(np.abs(x - ref)) <= 1e-9 * np.max(np.abs(ref))
    assert np.max(np.abs(A @ x - d)) <= 1e-9 * np.max(np.abs(d))


def _oracle(nr, dr, max_mode, nxi=4, dxi=0.05):
    return po.Oracle(po.make_grid(nr, nxi, max_mode, dr, dxi))


def _lap_gauss(r, m, w):
    """(d2/dr2 + 1/r d/dr - m^2/r^2) [r^m exp(-r^2/w^2)]"""
    s = (r / w) ** 2
    return 4.0 / w ** 2 * r ** m * np.exp(-s) * (s - m - 1.0)


@pytest.mark.parametrize("m", [0, 1, 2, 3])
def test_psi_manufactured_second_order(m):
    """psi solve of q = -lap_m(r^m e^{-r^2/w^2}) converges to it at O(dr^2)."""
    w, R = 1.0, 6.0
    errs = []
    for nr in (64, 128, 256):
        dr = R / nr
        o = _oracle(nr, dr, 3)
        q = o.new_field(1)
        psi = o.new_field(1)
        r = (np.arange(nr + 2) - 1) * dr            # node i at r = (i-1) dr
        k = 0 if m == 0 else 2 * m - 1
        src = -_lap_gauss(np.maximum(r, 0.0), m, w)
        if m == 0:
            src[1] = 4.0 / w ** 2                 # -lap at r = 0 (limit)
        q[0, k, :] = src
        q[0, k, 0] = 0.0
        if m > 0:
            q[0, k + 1, :] = 0.5 * src            # Im part: half amplitude
        o.solve("psi", q, psi)
        exact = np.maximum(r, 0.0) ** m * np.exp(-(r / w) ** 2)
        e = np.max(np.abs(psi[0, k, 1:nr + 1] - exact[1:nr + 1]))
        if m > 0:
            e = max(e, np.max(np.abs(psi[0, k + 1, 1:nr + 1] - 0.5 * exact[1:nr + 1])))
        errs.append(e)
    assert errs[-1] < 2e-3
    # second order: halving dr divides the error by about four
    assert 3.0 < errs[0] / errs[1] < 5.0 and 3.0 < errs[1] / errs[2] < 5.0


@pytest.mark.parametrize("m", [1, 2, 3])
def test_open_boundary_vacuum_falloff(m):
    """Outside a compact source the m > 0 psi solution falls as r^-m: the open
    boundary row encodes x_{N+1} = (1 - m/N) x_N."""
    nr, R = 200, 10.0
    dr = R / nr
    o = _oracle(nr, dr, 3)
    q = o.new_field(1)
    psi = o.new_field(1)
    r = (np.arange(nr + 2) - 1) * dr
    k = 2 * m - 1
    q[0, k, 2:] = np.exp(-((r[2:] - 1.0) / 0.2) ** 2)
    o.solve("psi", q, psi)
    i1, i2 = int(5.0 / dr) + 1, int(9.5 / dr) + 1
    ratio = psi[0, k, i2] / psi[0, k, i1]
    assert ratio == pytest.approx((r[i1] / r[i2]) ** m, rel=0.02)


def test_beam_bt_open_boundary_log():
    """m = 0 beam potential outside the beam follows ln r (the bt boundary
    row folds x_{N+1} = (1 + 1/(N ln(N dr))) x_N), so B_phi ~ 1/r."""
    nr, R = 200, 10.0
    dr = R / nr
    o = _oracle(nr, dr, 1)
    q = o.new_field(1)
    b = o.new_field(3)
    r = (np.arange(nr + 2) - 1) * dr
    q[0, 0, 1:] = -np.exp(-(r[1:] / 0.5) ** 2)
    q[0, 0, 1] = -1.0
    o.solve("bt_beam", q, b)
    i1, i2 = int(4.0 / dr) + 1, int(9.0 / dr) + 1
    # B_phi r is constant in vacuum (enclosed current)
    assert b[1, 0, i2] * r[i2] == pytest.approx(b[1, 0, i1] * r[i1], rel=0.01)
    # Ampere for a beam moving at c: lap A = -rho and B_phi = -dA/dr give
    # r B_phi = int_0^r rho r' dr' (negative for this electron beam)
    enclosed = -0.5 ** 2 / 2.0
    assert b[1, 0, i1] * r[i1] == pytest.approx(enclosed, rel=0.02)


def test_mode_convention_deposit_and_gather():
    """A ring of unit charges modulated by 1 + a cos(phi - phi0) deposits
    m = 1 content (P a / 2) e^{-i phi0} / j; gathering a pure m = 1 field
    returns 2 (Re f cos phi - Im f sin phi)."""
    nr, dr, M = 32, 0.1, 2
    o = po.Oracle(po.make_grid(nr, 4, M, dr, 0.05))
    o.add_species(po.species_cfg(1, 1, 8))
    P, a, phi0, j0 = 64, 0.3, 0.7, 10
    phi = 2 * np.pi * (np.arange(P) + 0.25) / P
    r0 = (j0 - 1) * dr + 1e-12                  # on node j0
    arrs = {"x1": r0 * np.cos(phi), "x2": r0 * np.sin(phi),
            "p1": np.zeros(P), "p2": np.zeros(P), "p3": np.zeros(P),
            "q": 1.0 + a * np.cos(phi - phi0), "gamma": np.ones(P), "psi": np.zeros(P)}
    o.set_species(0, arrs)
    q = o.new_field(1)
    po.lib().ora_qdeposit(o.h, o.species_raw(0), po.ptr(q))
    expect = P * a / 2 / (j0 - 1)
    assert q[0, 1, j0] == pytest.approx(expect * math.cos(phi0), rel=1e-9)
    assert q[0, 2, j0] == pytest.approx(-expect * math.sin(phi0), rel=1e-9)
    assert q[0, 0, j0] == pytest.approx(P / (j0 - 1), rel=1e-9)
    assert abs(q[0, 3, j0]) < 1e-9 and abs(q[0, 4, j0]) < 1e-9
    # gather
    f = o.new_field(3)
    f[:, 1, :] = 0.4
    f[:, 2, :] = -0.9
    out = np.zeros(3)
    for ph in (0.1, 1.3, 2.9, 4.4):
        x1, x2 = 0.55 * math.cos(ph), 0.55 * math.sin(ph)
        po.lib().ora_interp_point(o.h, po.ptr(f), 3, x1, x2, po.ptr(out))
        assert out[0] == pytest.approx(2 * (0.4 * math.cos(ph) + 0.9 * math.sin(ph)), rel=1e-12)


def test_beam_charge_normalisation():
    """Deposit conserves charge exactly (axis node x8, node i / (i-1)) and the
    per-particle charge makes the m = 0 deposit integrate to total_charge:
    sum_p q_p 2 pi dr^2 dxi = -Q (session_beam.rst:145-150,
    fdist3d_rnd_class.f03:381)."""
    nr, nxi, dr, dxi = 64, 64, 0.05, 0.05
    o = po.Oracle(po.make_grid(nr, nxi, 1, dr, dxi))
    sig = (0.3, 0.3, 0.4)
    Q = 4.0 * (2 * np.pi) ** 1.5 * sig[0] * sig[1] * sig[2]
    cfg = po.beam_cfg(1 << 15, Q, (0.0, 0.0, 1.6), sig, (-1.5, -1.5, 0.0), (1.5, 1.5, 3.2))
    o.add_beam(cfg)
    o.step_begin()
    h = o.hist(po.Q_BEAM)
    raw = 0.0
    for s in range(1, nxi + 2):
        corr = s <= nxi                               # guard slice is not corrected
        raw += h[s, 0, 0, 1] / 8.0 if corr else h[s, 0, 0, 1]
        w = (np.arange(2, nr + 2) - 1.0) if corr else np.ones(nr)
        raw += np.sum(h[s, 0, 0, 2:nr + 2] * w)
    qsum = np.sum(o.beam(0)["q"])
    assert raw == pytest.approx(qsum, rel=1e-12)
    assert qsum * 2 * np.pi * dr * dr * dxi == pytest.approx(-Q, rel=1e-3)


def test_push_invariants():
    nr, dr, dxi = 16, 0.1, 0.05
    o = po.Oracle(po.make_grid(nr, 4, 1, dr, dxi))
    o.add_species(po.species_cfg(1, 1, 4))
    rng = np.random.default_rng(3)
    n = 32
    x1, x2 = rng.uniform(-0.8, 0.8, n), rng.uniform(-0.8, 0.8, n)
    p = rng.normal(0, 0.3, (3, n))
    arrs = {"x1": x1, "x2": x2, "p1": p[0], "p2": p[1], "p3": p[2], "q": np.ones(n),
            "gamma": np.sqrt(1 + (p ** 2).sum(0)), "psi": np.zeros(n)}
    e, b = o.new_field(3), o.new_field(3)
    # zero fields: momentum unchanged, x advances by p dxi / (gamma - pz)
    o.set_species(0, arrs)
    sp = o.species_raw(0)
    po.lib().ora_push_u_robust(o.h, sp, po.ptr(e), po.ptr(b), dxi)
    po.lib().ora_push_x(sp, dxi)
    s = o.species(0)
    g = np.sqrt(1 + (p ** 2).sum(0))
    assert np.allclose(s["p1"], p[0], rtol=0, atol=1e-15)
    assert np.allclose(s["x1"], x1 + p[0] * dxi / (g - p[2]), rtol=1e-14)
    # uniform Bz only: the Boris rotation conserves |p| and p_z
    b[2, 0, :] = 1.7
    o.set_species(0, arrs)
    po.lib().ora_push_u_robust(o.h, sp, po.ptr(e), po.ptr(b), dxi)
    s = o.species(0)
    pn = np.sqrt(s["p1"] ** 2 + s["p2"] ** 2 + s["p3"] ** 2)
    assert np.allclose(pn, np.sqrt((p ** 2).sum(0)), rtol=1e-13)
    assert np.allclose(s["p3"], p[2], rtol=1e-13)
    # electrons (q < 0) in E_z > 0 lose longitudinal momentum
    b[2, 0, :] = 0.0
    e[2, 0, :] = 0.5
    o.set_species(0, arrs)
    po.lib().ora_push_u_robust(o.h, sp, po.ptr(e), po.ptr(b), dxi)
    assert np.all(o.species(0)["p3"] < p[2])


def test_convergence_tester_hand_example():
    """record / compare on B_phi (dim = 2) over modes, infinity norm of
    sqrt(re^2 + im^2) (simulation_class.f03:549-599)."""
    nr = 3
    o = po.Oracle(po.make_grid(8, 4, 1, 0.1, 0.05))
    b = o.new_field(3)
    # nodes 1..3 only matter; use an 8-node grid and leave 4..8 zero
    b[1, 0, 1:4] = [1.0, -2.0, 0.5]     # Re m0
    b[1, 1, 1:4] = [0.0, 1.0, -1.0]     # Re m1
    b[1, 2, 1:4] = [3.0, 0.0, 0.0]      # Im m1
    po.lib().ora_conv_record(o.h, po.ptr(b))
    b2 = b.copy()
    b2[1, 0, 2] = -2.5
    b2[1, 2, 1] = 2.0
    rel, ab = np.zeros(1), np.zeros(1)
    po.lib().ora_conv_compare(o.h, po.ptr(b2), po.ptr(rel), po.ptr(ab))
    # old: re = [1, 3, 1.5], im = [3, 0, 0] -> norm = max sqrt(10, 9, 2.25) = sqrt(10)
    # diff: re = [0, -0.5, 0], im = [1, 0, 0] -> abs = 1
    assert ab[0] == pytest.approx(1.0, rel=1e-15)
    assert rel[0] == pytest.approx(1.0 / math.sqrt(10.0), rel=1e-15)
    del nr


def _form_factor(sigma_r):
    """On-axis linear response of a Gaussian transverse profile relative to
    the 1D limit: int_0^inf K0(r) r exp(-r^2 / 2 sigma^2) dr."""
    val, _ = integrate.quad(lambda r: special.k0(r) * r * math.exp(-r * r / (2 * sigma_r ** 2)),
                            0, 60, limit=400)
    return val


def test_linear_wake_weak_wide_beam():
    """Weak beam (n_b = 0.01), sigma_r = 2: on-axis Ez behind the beam follows
    linear theory E_z(xi) = F(sigma_r) int n_b(xi') cos(xi - xi') dxi'."""
    nr, rmax, nxi, dxi = 192, 16.0, 300, 0.05
    dr = rmax / nr
    o = po.Oracle(po.make_grid(nr, nxi, 0, dr, dxi, iter_max=20, reltol=1e-6, abstol=1e-9))
    o.add_species(po.species_cfg(2, 1, 8))
    nb0, sr, sz, zc = 0.01, 2.0, 0.5, 2.5
    o.step_begin()
    r = (np.arange(nr + 2) - 1) * dr
    xi = (np.arange(nxi + 2) - 1) * dxi
    h = o.hist(po.Q_BEAM)
    prof = np.exp(-(r ** 2) / (2 * sr ** 2))
    for s in range(1, nxi + 1):
        h[s, 0, 0, :] = -nb0 * prof * math.exp(-((xi[s] - zc) ** 2) / (2 * sz ** 2))
    h[:, 0, 0, 0] = 0.0
    for j in range(1, nxi + 1):
        o.slice(j)
    ez = o.hist(po.E)[1:nxi + 1, 2, 0, 1]
    nb_xi = nb0 * np.exp(-((xi[1:nxi + 1] - zc) ** 2) / (2 * sz ** 2))
    F = _form_factor(sr)
    lin = np.array([F * np.sum(nb_xi[:j + 1] * np.cos(xi[j + 1] - xi[1:j + 2])) * dxi
                    for j in range(nxi)])
    behind = xi[1:nxi + 1] > zc + 4 * sz
    amp = np.max(np.abs(lin[behind]))
    err = np.max(np.abs(ez[behind] - lin[behind])) / amp
    print(f"linear wake: amplitude {amp:.4e}, max deviation {err:.3%}")
    assert err < 0.05
    # the sign convention: the wake opposes the beam head (decelerating)
    assert ez[np.argmin(np.abs(xi[1:nxi + 1] - zc))] > 0
