"""Plasma density profiles and the deck's expression language on the host
(qpad_amd/profiles.py): fparser.f03's parsing rules and functions,
fdist2d_lib.f03's transverse and longitudinal profiles, each against values
fixed by hand or by the mathematics, and the profiled injection of the
oracle."""
import math

import numpy as np
import pytest

import pyoracle as po
from qpad_amd import QpadError
from qpad_amd.profiles import FParser, PlasmaProfile, plasma_lattice


def ev(expr, x=0.0, y=0.0, z=0.0):
    return float(FParser(expr)(x, y, z))


@pytest.mark.parametrize("expr,want", [
    ("1+2*3", 7.0), ("8/2/2", 2.0), ("2-3-4", -5.0), ("2^3^2", 64.0),        # left to right, '^' too
    ("-2^2", -4.0), ("(-2)^2", 4.0), ("2^2.9", 4.0), ("2^(-1.5)", 0.5),       # "^" takes int(exponent)
    ("pow(2,0.5)", math.sqrt(2.0)), ("-(1+2)*2", -6.0), ("-exp(0)", -1.0), ("+3", 3.0),
    ("1.5e-3*2", 3e-3), ("2.5E+2", 250.0), ("1d1+1", 11.0),
    ("3>2", 1.0), ("3<2", 0.0), ("2<=2", 1.0), ("2>=3", 0.0), ("2==2", 1.0), ("2!=2", 0.0),
    ("1&&0", 0.0), ("1||0", 1.0), ("1+1==2", 1.0), ("if(0,5,7)", 7.0), ("if(2>1,5,7)", 5.0),
    ("not(0)", 1.0), ("neg(3)", -3.0), ("abs(-2)", 2.0), ("int(-2.7)", -2.0), ("nint(2.5)", 3.0),
    ("nint(-2.5)", -3.0), ("ceiling(1.2)", 2.0), ("floor(-1.2)", -2.0), ("modulo(-1,3)", 2.0),
    ("modulo(1,-3)", -2.0), ("rect(0.5)", 1.0), ("rect(-0.51)", 0.0), ("step(0)", 1.0), ("step(-1e-9)", 0.0),
    ("min3(3,1,2)", 1.0), ("max3(3,1,2)", 3.0), ("min(4,5)", 4.0), ("max(4,5)", 5.0),
    ("atan2(1,-1)", 0.75 * math.pi), ("log10(1000)", 3.0), ("log(exp(2))", 2.0), ("sqrt(16)", 4.0),
    ("sinh(1)", math.sinh(1.0)), ("cosh(1)", math.cosh(1.0)), ("tanh(1)", math.tanh(1.0)),
    ("asin(1)", math.pi / 2), ("acos(1)", 0.0), ("atan(1)", math.pi / 4), ("tan(1)", math.tan(1.0)),
    ("sin(1)*cos(1)", math.sin(1.0) * math.cos(1.0)),
])
def test_fparser_known_answers(expr, want):
    assert ev(expr) == pytest.approx(want, rel=1e-15, abs=1e-15), expr


def test_fparser_variables_and_the_deck_expression():
    f = FParser("z^2/(z + 10.0)^2 * exp(-(x^2 + y^2)/50.0)")
    x, y, z = np.array([0.0, 1.0, 3.0]), np.array([0.5, -2.0, 4.0]), np.array([2.0, 0.0, 7.5])
    want = z ** 2 / (z + 10.0) ** 2 * np.exp(-(x ** 2 + y ** 2) / 50.0)
    assert np.allclose(f(x, y, z), want, rtol=1e-15, atol=0)
    assert ev("-x", x=2.0) == -2.0 and ev("-x^2", x=3.0) == -9.0
    g = FParser("exp(-z^2/0.5) * exp(-(x^2 + y^2)/8.0)")
    assert g(1.0, 2.0, 0.5) == pytest.approx(math.exp(-0.5) * math.exp(-5.0 / 8.0), rel=1e-15)


@pytest.mark.parametrize("bad", ["", "foo(1)", "1+", "x y", "(1", "q+1", "2^-1.5", "2*-3"])
def test_fparser_refuses(bad):
    with pytest.raises(QpadError):
        FParser(bad)(0.0, 0.0, 0.0)


def test_transverse_profiles():
    """get_den_perp_para_chl / hllw_chl (fdist2d_lib.f03:66-136)"""
    p = PlasmaProfile({"profile": ["parabolic-channel", "uniform"], "channel_n0": 1.0, "channel_depth": 0.5,
                       "channel_r0": 2.0, "channel_width": 3.0, "channel_rmax": 4.0})
    r = np.array([0.0, 1.0, 2.9, 3.5, 3.99, 4.01])
    want = np.array([1.0, 1.0 + 0.5 / 4.0, 1.0 + 0.5 * 2.9 ** 2 / 4.0, 1.0 + 0.5 * 9.0 / 4.0,
                     1.0 + 0.5 * 9.0 / 4.0, 0.0])
    assert np.allclose(p.den_perp(r / math.sqrt(2), r / math.sqrt(2), 0.0), want, rtol=1e-14, atol=0)
    h = PlasmaProfile({"profile": ["hollow-channel", "uniform"], "channel_rmin": 1.0, "channel_rmax": 2.0,
                       "channel_depth": 3.0})
    assert list(h.den_perp(np.array([0.5, 1.0, 1.5, 2.0, 2.5]), np.zeros(5), 0.0)) == [0.0, 0.0, 3.0, 0.0, 0.0]
    a = PlasmaProfile({"profile": ["analytic", "analytic"], "math_func": "1+x*y+z"})
    assert a.den_perp(np.array([2.0]), np.array([3.0]), 0.5)[0] == 7.5 and a.den_lon(123.0) == 1.0


def test_longitudinal_profiles():
    """get_den_lon_pw_linear / sine / cubic_spline (fdist2d_lib.f03:236-415)"""
    pw = PlasmaProfile({"profile": ["uniform", "piecewise-linear"], "piecewise_s": [0.0, 10.0, 30.0],
                        "piecewise_fs": [0.0, 1.0, 0.5]})
    got = [pw.den_lon(s) for s in (-1.0, 0.0, 5.0, 10.0, 20.0, 30.0, 30.5)]
    assert got == pytest.approx([0.0, 0.0, 0.5, 1.0, 0.75, 0.5, 0.0], abs=1e-15)
    sn = PlasmaProfile({"profile": ["uniform", "sine-series"], "start": 2.0, "length": 4.0, "sine_coefs": [1.0, -3.0]})
    for s in (1.0, 2.0, 3.0, 5.5, 6.0, 6.5):
        want = abs(math.sin(math.pi * (s - 2) / 4) - 3 * math.sin(2 * math.pi * (s - 2) / 4)) if 2 < s <= 6 else 0.0
        assert sn.den_lon(s) == pytest.approx(want, abs=1e-15)
    # a cubic spline through points of a line is that line, for every end condition
    for kind in ("natural", "clamped", "not-a-knot"):
        sp = PlasmaProfile({"profile": ["uniform", "cubic-spline"], "piecewise_s": [0.0, 1.0, 2.0, 3.0, 4.0],
                            "piecewise_fs": [1.0, 3.0, 5.0, 7.0, 9.0], "endpoints_type": kind})
        for s in (0.5, 1.0, 1.7, 3.99):
            v = sp.den_lon(s)
            if kind == "clamped":
                assert 1.0 <= v <= 9.0 and (s != 1.0 or v == pytest.approx(3.0))
            else:
                assert v == pytest.approx(1.0 + 2.0 * s, rel=1e-13), (kind, s)
        assert sp.den_lon(0.0) == 0.0 and sp.den_lon(4.5) == 0.0
    # not-a-knot: the spline's node slopes solve QPAD's rows (set_prof_lon_cubic_spline,
    # fdist2d_lib.f03:360-366); its first row is the not-a-knot condition, which a
    # cubic's own slopes meet, while the last repeats the first row's stencil
    # unmirrored (-5 f(n-2) + 4 f(n-1) + f(n)), which they do not: restated as is
    s = np.arange(6.0)
    f = 0.3 * s ** 3 - s ** 2 + 2.0
    from qpad_amd.profiles import _spline_slopes
    d = _spline_slopes(s, f, "not-a-knot")
    A = np.diag(np.full(6, 4.0)) + np.diag(np.ones(5), 1) + np.diag(np.ones(5), -1)
    A[0, 0] = A[5, 5] = 2.0
    A[0, 1] = A[5, 4] = 4.0
    rhs = np.concatenate([[-5 * f[0] + 4 * f[1] + f[2]], 3 * (f[2:] - f[:-2]), [-5 * f[3] + 4 * f[4] + f[5]]])
    assert np.allclose(d, np.linalg.solve(A, rhs), rtol=1e-13, atol=1e-13)
    exact = 0.9 * s ** 2 - 2 * s
    assert 2 * exact[0] + 4 * exact[1] == pytest.approx(rhs[0]) and 4 * exact[4] + 2 * exact[5] != pytest.approx(rhs[5])
    # clamped: zero slopes at both ends
    cl = PlasmaProfile({"profile": ["uniform", "cubic-spline"], "piecewise_s": list(s), "piecewise_fs": list(f),
                        "endpoints_type": "clamped"})
    eps = 1e-6
    assert abs(cl.den_lon(eps) - f[0]) <= 1e-9 and abs(cl.den_lon(5.0) - f[5]) <= 1e-12
    with pytest.raises(QpadError, match="monotonically"):
        PlasmaProfile({"profile": ["uniform", "piecewise-linear"], "piecewise_s": [0.0, 0.0], "piecewise_fs": [1, 1]})


def test_oracle_profiled_injection():
    """the profiled injection (inject_fdist2d, fdist2d_class.f03:289-357):
    the lattice points with den_lon den_perp density >= den_min, in lattice
    order with consecutive ids, charge rn den_perp den_lon density coef, and
    a deposited charge density that follows n(r) to O(dr^2)"""
    nr, dr, ppc, nth = 40, 0.1, (2, 2), 8
    o = po.Oracle(po.make_grid(nr, 4, 1, dr, 0.05))
    isp = o.add_species(po.species_cfg(ppc[0], ppc[1], nth, den_min=0.2))
    prof = PlasmaProfile({"profile": ["parabolic-channel", "piecewise-linear"], "channel_n0": 0.2,
                          "channel_depth": 1.0, "channel_r0": 3.0, "channel_width": 3.5, "channel_rmax": 3.2,
                          "piecewise_s": [0.0, 10.0], "piecewise_fs": [0.0, 2.0]})
    x1, x2 = plasma_lattice(nr, dr, ppc[0], ppc[1], nth)
    s = 4.0
    den, dl = prof.den_perp(x1, x2, s), prof.den_lon(s)
    assert dl == pytest.approx(0.8)
    o.set_profile(isp, den, dl)
    o.renew_species(isp)
    sp = o.species(isp)
    keep = ~(dl * den * 1.0 < 0.2)
    assert 0 < keep.sum() < keep.size and len(sp["x1"]) == keep.sum()
    assert np.array_equal(sp["id"], np.arange(keep.sum()))
    assert np.array_equal(sp["x1"], x1[keep]) or np.allclose(sp["x1"], x1[keep], rtol=0, atol=1e-15)
    rn = np.sqrt(x1 ** 2 + x2 ** 2)[keep] / dr
    assert np.allclose(sp["q"], -rn * den[keep] * dl / (ppc[0] * ppc[1] * nth), rtol=1e-14, atol=0)
    q = np.zeros((1, o.nm, o.ld))
    o.L.ora_qdeposit(o.h, o.species_raw(isp), po.ptr(q))
    r = (np.arange(o.ld) - 1) * dr
    n = dl * (0.2 + r ** 2 / 9.0)
    inner = (r > 0.9) & (r < 3.0)        # every particle near these nodes kept
    assert np.max(np.abs(-q[0, 0, inner] - n[inner])) <= 2.0 * dr ** 2 / 9.0 + 1e-12
    o.close()


def test_beam_profiles():
    """get_den_super_gauss / parabolic / rational / analytic
    (fdist3d_std_lib.f03:46-286): closed forms, z of the analytic function
    absolute (the box start added back)"""
    from qpad_amd.profiles import BeamProfile
    bm = {"profile": ["super-gaussian", "parabolic", "rational"], "super_gauss_center": [0.5, 0, 0],
          "super_gauss_sigma": [2.0, 1, 1], "super_gauss_order": [2.0, 1, 1], "parabolic_center": [0, 1.0, 0],
          "parabolic_radius": [1, 2.0, 1], "poly_x0": [0, 0, 3.0], "poly_numerator3": [1.0, 2.0],
          "poly_denominator3": [2.0, 0.0, 1.0]}
    p = BeamProfile(bm, z0=-2.0)
    x = np.array([0.5, 2.5, -1.5])
    y = np.array([1.0, 2.0, 3.5])
    z = np.array([5.0, 6.0, 4.0])            # box-relative: absolute z = z - 2
    zr = z - 5.0                              # x - x0 with x0 = 3 - z0
    want = (np.exp(-(0.5 * ((x - 0.5) / 2.0) ** 2) ** 2) * np.where((y - 1.0) ** 2 / 4.0 < 1, 1 - (y - 1.0) ** 2 / 4.0, 0)
            * (1.0 + 2.0 * zr) / (2.0 + zr ** 2))
    assert np.allclose(p.density(x, y, z), want, rtol=1e-14, atol=0) and p.tabulated
    a = BeamProfile({"profile": ["analytic"] * 3, "math_func": "x + 10*y + 100*z"}, z0=-2.0)
    assert a.density(np.array([1.0]), np.array([2.0]), np.array([3.0]))[0] == pytest.approx(1 + 20 + 100 * 1.0)
    with pytest.raises(QpadError, match="super-Gaussian"):
        BeamProfile(dict(bm, super_gauss_order=[0.5, 1, 1]), z0=0.0)


def test_oracle_tabulated_beam_equals_profile_beam():
    """a lattice beam given by its density table (the host's Gaussian) is the
    beam the oracle's own Gaussian profile injects: same particles, charges
    to rounding"""
    from qpad_amd import configs
    from qpad_amd.input import _lattice_beam, grid_from_deck
    from qpad_amd.profiles import BeamProfile, beam_lattice_box, beam_lattice_density
    import twin
    deck = configs.tri_gaussian()
    g, _, z0 = grid_from_deck(deck)
    bm = deck["beam"][0]
    cfg = _lattice_beam(bm, z0, seed=17)
    prof = BeamProfile(bm, z0)
    box = beam_lattice_box(g.nr, g.nxi, g.dr, g.dxi, cfg)
    den = beam_lattice_density(prof, cfg, g.nr, g.dr, g.dxi, box)
    og = twin._as_oracle_struct(po.Grid, g)
    a, b = po.Oracle(og), po.Oracle(og)
    a.add_beam_std(twin._as_oracle_struct(po.BeamStdCfg, cfg))
    b.add_beam_std_table(twin._as_oracle_struct(po.BeamStdCfg, cfg), den, *box)
    ba, bb = a.beam(0), b.beam(0)
    assert len(ba["x1"]) == len(bb["x1"]) > 1000
    for k in ("x1", "x2", "x3", "p1", "p2", "p3"):
        assert np.array_equal(ba[k], bb[k]), k
    assert np.allclose(ba["q"], bb["q"], rtol=1e-14, atol=0)
    a.close()
    b.close()
