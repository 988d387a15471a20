"""Plasma density profiles of the input deck, evaluated on the host at the
injection lattice (fdist2d_lib.f03, fdist2d_class.f03:289-357), and the
analytic-expression language of the decks' `math_func` (fparser.f03).

The library injects the plasma lattice itself; a non-uniform species hands it
the transverse density of every lattice point and the longitudinal factor of
the injection's propagation distance s (qpad_species_set_profile), which the
host re-evaluates before each renew (simulation_class.f03:497-501).
"""
import math

import numpy as np

from ._lib import QpadError

# ---------------------------------------------------------------------------
# fparser.f03: an expression of the variables (x, y, z), compiled by the same
# rules: blanks removed; a leading '+' dropped; '(...)', 'f(...)', '-(...)'
# and '-f(...)' recognised first; otherwise the string is split at the
# rightmost operator outside parentheses of the lowest-priority group that
# occurs (so operators associate to the left, '^' included), a leading '-'
# before a '^' split negating the whole power; numbers with a signed exponent
# are not split.  '^' raises to the integer part of its right operand; the
# comparisons and logical operators give 1 or 0.
# ---------------------------------------------------------------------------
_OPER_GROUPS = [["||"], ["&&"], ["==", "!="], ["<=", ">=", "<", ">"], ["+", "-"], ["*", "/"], ["^"]]

_FUNCS = [("abs", 1), ("sinh", 1), ("sin", 1), ("cosh", 1), ("cos", 1), ("tanh", 1), ("tan", 1), ("exp", 1),
          ("log10", 1), ("log", 1), ("asin", 1), ("acos", 1), ("atan2", 2), ("atan", 1), ("sqrt", 1), ("not", 1),
          ("neg", 1), ("if", 3), ("pow", 2), ("int", 1), ("nint", 1), ("ceiling", 1), ("floor", 1),
          ("modulo", 2), ("rect", 1), ("step", 1), ("min3", 3), ("min", 2), ("max3", 3), ("max", 2)]


def _b(v):
    return np.where(v, 1.0, 0.0)


def _fortran_int(v):
    return np.trunc(v)


def _nint(v):
    # Fortran nint: halves away from zero
    return np.sign(v) * np.floor(np.abs(v) + 0.5)


def _modulo(a, b):
    # Fortran modulo: the result has the sign of b
    return a - np.floor(a / b) * b


def _ipow(a, b):
    e = np.trunc(b)
    return np.power(a, e)


_OPS = {"||": lambda a, b: _b((a != 0) | (b != 0)), "&&": lambda a, b: _b((a != 0) & (b != 0)),
        "==": lambda a, b: _b(a == b), "!=": lambda a, b: _b(a != b), "<=": lambda a, b: _b(a <= b),
        ">=": lambda a, b: _b(a >= b), "<": lambda a, b: _b(a < b), ">": lambda a, b: _b(a > b),
        "+": lambda a, b: a + b, "-": lambda a, b: a - b, "*": lambda a, b: a * b, "/": lambda a, b: a / b,
        "^": _ipow}

_FN = {"abs": np.abs, "sinh": np.sinh, "sin": np.sin, "cosh": np.cosh, "cos": np.cos, "tanh": np.tanh,
       "tan": np.tan, "exp": np.exp, "log10": np.log10, "log": np.log, "asin": np.arcsin, "acos": np.arccos,
       "atan2": np.arctan2, "atan": np.arctan, "sqrt": np.sqrt, "not": lambda a: _b(a == 0),
       "neg": np.negative, "if": lambda c, a, b: np.where(c != 0, a, b), "pow": np.power,
       "int": _fortran_int, "nint": _nint, "ceiling": np.ceil, "floor": np.floor, "modulo": _modulo,
       "rect": lambda a: _b(np.abs(a) <= 0.5), "step": lambda a: _b(a >= 0),
       "min3": lambda a, b, c: np.minimum(np.minimum(a, b), c), "min": np.minimum,
       "max3": lambda a, b, c: np.maximum(np.maximum(a, b), c), "max": np.maximum}


def _enclosed(s):
    """if_enclosed_parent: '(' ... ')' with the outer pair matching"""
    if len(s) < 2 or s[0] != "(" or s[-1] != ")":
        return False
    k = 0
    for ch in s[1:-1]:
        k += ch == "("
        k -= ch == ")"
        if k < 0:
            return False
    return k == 0


def _function_at(s):
    for name, n in _FUNCS:
        if s.startswith(name):
            return name, n
    return None, 0


def _split_params(body, n):
    """the n comma-separated parameters of a function's argument list, the
    commas at the first parenthesis level"""
    out, k, start = [], 0, 0
    for i, ch in enumerate(body):
        if ch == "(":
            k += 1
        elif ch == ")":
            k -= 1
        elif ch == "," and k == 0 and len(out) < n - 1:
            out.append(body[start:i])
            start = i + 1
    out.append(body[start:])
    if len(out) != n:
        raise QpadError(f"math_func: {n} parameters expected in '{body}'")
    return out


def _signed_exp(s, i):
    return i > 1 and s[i] in "+-" and s[i - 1] in "edED" and s[i - 2] in "0123456789."


class FParser:
    """A compiled math_func of the variables `names` (default x, y, z)."""

    def __init__(self, expr, names=("x", "y", "z")):
        self.expr = "".join(str(expr).split())
        self.names = tuple(names)
        if not self.expr:
            raise QpadError("math_func: empty expression")
        self.tree = self._compile(self.expr)

    def _compile(self, s):
        if not s:
            raise QpadError(f"math_func: missing operand in '{self.expr}'")
        if s[0] == "+":
            return self._compile(s[1:])
        if _enclosed(s):
            return self._compile(s[1:-1])
        if s[0].isalpha():
            name, n = _function_at(s)
            if name and _enclosed(s[len(name):]):
                return ("fn", name, [self._compile(p) for p in _split_params(s[len(name) + 1:-1], n)])
        elif s[0] == "-":
            if _enclosed(s[1:]):
                return ("fn", "neg", [self._compile(s[2:-1])])
            if len(s) > 1 and s[1].isalpha():
                name, n = _function_at(s[1:])
                if name and _enclosed(s[1 + len(name):]):
                    inner = ("fn", name, [self._compile(p) for p in _split_params(s[2 + len(name):-1], n)])
                    return ("fn", "neg", [inner])
        for gi, group in enumerate(_OPER_GROUPS):
            k = 0
            for i in range(len(s) - 1, 0, -1):
                ch = s[i]
                if ch == ")":
                    k += 1
                elif ch == "(":
                    k -= 1
                if k != 0:
                    continue
                op = next((o for o in group if s.startswith(o, i)), None)
                if op is None or _signed_exp(s, i):
                    continue
                if s[0] == "-" and gi == len(_OPER_GROUPS) - 1:
                    return ("fn", "neg", [("op", op, self._compile(s[1:i]), self._compile(s[i + len(op):]))])
                return ("op", op, self._compile(s[:i]), self._compile(s[i + len(op):]))
        neg = s[0] == "-"
        body = s[1:] if neg else s
        var = next((v for v in self.names if body.startswith(v)), None)
        if var is not None:
            if body != var:
                raise QpadError(f"math_func: cannot read '{s}' in '{self.expr}'")
            node = ("var", self.names.index(var))
            return ("fn", "neg", [node]) if neg else node
        try:
            return ("num", float(s.replace("d", "e").replace("D", "e")))
        except ValueError:
            raise QpadError(f"math_func: cannot read '{s}' in '{self.expr}'") from None

    def _eval(self, node, vals):
        kind = node[0]
        if kind == "num":
            return node[1]
        if kind == "var":
            return vals[node[1]]
        if kind == "op":
            return _OPS[node[1]](np.asarray(self._eval(node[2], vals), dtype=float),
                                 np.asarray(self._eval(node[3], vals), dtype=float))
        return _FN[node[1]](*[np.asarray(self._eval(a, vals), dtype=float) for a in node[2]])

    def __call__(self, *vals):
        with np.errstate(all="ignore"):
            return np.asarray(self._eval(self.tree, [np.asarray(v, dtype=float) for v in vals]), dtype=float)


# ---------------------------------------------------------------------------
# transverse profiles (fdist2d_lib.f03:28-171): den_perp(x, y, s)
# ---------------------------------------------------------------------------
def _deck_get(sec, key):
    if key not in sec:
        raise QpadError(f"input deck: missing key '{key}'")
    return sec[key]


class PlasmaProfile:
    """species(i).profile = [transverse, longitudinal] with its parameters
    (fdist2d_class.f03:140-215)."""

    PERP = ("uniform", "parabolic-channel", "hollow-channel", "analytic")
    LON = ("uniform", "piecewise-linear", "sine-series", "cubic-spline", "analytic")

    def __init__(self, sec):
        prof = list(_deck_get(sec, "profile"))
        if len(prof) != 2 or prof[0] not in self.PERP or prof[1] not in self.LON:
            raise QpadError("Invalid density profile! transverse: uniform, parabolic-channel, hollow-channel, "
                            "analytic; longitudinal: uniform, piecewise-linear, sine-series, cubic-spline, "
                            f"analytic (got {prof})")
        self.perp, self.lon = prof
        self.func = None
        if self.perp == "parabolic-channel":
            self.pp = [float(_deck_get(sec, k)) for k in ("channel_n0", "channel_depth", "channel_r0",
                                                           "channel_width", "channel_rmax")]
        elif self.perp == "hollow-channel":
            self.pp = [float(_deck_get(sec, k)) for k in ("channel_rmin", "channel_rmax", "channel_depth")]
        elif self.perp == "analytic" or self.lon == "analytic":
            self.func = FParser(_deck_get(sec, "math_func"))
        if self.lon in ("piecewise-linear", "cubic-spline"):
            self.s = np.asarray(_deck_get(sec, "piecewise_s"), dtype=float)
            self.fs = np.asarray(_deck_get(sec, "piecewise_fs"), dtype=float)
            if len(self.s) != len(self.fs) or len(self.s) < 2 or np.any(np.diff(self.s) <= 0):
                raise QpadError("Array of longitudinal position must be monotonically increasing!")
            if self.lon == "cubic-spline":
                self.dfs = _spline_slopes(self.s, self.fs, _deck_get(sec, "endpoints_type"))
        elif self.lon == "sine-series":
            self.start = float(_deck_get(sec, "start"))
            self.length = float(_deck_get(sec, "length"))
            self.coefs = np.asarray(_deck_get(sec, "sine_coefs"), dtype=float)

    @property
    def uniform(self):
        return self.perp == "uniform" and self.lon in ("uniform", "analytic")

    def den_perp(self, x1, x2, s):
        """get_den_perp_* (fdist2d_lib.f03:37-171) at the points (x1, x2)"""
        x1 = np.asarray(x1, dtype=float)
        x2 = np.asarray(x2, dtype=float)
        if self.perp == "uniform":
            return np.ones_like(x1)
        if self.perp == "parabolic-channel":
            n0, depth, r0, width, rmax = self.pp
            r2 = x1 ** 2 + x2 ** 2
            r02, r2max = r0 ** 2, width ** 2
            d = np.where(r2 < r2max, n0 + depth * r2 / r02, n0 + depth * r2max / r02)
            return np.where(r2 > rmax ** 2, 0.0, d)
        if self.perp == "hollow-channel":
            rmin, rmax, depth = self.pp
            r = np.sqrt(x1 ** 2 + x2 ** 2)
            return np.where((r < rmax) & (r > rmin), depth, 0.0)
        return np.broadcast_to(self.func(x1, x2, np.full_like(x1, float(s))), x1.shape).astype(float)

    def den_lon(self, s):
        """get_den_lon_* (fdist2d_lib.f03:186-415) at propagation distance s;
        the analytic profile puts its s dependence in the transverse one"""
        s = float(s)
        if self.lon in ("uniform", "analytic"):
            return 1.0
        if self.lon == "piecewise-linear":
            sa, fa = self.s, self.fs
            if s <= sa[0] or s > sa[-1]:
                return 0.0
            i = int(np.searchsorted(sa, s, side="left"))   # first sa[i] >= s, i >= 1
            return float(fa[i - 1] + (fa[i] - fa[i - 1]) / (sa[i] - sa[i - 1]) * (s - sa[i - 1]))
        if self.lon == "sine-series":
            v = 0.0
            if self.start < s <= self.start + self.length:
                for i, c in enumerate(self.coefs, start=1):
                    v = v + c * math.sin(math.pi * float(i) * (s - self.start) / self.length)
            return abs(v)
        sa, fa, da = self.s, self.fs, self.dfs
        if s <= sa[0] or s > sa[-1]:
            return 0.0
        i = int(np.searchsorted(sa, s, side="left")) - 1
        t = (s - sa[i]) / (sa[i + 1] - sa[i])
        return float(fa[i] + da[i] * t + (3.0 * (fa[i + 1] - fa[i]) - 2.0 * da[i] - da[i + 1]) * t ** 2
                     + (2.0 * (fa[i] - fa[i + 1]) + da[i] + da[i + 1]) * t ** 3)


def _tdma(lo, d, up, rhs):
    n = len(d)
    c, x = np.zeros(n), np.zeros(n)
    b = np.array(d, dtype=float)
    r = np.array(rhs, dtype=float)
    for i in range(1, n):
        w = lo[i] / b[i - 1]
        b[i] = b[i] - w * up[i - 1]
        r[i] = r[i] - w * r[i - 1]
    x[n - 1] = r[n - 1] / b[n - 1]
    for i in range(n - 2, -1, -1):
        x[i] = (r[i] - up[i] * x[i + 1]) / b[i]
    del c
    return x


def _spline_slopes(s, fs, kind):
    """set_prof_lon_cubic_spline (fdist2d_lib.f03:319-380): the node slopes
    (per unit of the normalised interval) of the cubic spline through fs"""
    n = len(fs)
    if n < 3 and kind != "natural":
        raise QpadError("cubic-spline profiles need at least 3 points")
    diag = np.full(n, 4.0)
    lo = np.ones(n)
    up = np.ones(n)
    rhs = np.zeros(n)
    rhs[1:n - 1] = 3.0 * (fs[2:] - fs[:n - 2])
    if kind == "natural":
        diag[0] = diag[-1] = 2.0
        rhs[0] = 3.0 * (fs[1] - fs[0])
        rhs[-1] = 3.0 * (fs[-1] - fs[-2])
        return _tdma(lo, diag, up, rhs)
    if kind == "clamped":
        d = np.zeros(n)
        d[1:n - 1] = _tdma(lo[1:n - 1], diag[1:n - 1], up[1:n - 1], rhs[1:n - 1])
        return d
    if kind == "not-a-knot":
        diag[0] = diag[-1] = 2.0
        lo[-1] = 4.0
        up[0] = 4.0
        rhs[0] = -5.0 * fs[0] + 4.0 * fs[1] + fs[2]
        rhs[-1] = -5.0 * fs[-3] + 4.0 * fs[-2] + fs[-1]
        return _tdma(lo, diag, up, rhs)
    raise QpadError("Invaild endpoints type of cubic spline profile.")


def plasma_lattice(nr, dr, ppc1, ppc2, num_theta):
    """the injection lattice in inject_fdist2d's order (theta bin j, radial
    cell i, i1, i2; fdist2d_class.f03:317-333): x1, x2 of every point"""
    j, i, i1, i2 = np.meshgrid(np.arange(1, num_theta + 1), np.arange(1, nr + 1), np.arange(1, ppc1 + 1),
                               np.arange(1, ppc2 + 1), indexing="ij")
    rn = (i1 - 0.5) / ppc1 + (i - 1.0)
    theta = ((i2 - 0.5) / ppc2 + j - 1.0) * (2.0 * math.pi / num_theta)
    return (rn * dr * np.cos(theta)).ravel(), (rn * dr * np.sin(theta)).ravel()


# ---------------------------------------------------------------------------
# lattice beam profiles (fdist3d_std_lib.f03) the library does not evaluate
# itself: per direction super-gaussian, parabolic, rational (besides uniform,
# gaussian, piecewise-linear), or one analytic math_func of (x, y, z) with z
# absolute (the box start added back)
# ---------------------------------------------------------------------------
class BeamProfile:
    """beam(i).profile of a lattice ("standard") beam, evaluated on the host
    (fdist3d_std_class.f03:160-300, fdist3d_std_lib.f03:20-363); z0 is the
    box start, which the deck's longitudinal positions are relative to."""

    TABULATED = ("super-gaussian", "parabolic", "rational", "analytic")

    def __init__(self, bm, z0):
        prof = list(_deck_get(bm, "profile"))
        self.kinds = prof
        self.z0 = float(z0)
        self.func = None
        self.pars = []
        if "analytic" in prof:
            self.func = FParser(_deck_get(bm, "math_func"))
        for d, name in enumerate(prof):
            shift = self.z0 if d == 2 else 0.0
            dim = str(d + 1)
            if name == "gaussian":
                self.pars.append((float(bm["gauss_center"][d]) - shift, float(bm["gauss_sigma"][d])))
            elif name == "super-gaussian":
                order = float(bm["super_gauss_order"][d])
                if order < 1.0:
                    raise QpadError("The order of super-Gaussian profile must be >= 1.")
                self.pars.append((float(bm["super_gauss_center"][d]) - shift, float(bm["super_gauss_sigma"][d]),
                                  order))
            elif name == "parabolic":
                self.pars.append((float(bm["parabolic_center"][d]) - shift, float(bm["parabolic_radius"][d])))
            elif name == "rational":
                self.pars.append((float(bm["poly_x0"][d]) - shift,
                                  np.asarray(_deck_get(bm, "poly_numerator" + dim), dtype=float),
                                  np.asarray(_deck_get(bm, "poly_denominator" + dim), dtype=float)))
            elif name == "piecewise-linear":
                xs = np.asarray(_deck_get(bm, "piecewise_x" + dim), dtype=float) - shift
                if np.any(np.diff(xs) <= 0):
                    raise QpadError("Position array must be monotonically increasing!")
                self.pars.append((xs, np.asarray(_deck_get(bm, "piecewise_fx" + dim), dtype=float)))
            elif name in ("uniform", "analytic"):
                self.pars.append(())
            else:
                raise QpadError(f"Invalid beam density profile '{name}'")

    @property
    def tabulated(self):
        return any(k in self.TABULATED for k in self.kinds)

    def _dim(self, d, x):
        kind, p = self.kinds[d], self.pars[d]
        if kind == "uniform":
            return np.ones_like(x)
        if kind == "gaussian":
            return np.exp(-0.5 * ((x - p[0]) / p[1]) ** 2)
        if kind == "super-gaussian":
            return np.exp(-(0.5 * ((x - p[0]) / p[1]) ** 2) ** p[2])
        if kind == "parabolic":
            t = (x - p[0]) ** 2 / (p[1] * p[1])
            return np.where(t < 1.0, 1.0 - t, 0.0)
        if kind == "rational":
            x0, pn, pd = p
            num = np.zeros_like(x)
            den = np.zeros_like(x)
            for i, c in enumerate(pn):
                num = num + c * (x - x0) ** i
            for i, c in enumerate(pd):
                den = den + c * (x - x0) ** i
            return num / den
        xs, fx = p
        out = np.zeros_like(x)
        inside = (x >= xs[0]) & (x <= xs[-1])
        i = np.clip(np.searchsorted(xs, x, side="left"), 1, len(xs) - 1)
        lin = fx[i - 1] + (fx[i] - fx[i - 1]) / (xs[i] - xs[i - 1]) * (x - xs[i - 1])
        return np.where(inside, lin, out)

    def density(self, x1, x2, x3):
        """the density at lattice positions in the beam's geometry
        coordinates (x3 box-relative), without the beam's density factor"""
        with np.errstate(all="ignore"):
            if self.func is not None:
                return np.asarray(self.func(x1, x2, x3 + self.z0), dtype=float) * np.ones_like(x1)
            return self._dim(0, x1) * self._dim(1, x2) * self._dim(2, x3)


def beam_lattice_box(nr, nxi, dr, dxi, cfg):
    """the lattice box that holds every point inside the beam's ranges (the
    library's own choice, qpad_beam_lattice_box): first xi cell k0 (1-based),
    xi cells nk, r cells ni"""
    if cfg.geometry == 1:
        rmax = cfg.range_hi[0]
    else:
        mx = max(abs(cfg.range_lo[0]), abs(cfg.range_hi[0]))
        my = max(abs(cfg.range_lo[1]), abs(cfg.range_hi[1]))
        rmax = math.sqrt(mx * mx + my * my)
    ni = int(min(float(nr), max(math.floor(rmax / dr) + 2.0, 0.0)))
    k0 = math.floor(cfg.range_lo[2] / dxi) - 1.0
    k1 = math.floor(cfg.range_hi[2] / dxi) + 2.0
    kk0 = int(max(1.0, min(k0, nxi + 1.0)))
    kk1 = int(max(0.0, min(k1, float(nxi))))
    nk = kk1 - kk0 + 1 if kk1 >= kk0 else 0
    if ni <= 0:
        nk = 0
    return kk0, nk, ni


def beam_lattice_density(profile, cfg, nr, dr, dxi, box):
    """the density of every point of the lattice box in inject_fdist3d_std's
    order (k > j > i > i3 > i2 > i1, fdist3d_std_class.f03:439-490)"""
    k0, nk, ni = box
    p1, p2, p3, nth = cfg.ppc[0], cfg.ppc[1], cfg.ppc[2], cfg.num_theta
    k, j, i, i3, i2, i1 = np.meshgrid(np.arange(k0, k0 + nk), np.arange(1, nth + 1), np.arange(1, ni + 1),
                                      np.arange(1, p3 + 1), np.arange(1, p2 + 1), np.arange(1, p1 + 1),
                                      indexing="ij")
    zn = (i3 - 0.5) / p3 + (k - 1.0)
    theta = ((i2 - 0.5) / p2 + j - 1.0) * (2.0 * math.pi / nth)
    rn = (i1 - 0.5) / p1 + (i - 1.0)
    if cfg.geometry == 1:
        xa, xb = rn * dr, theta
    else:
        xa, xb = rn * dr * np.cos(theta), rn * dr * np.sin(theta)
    return np.ascontiguousarray(profile.density(xa, xb, zn * dxi).ravel(), dtype=np.float64)
