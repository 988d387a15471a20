"""qpinput.json reader for the GPU path.

Maps QPAD's input deck (docs/source/input_file/*.rst; read by
source/input_class.f03:49-104 and the init routines it feeds) onto the C ABI
structures.  Only the features on the accelerated path are accepted; anything
else raises QpadError, as QPAD's write_err would stop on an invalid deck.
"""
import dataclasses
import json
import math

import numpy as np

from ._lib import (BND_OPEN, BND_ZERO, GEOM_CARTESIAN, GEOM_CYLINDRICAL, OFFSET_MAX, PROF_GAUSSIAN,
                   PROF_PIECEWISE, PROF_UNIFORM, PUSH2_ROBUST, PUSH2_ROBUST_PGC, PUSH2_STD, PUSH2_STD_PGC,
                   PW_MAX, BeamCfg, BeamStdCfg, Grid, NeutralCfg, QpadError, SpeciesCfg)


def _get(d, key, default=None, required=True):
    if key in d:
        return d[key]
    if required and default is None:
        raise QpadError(f"input deck: missing key '{key}'")
    return default


def strip_comments(text):
    """QPAD reads its deck with json-fortran and comment_char '!'
    (input_class.f03:63): drop '!' to the end of a line outside strings."""
    out = []
    for line in text.splitlines():
        in_str, esc = False, False
        for i, ch in enumerate(line):
            if esc:
                esc = False
            elif ch == "\\":
                esc = in_str
            elif ch == '"':
                in_str = not in_str
            elif ch == "!" and not in_str:
                line = line[:i]
                break
        out.append(line)
    return "\n".join(out)


def load_deck(path):
    with open(path) as f:
        return json.loads(strip_comments(f.read()))


def grid_from_deck(deck):
    """options_class.f03:44-117 and simulation_class.f03:141-172."""
    sim = deck["simulation"]
    nr, nz = _get(sim, "grid")
    box = _get(sim, "box")
    r0, r1 = box["r"]
    z0, z1 = box["z"]
    dr = (r1 - r0) / nr
    dxi = (z1 - z0) / nz
    bnd_name = _get(sim, "field_boundary", "open", required=False)
    if bnd_name not in ("open", "zero"):
        raise QpadError('Invalid field boundary type! Only "open" and "zero" are supported currently.')
    if _get(sim, "interpolation", "linear", required=False) != "linear":
        raise QpadError('Invalid interpolation type! Only "linear" are supported currently.')
    if _get(sim, "algorithm", "standard", required=False) not in ("standard", "subcycling"):
        raise QpadError("only algorithms 'standard' and 'subcycling' are on the GPU path")
    if _get(sim, "nlasers", 0, required=False) != 0:
        raise QpadError("nlasers > 0: the laser solver is not on the GPU path; give the envelope to "
                        "Simulation.set_laser (the PGC pushers and the chi deposit are)")
    relax = _get(sim, "relax_fac", 1.0e-3 * (dr / 0.02) ** 2, required=False)  # sim_fields_class.f03:137
    g = Grid(nr, nz, _get(sim, "max_mode"), dr, dxi, relax,
             BND_OPEN if bnd_name == "open" else BND_ZERO,
             _get(sim, "iter_max"), _get(sim, "iter_reltol"), _get(sim, "iter_abstol"),
             _get(sim, "dt"))
    nstep3d = int(_get(sim, "time") / _get(sim, "dt"))
    return g, nstep3d, z0


def subcycling_from_deck(deck):
    """simulation.algorithm "subcycling" (options_class.f03:55-69,
    proj_subcyc/simulation_subcyc_class.f03:84-86): (expansion_fac_max,
    expansion_fac_clamped, dt_2d_min); None for the standard algorithm."""
    sim = deck["simulation"]
    if _get(sim, "algorithm", "standard", required=False) != "subcycling":
        return None
    return (float(_get(sim, "expansion_fac_max")), float(_get(sim, "expansion_fac_clamped")),
            float(_get(sim, "dt_2d_min")))


# species(i).push_type (sim_plasma_class.f03:123-137)
_PUSHERS = {"standard": PUSH2_STD, "robust": PUSH2_ROBUST, "standard_pgc": PUSH2_STD_PGC,
            "robust_pgc": PUSH2_ROBUST_PGC}


# neutrals(i).element: atomic numbers with ADK coefficients
# (neutral_class.f03:23-299; param.f03:135-148)
NEUTRAL_ELEMENTS = {1: 1, 2: 2, 3: 3, 6: 6, 7: 7, 8: 8, 10: 10, 18: 18}


def neutral_from_deck(nt, n0):
    """neutrals(i) section (sim_plasma_class.f03:148-176, fdist2d for the
    'neutrals' section): q, m, element, ion_max, push_type ("standard" or
    "robust"), ppc, num_theta, density; n0 of the simulation section."""
    prof = _get(nt, "profile")
    if list(prof) != ["uniform", "uniform"]:
        raise QpadError("only uniform neutral profiles are on the GPU path")
    _no_smoothing(nt, "neutrals")
    push = _get(nt, "push_type")
    if push not in ("standard", "robust"):
        raise QpadError('Invalid pusher type! Only "std", "robust" are supported currently.')
    elem = int(_get(nt, "element"))
    if elem not in NEUTRAL_ELEMENTS:
        raise QpadError(f"neutral element {elem}: the GPU path has the ADK coefficients of "
                        f"{sorted(NEUTRAL_ELEMENTS)} (atomic numbers)")
    ion_max = int(_get(nt, "ion_max"))
    if not 1 <= ion_max <= NEUTRAL_ELEMENTS[elem]:
        raise QpadError(f"neutral ion_max {ion_max}: element {elem} has {NEUTRAL_ELEMENTS[elem]} levels")
    if _get(nt, "random_theta", False, required=False):
        raise QpadError("random_theta is not on the GPU path")
    ppc = _get(nt, "ppc")
    return NeutralCfg(int(ppc[0]), int(ppc[1]), int(_get(nt, "num_theta")), float(_get(nt, "q")),
                      float(_get(nt, "m")), float(_get(nt, "density")),
                      float(_get(nt, "den_min", 1.0e-10, required=False)), elem, ion_max,
                      PUSH2_STD if push == "standard" else PUSH2_ROBUST, float(n0))


def profile_from_deck(sp):
    """species(i).profile and its parameters (fdist2d_class.f03:140-215): a
    PlasmaProfile, or None for a uniform plasma"""
    from .profiles import PlasmaProfile
    prof = PlasmaProfile(sp)
    return None if prof.uniform else prof


def _no_smoothing(sec, what):
    # species(i) / neutrals(i).smooth_order (sim_plasma_class.f03:140-141,
    # 169-170): the binomial smoothing of the species' deposits
    # (field_src_class.f03:102-265) is not on this path
    if int(_get(sec, "smooth_order", 0, required=False)) > 0:
        raise QpadError(f"{what}.smooth_order > 0: the smoothing of the deposits is not on the GPU path")


def species_from_deck(sp, seed=0):
    """fdist2d_class.f03:117-272, sim_plasma_class.f03:116-157 (the density
    profile: profile_from_deck)."""
    profile_from_deck(sp)   # checks the profile section
    _no_smoothing(sp, "species")
    push = _get(sp, "push_type")
    if push not in _PUSHERS:
        raise QpadError('Invalid pusher type! Only "standard", "robust", "standard_pgc" and "robust_pgc" '
                        'are supported currently.')
    ppc = _get(sp, "ppc")
    uth = _get(sp, "uth", [0.0, 0.0, 0.0], required=False)
    if _get(sp, "random_theta", False, required=False):
        raise QpadError("random_theta is not on the GPU path")
    return SpeciesCfg(int(ppc[0]), int(ppc[1]), int(_get(sp, "num_theta")), float(_get(sp, "q")),
                      float(_get(sp, "m")), float(_get(sp, "density")), (ctypes_d3(uth)),
                      int(bool(_get(sp, "neutralized", True, required=False))),
                      float(_get(sp, "den_min", 1.0e-10, required=False)), int(seed), _PUSHERS[push])


def ctypes_d3(v):
    import ctypes as C
    return (C.c_double * 3)(*[float(x) for x in v])


def beam_from_deck(bm, z0, seed=1):
    """beam(i) section: profile_type "standard" (the default: the lattice of
    fdist3d_std_class.f03:150-436), "random" (fdist3d_rnd_class.f03:
    104-356) or "file" (fdist3d_file_class.f03:52-343).  xi positions are
    made relative to the box start, as QPAD stores them."""
    kind = _get(bm, "profile_type", "standard", required=False)
    if _get(bm, "push_type", "reduced", required=False) != "reduced":
        raise QpadError('only the "reduced" beam pusher is on the GPU path')
    if _get(bm, "has_spin", False, required=False):
        raise QpadError("spin push is not on the GPU path")
    if kind == "standard":
        return _lattice_beam(bm, z0, seed)
    if kind == "file":
        return _file_beam(bm, z0)
    if kind != "random":
        raise QpadError(f"Invalid beam profile type '{kind}'")
    if _get(bm, "geometry", "cartesian", required=False) != "cartesian":
        raise QpadError("only Cartesian random beams are on the GPU path")
    if list(_get(bm, "profile")) != ["gaussian", "gaussian", "gaussian"]:
        raise QpadError("only Gaussian random beam profiles are on the GPU path")
    for key in ("alpha", "perp_offset_x", "perp_offset_y"):
        if key in bm:
            raise QpadError(f"beam key '{key}' is not on the GPU path for random beams")
    ctr = list(_get(bm, "gauss_center"))
    ctr[2] = ctr[2] - z0
    lo = [bm["range1"][0], bm["range2"][0], bm["range3"][0] - z0]
    hi = [bm["range1"][1], bm["range2"][1], bm["range3"][1] - z0]
    return BeamCfg(int(_get(bm, "total_num")), float(_get(bm, "total_charge")), float(_get(bm, "q")),
                   float(_get(bm, "m")), float(_get(bm, "gamma")), ctypes_d3(ctr),
                   ctypes_d3(_get(bm, "gauss_sigma")), ctypes_d3(lo), ctypes_d3(hi),
                   ctypes_d3(_get(bm, "uth", [0.0, 0.0, 0.0], required=False)),
                   int(bool(_get(bm, "quiet_start", False, required=False))),
                   int(bool(_get(bm, "evolution", True, required=False))), int(seed))


@dataclasses.dataclass
class FileBeam:
    """A beam read from a particle file (init_fdist3d_file,
    fdist3d_file_class.f03:52-130): datasets x1..x3, p1..p3, q of the file's
    root group; beam_center[2] relative to the box start."""
    filename: str
    beam_center: tuple
    file_center: tuple
    length_conv: float
    charge_conv: float
    qm: float
    mass: float
    evol: bool


def _file_beam(bm, z0):
    if _get(bm, "has_spin", False, required=False):
        raise QpadError("spin is not on the GPU path")
    bc = [float(v) for v in _get(bm, "beam_center")]
    bc[2] -= z0
    return FileBeam(str(_get(bm, "filename")), tuple(bc), tuple(float(v) for v in _get(bm, "file_center")),
                    float(_get(bm, "length_conv_fac", 1.0, required=False)),
                    float(_get(bm, "charge_conv_fac", 1.0, required=False)), float(_get(bm, "q")),
                    float(_get(bm, "m")), bool(_get(bm, "evolution", True, required=False)))


def place_file_beam(raw, fb, nr, dr, nxi, dxi):
    """Particles of the file in box coordinates (inject_fdist3d_file,
    fdist3d_file_class.f03:300-330): x_perp = center + (x - file_center) *
    length_conv, xi = center_3 - (x3 - file_center_3) * length_conv (the
    file's x3 grows toward the head), q * charge_conv; only particles with
    0 <= r < nr*dr and 0 <= xi < nxi*dxi are kept, in file order."""
    c, f, k = fb.beam_center, fb.file_center, fb.length_conv
    x1 = c[0] + (np.asarray(raw["x1"], dtype=np.float64).ravel() - f[0]) * k
    x2 = c[1] + (np.asarray(raw["x2"], dtype=np.float64).ravel() - f[1]) * k
    x3 = c[2] - (np.asarray(raw["x3"], dtype=np.float64).ravel() - f[2]) * k
    r = np.sqrt(x1 * x1 + x2 * x2)
    keep = (r >= 0.0) & (r < nr * dr) & (x3 >= 0.0) & (x3 < nxi * dxi)
    out = {"x1": x1[keep], "x2": x2[keep], "x3": x3[keep]}
    for name in ("p1", "p2", "p3"):
        out[name] = np.asarray(raw[name], dtype=np.float64).ravel()[keep]
    out["q"] = np.asarray(raw["q"], dtype=np.float64).ravel()[keep] * fb.charge_conv
    return out


_PROFILES = {"uniform": PROF_UNIFORM, "gaussian": PROF_GAUSSIAN, "piecewise-linear": PROF_PIECEWISE}


def _lattice_beam(bm, z0, seed):
    """fdist3d_std_class.f03:150-436 (settings read by init_fdist3d_std) and
    fdist3d_std_lib.f03 (profile parameters)."""
    c = BeamStdCfg()
    geom = _get(bm, "geometry", "cartesian", required=False)
    if geom not in ("cartesian", "cylindrical"):
        raise QpadError(f"Invalid beam geometry '{geom}'")
    c.geometry = GEOM_CARTESIAN if geom == "cartesian" else GEOM_CYLINDRICAL
    prof = list(_get(bm, "profile"))
    from .profiles import BeamProfile
    tab = BeamProfile(bm, z0)     # checks every direction's profile and its parameters
    for d, name in enumerate(prof):
        if name not in _PROFILES:
            continue              # evaluated on the host (BeamProfile)
        c.profile[d] = _PROFILES[name]
        shift = z0 if d == 2 else 0.0
        if name == "gaussian":
            c.gauss_center[d] = float(bm["gauss_center"][d]) - shift
            c.gauss_sigma[d] = float(bm["gauss_sigma"][d])
        elif name == "piecewise-linear":
            xs = [float(v) - shift for v in _get(bm, f"piecewise_x{d + 1}")]
            fx = [float(v) for v in _get(bm, f"piecewise_fx{d + 1}")]
            if len(xs) != len(fx) or not 2 <= len(xs) <= PW_MAX:
                raise QpadError(f"piecewise_x{d + 1} / piecewise_fx{d + 1}: 2 .. {PW_MAX} points of equal count")
            c.pw_n[d] = len(xs)
            for i, (x, f) in enumerate(zip(xs, fx)):
                c.pw_x[d][i] = x
                c.pw_fx[d][i] = f
    ppc = _get(bm, "ppc")
    for d in range(3):
        c.ppc[d] = int(ppc[d])
    c.num_theta = int(_get(bm, "num_theta"))
    c.qm = float(_get(bm, "q"))
    c.mass = float(_get(bm, "m"))
    c.density = float(_get(bm, "density"))
    c.gamma = float(_get(bm, "gamma"))
    for d in range(3):
        r = _get(bm, f"range{d + 1}")
        shift = z0 if d == 2 else 0.0
        c.range_lo[d] = float(r[0]) - shift
        c.range_hi[d] = float(r[1]) - shift
    c.evol = int(bool(_get(bm, "evolution", True, required=False)))
    c.quiet = int(bool(_get(bm, "quiet_start", False, required=False)))
    c.den_min = float(_get(bm, "den_min", 1.0e-10, required=False))
    uth = _get(bm, "uth", [0.0, 0.0, 0.0], required=False)
    for d in range(3):
        c.uth[d] = float(uth[d])
    if "alpha" in bm:
        # Twiss parameters (fdist3d_std_class.f03:367-386)
        for d in range(2):
            if prof[d] != "gaussian" or geom != "cartesian":
                raise QpadError("Twiss parameter alpha is only available for Gaussian profile in the "
                                "Cartesian geometry.")
            c.alpha[d] = float(bm["alpha"][d])
            c.twiss_ctr[d] = float(bm["gauss_center"][d])
            c.twiss_beta[d] = c.gamma * float(bm["gauss_sigma"][d]) / float(uth[d])
    for d, key in enumerate(("perp_offset_x", "perp_offset_y")):
        if key in bm:
            v = [float(x) for x in bm[key]]
            if not 2 <= len(v) <= OFFSET_MAX + 1:
                raise QpadError(f"{key}: origin and 1 .. {OFFSET_MAX} coefficients")
            c.offset_xi0[d] = v[0] - z0
            c.offset_n[d] = len(v) - 1
            for k, x in enumerate(v[1:]):
                c.offset_c[d][k] = x
    c.seed = int(seed)
    return TabulatedBeam(c, tab) if tab.tabulated else c


class TabulatedBeam:
    """A lattice beam whose density the host evaluates (BeamProfile): cfg
    holds the lattice settings, profile the density"""

    def __init__(self, cfg, profile):
        self.cfg = cfg
        self.profile = profile


def tri_gaussian_charge(n_peak, sigma):
    """total_charge = n_b (2 pi)^(3/2) sigma_x sigma_y sigma_z
    (docs/source/input_file/session_beam.rst:145-150)."""
    return n_peak * (2.0 * math.pi) ** 1.5 * sigma[0] * sigma[1] * sigma[2]
