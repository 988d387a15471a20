"""qpinput.json reader for the GPU path.

Maps QPAD's input deck (docs/source/input_file/*.rst; read by
source/input_class.f03:49-104 and the init routines it feeds) onto the C ABI
structures.  Only the features on the accelerated path are accepted; anything
else raises QpadError, as QPAD's write_err would stop on an invalid deck.
"""
import json
import math

from ._lib import BND_OPEN, BND_ZERO, PUSH2_ROBUST, BeamCfg, Grid, QpadError, SpeciesCfg


def _get(d, key, default=None, required=True):
    if key in d:
        return d[key]
    if required and default is None:
        raise QpadError(f"input deck: missing key '{key}'")
    return default


def load_deck(path):
    with open(path) as f:
        return json.load(f)


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
    if _get(sim, "algorithm", "standard", required=False) != "standard":
        raise QpadError("only algorithm 'standard' is on the GPU path")
    for key in ("nneutrals", "nlasers"):
        if _get(sim, key, 0, required=False) != 0:
            raise QpadError(f"{key} > 0 is not on the GPU path")
    relax = _get(sim, "relax_fac", 1.0e-3 * (dr / 0.02) ** 2, required=False)  # sim_fields_class.f03:137
    g = Grid(nr, nz, _get(sim, "max_mode"), dr, dxi, relax,
             BND_OPEN if bnd_name == "open" else BND_ZERO,
             _get(sim, "iter_max"), _get(sim, "iter_reltol"), _get(sim, "iter_abstol"),
             _get(sim, "dt"))
    nstep3d = int(_get(sim, "time") / _get(sim, "dt"))
    return g, nstep3d, z0


def species_from_deck(sp, seed=0):
    """fdist2d_class.f03:117-272, sim_plasma_class.f03:116-157."""
    prof = _get(sp, "profile")
    if list(prof) != ["uniform", "uniform"]:
        raise QpadError("only uniform plasma profiles are on the GPU path")
    push = _get(sp, "push_type", "robust", required=False)
    if push != "robust":
        raise QpadError('Invalid pusher type! only "robust" is on the GPU path')
    ppc = _get(sp, "ppc")
    uth = _get(sp, "uth", [0.0, 0.0, 0.0], required=False)
    if _get(sp, "random_theta", False, required=False):
        raise QpadError("random_theta is not on the GPU path")
    return SpeciesCfg(int(ppc[0]), int(ppc[1]), int(_get(sp, "num_theta")), float(_get(sp, "q")),
                      float(_get(sp, "m")), float(_get(sp, "density")), (ctypes_d3(uth)),
                      int(bool(_get(sp, "neutralized", True, required=False))),
                      float(_get(sp, "den_min", 1.0e-10, required=False)), int(seed), PUSH2_ROBUST)


def ctypes_d3(v):
    import ctypes as C
    return (C.c_double * 3)(*[float(x) for x in v])


def beam_from_deck(bm, z0, seed=1):
    """fdist3d_rnd_class.f03:104-356 (profile_type 'random', Cartesian Gaussian)."""
    if _get(bm, "profile_type", "standard", required=False) != "random":
        raise QpadError("only profile_type 'random' beams are on the GPU path")
    if _get(bm, "geometry", "cartesian", required=False) != "cartesian":
        raise QpadError("only Cartesian beam geometry is on the GPU path")
    if list(_get(bm, "profile")) != ["gaussian", "gaussian", "gaussian"]:
        raise QpadError("only Gaussian beam profiles are on the GPU path")
    if _get(bm, "push_type", "reduced", required=False) != "reduced":
        raise QpadError('only the "reduced" beam pusher is on the GPU path')
    if _get(bm, "has_spin", False, required=False):
        raise QpadError("spin push is not on the GPU path")
    for key in ("alpha", "perp_offset_x", "perp_offset_y"):
        if key in bm:
            raise QpadError(f"beam key '{key}' is not on the GPU path")
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


def tri_gaussian_charge(n_peak, sigma):
    """total_charge = n_b (2 pi)^(3/2) sigma_x sigma_y sigma_z
    (docs/source/input_file/session_beam.rst:145-150)."""
    return n_peak * (2.0 * math.pi) ** 1.5 * sigma[0] * sigma[1] * sigma[2]
