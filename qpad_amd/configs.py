"""The benchmark workloads as qpinput.json-style decks.

C1: blowout wake, Nr=256, r_max=6, m=0..1, Nxi=256, dxi=0.04, ppc 2x8,
    tri-Gaussian beam sigma_r=0.25, sigma_xi=0.5, n_b/n_p=4, gamma=2e4,
    2^18 macro-particles, 1 3D step.
C2: Nr=4096, r_max=12, m=0..2, Nxi=1024, dxi=0.04, ppc 4x32 (524,288 plasma
    particles per slice), sigma_r=0.5, sigma_xi=1, n_b/n_p=8, 2^22 beam
    particles, 10 3D steps with dt=10.
C3: as C2 with Nr=2048, m=0..4, electrons and mobile ions (m=1836) at ppc
    4x16, both not neutralised, beam offset by 0.1 sigma_r, 20 3D steps.
C4: a parameter scan of 16 independent instances stepped as one batch:
    Nr uniform in {512, 1024, 2048, 4096} (r_max 12), max mode uniform in
    1..4, beam n_b/n_p uniform in [1, 16] and transverse offset uniform in
    [0, 0.2] sigma_r, Nxi=512, dxi=0.04, electrons at ppc 4x16, drawn from a
    recorded seed.
"""
import copy

import numpy as np

from .input import tri_gaussian_charge


def _deck(nr, rmax, max_mode, nxi, dxi, species, beam, dt=10.0, nstep=1, iter_max=10, tol=1e-3):
    zlen = nxi * dxi
    return {
        "simulation": {
            "nodes": [1, 1], "grid": [nr, nxi],
            "box": {"r": [0.0, rmax], "z": [-zlen / 2, zlen / 2]},
            "field_boundary": "open", "max_mode": max_mode, "interpolation": "linear",
            "n0": 1.0e16, "time": dt * nstep, "dt": dt,
            "nbeams": 1, "nspecies": len(species), "nneutrals": 0, "nlasers": 0,
            "iter_max": iter_max, "iter_reltol": tol, "iter_abstol": tol,
            "random_seed": 10, "algorithm": "standard",
        },
        "beam": [beam],
        "species": species,
    }


def _beam(zlen, sigma_r, sigma_xi, n_peak, total_num, offset_x=0.0, gamma=20000.0):
    z_front = -zlen / 2
    zc = z_front + 5.0 * sigma_xi
    sig = [sigma_r, sigma_r, sigma_xi]
    return {
        "profile_type": "random", "geometry": "cartesian",
        "profile": ["gaussian", "gaussian", "gaussian"], "evolution": True,
        "push_type": "reduced", "has_spin": False, "q": -1.0, "m": 1.0, "gamma": gamma,
        "total_num": int(total_num), "total_charge": tri_gaussian_charge(n_peak, sig),
        "quiet_start": False, "gauss_center": [offset_x, 0.0, zc], "gauss_sigma": sig,
        "range1": [offset_x - 5 * sigma_r, offset_x + 5 * sigma_r],
        "range2": [-5 * sigma_r, 5 * sigma_r],
        "range3": [z_front, zc + 5.0 * sigma_xi],
        "uth": [5.0, 5.0, 0.0],
    }


def _electrons(ppc_r, ppc_phi, neutralized=True, push_type="robust"):
    return {"profile": ["uniform", "uniform"], "ppc": [ppc_r, 1], "num_theta": ppc_phi,
            "q": -1.0, "m": 1.0, "density": 1.0, "push_type": push_type,
            "neutralized": neutralized}


def c1(total_num=1 << 18):
    nxi, dxi = 256, 0.04
    return _deck(256, 6.0, 1, nxi, dxi, [_electrons(2, 8)],
                 _beam(nxi * dxi, 0.25, 0.5, 4.0, total_num), nstep=1)


def c2(total_num=1 << 22, nstep=10):
    nxi, dxi = 1024, 0.04
    return _deck(4096, 12.0, 2, nxi, dxi, [_electrons(4, 32)],
                 _beam(nxi * dxi, 0.5, 1.0, 8.0, total_num), nstep=nstep)


def c3(total_num=1 << 22, nstep=20):
    nxi, dxi = 1024, 0.04
    ions = _electrons(4, 16, neutralized=False)
    ions.update({"q": 1.0, "m": 1836.0})
    return _deck(2048, 12.0, 4, nxi, dxi, [_electrons(4, 16, neutralized=False), ions],
                 _beam(nxi * dxi, 0.5, 1.0, 8.0, total_num, offset_x=0.05), nstep=nstep)


def c4(n=16, seed=2024, total_num=1 << 18, nxi=512):
    """The C4 scan: a list of n decks (sigma_r 0.5, sigma_xi 1, beam of
    total_num macro-particles each) and the drawn parameters."""
    rng = np.random.default_rng(seed)
    dxi, sigma_r = 0.04, 0.5
    decks, params = [], []
    for _ in range(n):
        nr = int(rng.choice([512, 1024, 2048, 4096]))
        m = int(rng.integers(1, 5))
        nb = float(rng.uniform(1.0, 16.0))
        off = float(rng.uniform(0.0, 0.2)) * sigma_r
        decks.append(_deck(nr, 12.0, m, nxi, dxi, [_electrons(4, 16)],
                           _beam(nxi * dxi, sigma_r, 1.0, nb, total_num, offset_x=off), nstep=1))
        params.append({"nr": nr, "max_mode": m, "n_b": nb, "offset_x": off})
    return decks, params


def small(nr=64, nxi=48, max_mode=1, ppc=(2, 4), total_num=1 << 14, rmax=4.0, dxi=0.05,
          n_peak=4.0, sigma_r=0.3, sigma_xi=0.4, offset_x=0.0, ions=False, ion_ppc=None,
          push_type="robust", ion_push_type=None):
    """A reduced deck with the same physics for fast parity checks; with
    ions, ion_ppc (default: the electrons' ppc) sets the ions' (r, phi)
    particles per cell, so that the two species can differ in count.
    push_type / ion_push_type (default: push_type): "robust" or "standard"."""
    species = [_electrons(ppc[0], ppc[1], neutralized=not ions, push_type=push_type)]
    if ions:
        ip = ppc if ion_ppc is None else ion_ppc
        ion = _electrons(ip[0], ip[1], neutralized=False, push_type=ion_push_type or push_type)
        ion.update({"q": 1.0, "m": 1836.0})
        species.append(ion)
    return _deck(nr, rmax, max_mode, nxi, dxi, species,
                 _beam(nxi * dxi, sigma_r, sigma_xi, n_peak, total_num, offset_x=offset_x), nstep=1)


def ionization(nr=48, nxi=40, max_mode=1, element=3, ion_max=3, ppc=(2, 2), num_theta=8, plasma=False,
               n_peak=6.0, sigma_r=0.25, total_num=1 << 12, n0=1.0e17, push_type="robust", offset_x=0.0):
    """A reduced ionisation deck (QPAD's input_file/ionization example in
    small): a uniform neutral gas of `element` followed up to ion_max
    levels behind a Gaussian driver, with (plasma) or without a plasma
    species; n0 sets the field unit of the ADK rates."""
    species = [_electrons(2, 4, push_type="robust")] if plasma else []
    d = _deck(nr, 4.0, max_mode, nxi, 0.05, species,
              _beam(nxi * 0.05, sigma_r, 0.4, n_peak, total_num, offset_x=offset_x), nstep=1)
    d["simulation"]["n0"] = n0
    d["simulation"]["nneutrals"] = 1
    d["neutrals"] = [{"profile": ["uniform", "uniform"], "ppc": list(ppc), "num_theta": num_theta,
                      "q": -1.0, "m": 1.0, "density": 1.0, "element": element, "ion_max": ion_max,
                      "push_type": push_type}]
    return d


def subcycling(nr=32, nxi=40, max_mode=1, n_peak=40.0, sigma_xi=0.2, exp_fac_max=1.1, exp_fac_clamped=1.5,
               dt_2d_min=0.01, density=1.0, push_type="robust", neutral=None, offset_x=0.0, nstep=1):
    """A reduced deck of QPAD's sub-cycling algorithm (simulation.algorithm
    "subcycling", proj_subcyc): a strong narrow driver (n_b/n_p = n_peak,
    sigma_r 0.25, sigma_xi) whose blowout drives gamma / (gamma - p_z) of the plasma
    electrons above expansion_fac_max behind it, so that those slices are
    pushed in several 2D steps; the clamp caps the factor at
    expansion_fac_clamped.  density scales the plasma (and the gas); neutral:
    (element, ion_max) of a uniform gas beside the plasma, n0 1e18."""
    d = _deck(nr, 4.0, max_mode, nxi, 0.05, [_electrons(2, 4, push_type=push_type)],
              _beam(nxi * 0.05, 0.25, sigma_xi, n_peak, 1 << 12, offset_x=offset_x), nstep=nstep)
    d["species"][0]["density"] = density
    sim = d["simulation"]
    sim.update({"algorithm": "subcycling", "expansion_fac_max": exp_fac_max,
                "expansion_fac_clamped": exp_fac_clamped, "dt_2d_min": dt_2d_min})
    if neutral is not None:
        sim["n0"] = 1.0e18
        sim["nneutrals"] = 1
        d["neutrals"] = [{"profile": ["uniform", "uniform"], "ppc": [2, 2], "num_theta": 8, "q": -1.0, "m": 1.0,
                          "density": density, "element": neutral[0], "ion_max": neutral[1], "push_type": push_type}]
    return d


def profiled(nr=48, nxi=32, max_mode=1, profile=None, den_min=1e-10, uth=(0.0, 0.0, 0.0), n_peak=6.0, nstep=2,
             dt=10.0, push_type="robust"):
    """A reduced deck with a non-uniform plasma (species(i).profile and its
    parameters, fdist2d_lib.f03): by default a parabolic channel cut at
    r = 3 whose density rises along the propagation distance (piecewise
    linear from 0.5 at s = 0 to 1.5 at s = 20); `profile` replaces those
    keys."""
    sp = _electrons(2, 4, push_type=push_type)
    sp.update(profile or {"profile": ["parabolic-channel", "piecewise-linear"], "channel_n0": 0.6,
                          "channel_depth": 0.4, "channel_r0": 1.0, "channel_width": 1.5, "channel_rmax": 3.0,
                          "piecewise_s": [-1.0, 20.0], "piecewise_fs": [0.5, 1.5]})
    sp["den_min"] = den_min
    sp["uth"] = list(uth)
    d = _deck(nr, 4.0, max_mode, nxi, 0.05, [sp], _beam(nxi * 0.05, 0.3, 0.3, n_peak, 1 << 12), dt=dt,
              nstep=nstep)
    return d


def analytic_profile(nr=48, nxi=40, max_mode=1):
    """QPAD's analytic-profile example (input_file/analytic_density_profiles/
    qpinput.json) in small: the plasma density
    z^2/(z + 10)^2 exp(-(x^2 + y^2)/50) of the propagation distance z and the
    transverse position, behind a Gaussian driver (the example's beam is an
    analytic lattice beam, which stays off this path)."""
    return profiled(nr=nr, nxi=nxi, max_mode=max_mode, dt=2.0, profile={
        "profile": ["analytic", "analytic"], "math_func": "z^2/(z + 10.0)^2 * exp(-(x^2 + y^2)/50.0)"})


def analytic_example():
    """QPAD's analytic-profile example, restated as data
    (input_file/analytic_density_profiles/qpinput.json, diagnostics and
    restart settings left out): Nr 1024, r_max 60, Nxi 1024 over [-2, 12],
    modes 0..1, a lattice beam of density exp(-z^2/0.5) exp(-(x^2+y^2)/8)
    (ppc 2x2x2, 8 theta sectors, quiet start) and electrons of density
    z^2/(z+10)^2 exp(-(x^2+y^2)/50) (ppc 2x2, 8 theta sectors), 100 3D
    steps of dt 2 (run here for as many as asked)."""
    return {
        "simulation": {
            "nodes": [1, 1], "grid": [1024, 1024], "box": {"r": [0.0, 60.0], "z": [-2.0, 12.0]},
            "field_boundary": "open", "max_mode": 1, "interpolation": "linear", "n0": 1.0e16,
            "time": 200.1, "dt": 2.0, "nbeams": 1, "nspecies": 1, "nneutrals": 0, "nlasers": 0,
            "iter_max": 10, "iter_reltol": 1.0e-2, "iter_abstol": 1.0e-3, "random_seed": 10,
            "algorithm": "standard",
        },
        "beam": [{
            "npmax": 1000000, "ppc": [2, 2, 2], "num_theta": 8, "evolution": True, "quiet_start": True,
            "geometry": "cartesian", "math_func": "exp(-z^2/0.5) * exp(-(x^2 + y^2)/8.0)",
            "profile": ["analytic", "analytic", "analytic"], "q": -1.0, "m": 1.0, "density": 1.0,
            "gamma": 20000.0, "uth": [10.0, 10.0, 0.0], "range1": [-10, 10], "range2": [-10, 10],
            "range3": [-2, 2],
        }],
        "species": [{
            "profile": ["analytic", "analytic"], "math_func": "z^2/(z + 10.0)^2 * exp(-(x^2 + y^2)/50.0)",
            "ppc": [2, 2], "num_theta": 8, "q": -1.0, "m": 1.0, "density": 1.0, "push_type": "robust",
        }],
    }


def ionization_example():
    """QPAD's own ionisation example, restated as data
    (input_file/ionization/qpinput.json, diagnostics left out): Nr 500,
    r_max 5, Nxi 1000 over [0, 10], modes 0..1, n0 1e17, no plasma species,
    a lithium gas (ion_max 3, ppc 8 x 8, 32 theta bins, robust pusher), a
    lattice tri-Gaussian beam (n_b/n_p = 4, sigma (0.25, 0.25, 0.5), quiet
    start), one 3D step of dt 10."""
    return {
        "simulation": {
            "nodes": [1, 2], "grid": [500, 1000], "box": {"r": [0.0, 5.0], "z": [0.0, 10.0]},
            "field_boundary": "open", "max_mode": 1, "interpolation": "linear", "n0": 1.0e17,
            "time": 10.1, "dt": 10.0, "nbeams": 1, "nspecies": 0, "nneutrals": 1, "nlasers": 0,
            "iter_max": 10, "iter_reltol": 1.0e-3, "iter_abstol": 1.0e-3, "random_seed": 10,
            "algorithm": "standard",
        },
        "beam": [{
            "geometry": "cartesian", "profile": ["gaussian", "gaussian", "gaussian"], "evolution": True,
            "push_type": "reduced", "has_spin": False, "ppc": [1, 1, 1], "num_theta": 16,
            "npmax": 2000000, "q": -1.0, "m": 1.0, "gamma": 20000, "density": 4.0, "quiet_start": True,
            "gauss_center": [0.0, 0.0, 2.5], "gauss_sigma": [0.25, 0.25, 0.5],
            "range1": [-1.25, 1.25], "range2": [-1.25, 1.25], "range3": [0.0, 5.0],
            "uth": [5.0, 5.0, 0.0], "den_min": 1e-10,
        }],
        "species": [],
        "neutrals": [{
            "profile": ["uniform", "uniform"], "ppc": [8, 8], "num_theta": 32, "q": -1.0, "m": 1.0,
            "density": 1.0, "element": 3, "ion_max": 3, "push_type": "robust",
        }],
    }


def tri_gaussian():
    """QPAD's own blowout-regime example, restated as data
    (input_file/blowout_regime/qpinput_tri-gaussian.json, diagnostics and
    restart settings left out): Nr 250, r_max 5, Nxi 500 over [-5, 5], modes
    0..1, electrons at ppc 2 x 2 x 16, a lattice ("standard") tri-Gaussian
    beam with quiet start, n_b/n_p = 4, sigma (0.25, 0.25, 0.5), one 3D step
    of dt 10."""
    return {
        "simulation": {
            "nodes": [1, 1], "grid": [250, 500], "box": {"r": [0.0, 5.0], "z": [-5.0, 5.0]},
            "field_boundary": "open", "max_mode": 1, "interpolation": "linear", "n0": 1.0e16,
            "time": 10.1, "dt": 10.0, "nbeams": 1, "nspecies": 1, "nneutrals": 0, "nlasers": 0,
            "iter_max": 10, "iter_reltol": 1.0e-3, "iter_abstol": 1.0e-3, "random_seed": 10,
            "algorithm": "standard",
        },
        "beam": [{
            "geometry": "cartesian", "profile": ["gaussian", "gaussian", "gaussian"], "evolution": True,
            "push_type": "reduced", "has_spin": False, "ppc": [2, 2, 2], "num_theta": 16,
            "npmax": 1000000, "q": -1.0, "m": 1.0, "gamma": 20000, "density": 4.0, "quiet_start": True,
            "gauss_center": [0.0, 0.0, -2.5], "gauss_sigma": [0.25, 0.25, 0.5],
            "range1": [-1.25, 1.25], "range2": [-1.25, 1.25], "range3": [-5.0, 0.0],
            "uth": [5.0, 5.0, 0.0], "den_min": 1e-10,
        }],
        "species": [{
            "profile": ["uniform", "uniform"], "ppc": [2, 2], "num_theta": 16, "q": -1.0, "m": 1.0,
            "density": 1.0, "push_type": "robust",
        }],
    }


def with_overrides(deck, **sim):
    d = copy.deepcopy(deck)
    d["simulation"].update(sim)
    return d
