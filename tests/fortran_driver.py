"""Runs qpad_amd/fortran/qpad_slice_driver (the Fortran program that drives a
3D step through qpad_hip_mod in QPAD's order) on a deck and reads back what
it wrote."""
import os
import subprocess

import numpy as np

from qpad_amd import _lib as L
from qpad_amd.input import beam_from_deck, grid_from_deck, species_from_deck

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "qpad_amd", "fortran", "qpad_slice_driver")
GRID_KEYS = ("nr", "nxi", "max_mode", "dr", "dxi", "relax_fac", "bnd", "iter_max", "iter_reltol", "iter_abstol",
             "dt")
DUMP = ("psi", "e", "b", "b_spe", "cu", "q_spe")


def _fmt(v):
    return repr(float(v)) if isinstance(v, float) else str(int(v))


def write_parameters(path, deck):
    """the list-directed parameter file of the driver, from a deck as
    qpad_amd.Simulation reads it (same seeds)"""
    g, _, z0 = grid_from_deck(deck)
    sim = deck["simulation"]
    seed = int(sim.get("random_seed", 10))
    species = [species_from_deck(sp, seed=seed + 1000 * k)
               for k, sp in enumerate(deck["species"][:sim.get("nspecies", len(deck["species"]))])]
    beam = beam_from_deck(deck["beam"][0], z0, seed=seed + 7)
    assert isinstance(beam, L.BeamCfg), "the driver takes one random Gaussian beam"
    lines = [" ".join(_fmt(getattr(g, k)) for k in GRID_KEYS), str(len(species))]
    for c in species:
        lines.append(" ".join([_fmt(c.ppc1), _fmt(c.ppc2), _fmt(c.num_theta), _fmt(c.qm), _fmt(c.mass),
                               _fmt(c.density)] + [_fmt(float(x)) for x in c.uth]
                              + [_fmt(c.neutralized), _fmt(c.den_min), _fmt(c.seed), _fmt(c.push_type)]))
    b = beam
    vals = [_fmt(b.total_num), _fmt(b.total_charge), _fmt(b.qm), _fmt(b.mass), _fmt(b.gamma)]
    for arr in (b.center, b.sigma, b.range_lo, b.range_hi, b.uth):
        vals += [_fmt(float(x)) for x in arr]
    vals += [_fmt(b.quiet), _fmt(b.evol), _fmt(b.seed)]
    lines.append(" ".join(vals))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return g


def run(deck, workdir, timeout=120):
    """(grid as the program read it, iterations per slice, fields, species-0
    particles); raises CalledProcessError with the program's message"""
    par = os.path.join(workdir, "driver.par")
    out = os.path.join(workdir, "driver.out")
    write_parameters(par, deck)
    subprocess.run([DRIVER, par, out], check=True, capture_output=True, text=True, timeout=timeout)
    raw = open(out, "rb").read()
    i32, f64 = np.dtype("<i4"), np.dtype("<f8")
    off = 0

    def take(dt, n):
        nonlocal off
        a = np.frombuffer(raw, dtype=dt, count=n, offset=off)
        off += a.nbytes
        return a

    g = {}
    for k in GRID_KEYS:
        g[k] = take(i32, 1)[0] if k in ("nr", "nxi", "max_mode", "bnd", "iter_max") else take(f64, 1)[0]
    nr, nxi, nm = int(g["nr"]), int(g["nxi"]), 2 * int(g["max_mode"]) + 1
    iters = take(i32, nxi).copy()
    fields = {}
    for name in DUMP:
        nc = L.NCOMP[L.FIELD_IDS[name]]
        fields[name] = take(f64, nc * nm * (nr + 2)).reshape(nc, nm, nr + 2).copy()
    npp = int(take(np.dtype("<i8"), 1)[0])
    parts = {k: take(f64, npp).copy() for k in ("x1", "x2", "p1", "p2", "p3")}
    parts["id"] = take(i32, npp).copy()
    assert off == len(raw)
    return g, iters, fields, parts
