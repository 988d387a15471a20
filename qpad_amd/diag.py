"""QPAD's HDF5 diagnostics from the GPU state (libqpad_diag.so, include/qpad_diag.h).

Diagnostics(sim) reads the deck's diagnostic sections as QPAD's sim_diag
does (diagnostics_class.f03: field.diag :396-606, species(i).diag and
beam(i).diag :180-330; every entry with ndump > 0) and writes the files due
at a 3D step; Simulation.run(diag_dir=...) calls it where QPAD's run loop
does, after the beam push and before the species are renewed
(simulation_class.f03:485-501).
"""
import ctypes as C
import os

import numpy as np

from . import _lib as L

DIAG_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libqpad_diag.so")

FIELD_DIAGS = ("er_cyl_m", "ephi_cyl_m", "ez_cyl_m", "br_cyl_m", "bphi_cyl_m", "bz_cyl_m", "spec_er_cyl_m",
               "spec_ephi_cyl_m", "spec_br_cyl_m", "spec_bphi_cyl_m", "spec_bz_cyl_m", "charge_cyl_m",
               "jr_cyl_m", "jphi_cyl_m", "jz_cyl_m", "psi_cyl_m")


class Box(C.Structure):
    _fields_ = [("rmin", C.c_double), ("rmax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double),
                ("dt", C.c_double)]


_dlib = None


def load():
    """ctypes handle on libqpad_diag.so (raises QpadError when it is not built)."""
    global _dlib
    if _dlib is None:
        L.load()
        if not os.path.exists(DIAG_LIB):
            raise L.QpadError(f"{DIAG_LIB} is not built (python -m qpad_amd.build; needs HDF5)")
        d = C.CDLL(DIAG_LIB)
        d.qpad_diag_last_error.restype = C.c_char_p
        d.qpad_diag_write_grid.argtypes = [C.c_char_p, C.c_char_p, L._dp, C.c_int32, C.c_int32, C.c_int32,
                                           C.POINTER(Box), C.c_char_p, C.c_char_p]
        d.qpad_diag_field.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int32, C.POINTER(Box)]
        for name in ("qpad_diag_species_charge", "qpad_diag_beam_charge"):
            getattr(d, name).argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32, C.POINTER(Box)]
        for name in ("qpad_diag_species_raw", "qpad_diag_beam_raw"):
            getattr(d, name).argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_char_p, C.c_int32, C.POINTER(Box)]
        for name in ("qpad_diag_neutral_charge", "qpad_diag_neutral_ion"):
            getattr(d, name).argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_char_p, C.c_int32, C.POINTER(Box)]
        d.qpad_diag_neutral_raw.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_char_p, C.c_int32,
                                            C.POINTER(Box)]
        d.qpad_diag_read_beam_file.argtypes = [C.c_char_p, C.POINTER(C.c_int64)] + [L._dp] * 7
        _dlib = d
    return _dlib


def check(rc):
    if rc != 0:
        msg = load().qpad_diag_last_error()
        raise L.QpadError(f"libqpad_diag error {rc}: {msg.decode() if msg else ''}")


def box_of(deck):
    sim = deck["simulation"]
    r, z = sim["box"]["r"], sim["box"]["z"]
    return Box(float(r[0]), float(r[1]), float(z[0]), float(z[1]), float(sim["dt"]))


def write_grid(path, dataname, data, iteration, box, units="", label=""):
    """one QPAD grid file of a (nxi, nr) array"""
    a = np.ascontiguousarray(data, dtype=np.float64)
    nxi, nr = a.shape
    check(load().qpad_diag_write_grid(path.encode(), dataname.encode(), L.dptr(a), nxi, nr, int(iteration),
                                      C.byref(box), units.encode(), label.encode()))


BEAM_FILE_KEYS = ("x1", "x2", "x3", "p1", "p2", "p3", "q")


def read_beam_file(path):
    """the datasets x1..x3, p1..p3, q of a "file" beam's particle file
    (fdist3d_file_class.f03:200-290) as float64 arrays"""
    d, n = load(), C.c_int64(0)
    check(d.qpad_diag_read_beam_file(path.encode(), C.byref(n), *([None] * 7)))
    out = {k: np.zeros(n.value) for k in BEAM_FILE_KEYS}
    check(d.qpad_diag_read_beam_file(path.encode(), C.byref(n), *[L.dptr(out[k]) for k in BEAM_FILE_KEYS]))
    return out


def _entries(section):
    """(name, ndump, psample) of every diag entry with ndump > 0"""
    out = []
    for d in section.get("diag", []) if section else []:
        nd = int(d.get("ndump", 0))
        if nd <= 0:
            continue
        names = d.get("name", [])
        for n in ([names] if isinstance(names, str) else names):
            out.append((n, nd, int(d.get("psample", 1))))
    return out


class Diagnostics:
    """The diagnostics a deck asks for, bound to a Simulation."""

    def __init__(self, sim, directory="."):
        self.sim = sim
        self.dir = directory
        self.box = box_of(sim.deck)
        deck = sim.deck
        self.fields = _entries(deck.get("field"))
        for name, _, _ in self.fields:
            if name not in FIELD_DIAGS:
                raise L.QpadError(f"field diagnostic '{name}' is not on the GPU path")
        self.species = [(k, e) for k, sp in enumerate(deck.get("species", [])[:len(sim.spe)]) for e in _entries(sp)]
        self.beams = [(k, e) for k, bm in enumerate(deck.get("beam", [])[:len(sim.beam)]) for e in _entries(bm)]
        self.neutrals = [(k, e) for k, nt in enumerate(deck.get("neutrals", [])[:len(sim.neut)]) for e in _entries(nt)]
        for _, (name, _, _) in self.species + self.beams:
            if name not in ("charge_cyl_m", "raw"):
                raise L.QpadError(f"particle diagnostic '{name}' is not on the GPU path")
        for _, (name, _, _) in self.neutrals:
            if name not in ("charge_cyl_m", "ion_cyl_m", "raw"):
                raise L.QpadError(f"neutral diagnostic '{name}' is not on the GPU path")
        if any(name == "charge_cyl_m" for _, (name, _, _) in self.species + self.neutrals):
            L.check(sim.lib.qpad_species_charge_history(sim.ctx, 1))
        load()

    def run(self, tstep):
        """write the files due at 3D step tstep (run_sim_diag,
        diagnostics_class.f03:822-944)"""
        d, ctx, b, dirb = load(), self.sim.ctx, C.byref(self.box), self.dir.encode()
        for name, nd, _ in self.fields:
            if tstep % nd == 0:
                check(d.qpad_diag_field(ctx, name.encode(), dirb, tstep, b))
        for k, (name, nd, ps) in self.species:
            if tstep % nd:
                continue
            if name == "charge_cyl_m":
                check(d.qpad_diag_species_charge(ctx, k, dirb, tstep, b))
            else:
                check(d.qpad_diag_species_raw(ctx, k, ps, dirb, tstep, b))
        for k, (name, nd, ps) in self.neutrals:
            if tstep % nd:
                continue
            sp = self.sim.neut[k].id
            if name == "charge_cyl_m":
                check(d.qpad_diag_neutral_charge(ctx, sp, k, dirb, tstep, b))
            elif name == "ion_cyl_m":
                check(d.qpad_diag_neutral_ion(ctx, sp, k, dirb, tstep, b))
            else:
                check(d.qpad_diag_neutral_raw(ctx, sp, k, ps, dirb, tstep, b))
        for k, (name, nd, ps) in self.beams:
            if tstep % nd:
                continue
            if name == "charge_cyl_m":
                check(d.qpad_diag_beam_charge(ctx, k, dirb, tstep, b))
            else:
                check(d.qpad_diag_beam_raw(ctx, k, ps, dirb, tstep, b))
