"""ctypes binding of the CPU oracle (oracle/qpad_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / CPU baseline.  The product
(qpad_amd) never imports it.

Arrays returned by the accessors are numpy views on the oracle's own memory
(layout documented in qpad_oracle.h): slice fields are [ncomp, nm, ld] with
ld = nr + 2 (nodes 0..nr+1), histories are [nxi + 2, ncomp, nm, ld].
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libqpad_oracle.so")

PSI, E, E_SPE, B, B_SPE, B_BEAM, CU, ACU, AMU, DCU, Q_SPE, Q_BEAM = range(12)
NCOMP = [1, 3, 3, 3, 3, 3, 3, 2, 3, 2, 1, 1]
FIELD_NAMES = ["psi", "e", "e_spe", "b", "b_spe", "b_beam", "cu", "acu", "amu",
               "dcu", "q_spe", "q_beam"]
FK_PSI, FK_EZ, FK_BZ, FK_BT, FK_BPLUS, FK_BMINUS = range(6)
# QPAD's timer event names (sysutil_module.f03:307-330), in ora_timers order
TIMER_NAMES = ["solve psi", "solve bz", "solve beam bt", "solve plasma bt", "solve ez",
               "solve plasma et", "set source", "copy slices", "arithmetics", "deposit 2D particles",
               "push 2D particles", "move 2D particles", "deposit 3D particles", "push 3D particles"]
BND_ZERO, BND_OPEN = 2, 3

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Grid(C.Structure):
    _fields_ = [("nr", C.c_int), ("nxi", C.c_int), ("max_mode", C.c_int),
                ("dr", C.c_double), ("dxi", C.c_double), ("relax_fac", C.c_double),
                ("bnd", C.c_int), ("iter_max", C.c_int), ("iter_reltol", C.c_double),
                ("iter_abstol", C.c_double), ("dt", C.c_double)]


class SpeciesCfg(C.Structure):
    _fields_ = [("ppc1", C.c_int), ("ppc2", C.c_int), ("num_theta", C.c_int),
                ("qm", C.c_double), ("mass", C.c_double), ("density", C.c_double),
                ("uth", C.c_double * 3), ("neutralized", C.c_int), ("den_min", C.c_double),
                ("seed", C.c_uint64), ("push_type", C.c_int)]


PUSH2_STD, PUSH2_ROBUST, PUSH2_STD_PGC, PUSH2_ROBUST_PGC = 0, 1, 4, 5
LASER_AR, LASER_AI, LASER_AR_GRAD, LASER_AI_GRAD, LASER_CHI, LASER_H_CHI, LASER_HIST_AR, LASER_HIST_AI = range(8)


class BeamCfg(C.Structure):
    _fields_ = [("total_num", C.c_int64), ("total_charge", C.c_double), ("qm", C.c_double),
                ("mass", C.c_double), ("gamma", C.c_double), ("center", C.c_double * 3),
                ("sigma", C.c_double * 3), ("range_lo", C.c_double * 3),
                ("range_hi", C.c_double * 3), ("uth", C.c_double * 3), ("quiet", C.c_int),
                ("evol", C.c_int), ("seed", C.c_uint64)]


class BeamStdCfg(C.Structure):
    _fields_ = [("ppc", C.c_int * 3), ("num_theta", C.c_int), ("geometry", C.c_int),
                ("profile", C.c_int * 3), ("qm", C.c_double), ("mass", C.c_double),
                ("density", C.c_double), ("gamma", C.c_double), ("den_min", C.c_double),
                ("gauss_center", C.c_double * 3), ("gauss_sigma", C.c_double * 3),
                ("pw_n", C.c_int * 3), ("pw_x", (C.c_double * 32) * 3),
                ("pw_fx", (C.c_double * 32) * 3), ("range_lo", C.c_double * 3),
                ("range_hi", C.c_double * 3), ("uth", C.c_double * 3), ("alpha", C.c_double * 2),
                ("twiss_ctr", C.c_double * 2), ("twiss_beta", C.c_double * 2),
                ("offset_n", C.c_int * 2), ("offset_xi0", C.c_double * 2),
                ("offset_c", (C.c_double * 8) * 2), ("quiet", C.c_int), ("evol", C.c_int),
                ("seed", C.c_uint64)]


class _Species(C.Structure):
    _fields_ = [("npp", C.c_int64), ("npmax", C.c_int64),
                ("x1", _dp), ("x2", _dp), ("p1", _dp), ("p2", _dp), ("p3", _dp),
                ("q", _dp), ("gamma", _dp), ("psi", _dp), ("id", _ip),
                ("qbm", C.c_double), ("qm", C.c_double), ("dt", C.c_double),
                ("cfg", SpeciesCfg),
                ("q_own", _dp), ("qn", _dp), ("cu_own", _dp), ("acu_own", _dp),
                ("amu_own", _dp), ("q_hist", _dp),
                ("is_neutral", C.c_int), ("multi_max", C.c_int), ("wp", C.c_double), ("rate", _dp),
                ("multi_ion", _dp), ("ion_old", _dp), ("ppc_add", C.POINTER(C.c_int)),
                ("add_x1", _dp), ("add_x2", _dp), ("add_q", _dp), ("n_add", C.c_int64),
                ("n_added", C.c_int64), ("rho_hist", _dp), ("den_perp", _dp), ("den_lon", C.c_double)]


class NeutralCfg(C.Structure):
    _fields_ = [("ppc1", C.c_int), ("ppc2", C.c_int), ("num_theta", C.c_int),
                ("qm", C.c_double), ("mass", C.c_double), ("density", C.c_double),
                ("den_min", C.c_double), ("element", C.c_int), ("ion_max", C.c_int),
                ("push_type", C.c_int), ("n0", C.c_double)]


def neutral_cfg(ppc1, ppc2, num_theta, element, ion_max, n0, qm=-1.0, mass=1.0, density=1.0,
                den_min=1e-10, push_type=PUSH2_ROBUST):
    return NeutralCfg(ppc1, ppc2, num_theta, qm, mass, density, den_min, element, ion_max, push_type, n0)


class _Beam(C.Structure):
    _fields_ = [("npp", C.c_int64), ("npmax", C.c_int64),
                ("x1", _dp), ("x2", _dp), ("x3", _dp), ("p1", _dp), ("p2", _dp),
                ("p3", _dp), ("q", _dp), ("id", _ip),
                ("qbm", C.c_double), ("dt", C.c_double), ("evol", C.c_int),
                ("q_own_hist", _dp)]


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load_library(LIB_PATH)
    return _lib


def load_library(path):
    """A ctypes handle on an oracle library at `path` (the built oracle, or
    a copy compiled from modified sources by the pin tests)."""
    L = C.CDLL(path)
    L.ora_create.restype = C.c_void_p
    L.ora_create.argtypes = [C.POINTER(Grid)]
    L.ora_destroy.argtypes = [C.c_void_p]
    L.ora_add_species.argtypes = [C.c_void_p, C.POINTER(SpeciesCfg)]
    L.ora_add_beam.argtypes = [C.c_void_p, C.POINTER(BeamCfg), C.c_int64]
    L.ora_add_beam_std.argtypes = [C.c_void_p, C.POINTER(BeamStdCfg)]
    L.ora_add_beam_std_table.argtypes = [C.c_void_p, C.POINTER(BeamStdCfg), _dp, C.c_int, C.c_int, C.c_int]
    L.ora_beam_set.argtypes = [C.c_void_p, C.c_int, C.c_int64] + [_dp] * 7
    L.ora_step3d_begin.argtypes = [C.c_void_p]
    L.ora_slice.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    L.ora_step3d_end.argtypes = [C.c_void_p, C.c_int]
    L.ora_run.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.ora_field.restype = _dp
    L.ora_field.argtypes = [C.c_void_p, C.c_int]
    L.ora_hist.restype = _dp
    L.ora_hist.argtypes = [C.c_void_p, C.c_int]
    L.ora_get_species.restype = C.POINTER(_Species)
    L.ora_get_species.argtypes = [C.c_void_p, C.c_int]
    L.ora_get_beam.restype = C.POINTER(_Beam)
    L.ora_get_beam.argtypes = [C.c_void_p, C.c_int]
    L.ora_matrix.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                             _dp, _dp, _dp]
    L.ora_thomas_factor.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    L.ora_thomas_solve.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
    for name in ("ora_solve_psi", "ora_solve_ez", "ora_solve_bz", "ora_solve_bt_beam"):
        getattr(L, name).argtypes = [C.c_void_p, _dp, _dp]
    for name in ("ora_solve_et", "ora_solve_bt_iter", "ora_solve_djdxi"):
        getattr(L, name).argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.ora_conv_record.argtypes = [C.c_void_p, _dp]
    L.ora_conv_compare.argtypes = [C.c_void_p, _dp, _dp, _dp]
    L.ora_qdeposit.argtypes = [C.c_void_p, C.POINTER(_Species), _dp]
    L.ora_amjdeposit_robust.argtypes = [C.c_void_p, C.POINTER(_Species), _dp, _dp, _dp, _dp,
                                        _dp, C.c_double]
    L.ora_push_u_robust.argtypes = [C.c_void_p, C.POINTER(_Species), _dp, _dp, C.c_double]
    L.ora_push_x.argtypes = [C.POINTER(_Species), C.c_double]
    L.ora_interp_psi.argtypes = [C.c_void_p, C.POINTER(_Species), _dp]
    L.ora_amjdeposit_std.argtypes = [C.c_void_p, C.POINTER(_Species), _dp, _dp, _dp, _dp, _dp, C.c_double]
    L.ora_push_u_std.argtypes = [C.c_void_p, C.POINTER(_Species), _dp, _dp, C.c_double]
    L.ora_set_laser.argtypes = [C.c_void_p, _dp, _dp]
    L.ora_laser_slice.argtypes = [C.c_void_p, C.c_int]
    L.ora_amjdeposit_pgc.argtypes = [C.c_void_p, C.POINTER(_Species), _dp, _dp, _dp, _dp, _dp, C.c_double,
                                     C.c_int]
    L.ora_push_u_pgc.argtypes = [C.c_void_p, C.POINTER(_Species), _dp, _dp, C.c_double]
    L.ora_deposit_chi.argtypes = [C.c_void_p, C.POINTER(_Species), _dp]
    L.ora_laser_array.restype = _dp
    L.ora_laser_array.argtypes = [C.c_void_p, C.c_int]
    L.ora_update_bound.argtypes = [C.c_void_p, C.POINTER(_Species)]
    L.ora_add_neutral.argtypes = [C.c_void_p, C.POINTER(NeutralCfg)]
    L.ora_adk.restype = _dp
    L.ora_adk.argtypes = [C.c_int, C.c_int]
    L.ora_neutral_ionize.argtypes = [C.c_void_p, C.POINTER(_Species), _dp]
    L.ora_neutral_add.argtypes = [C.c_void_p, C.POINTER(_Species)]
    L.ora_neutral_ion_deposit.argtypes = [C.c_void_p, C.POINTER(_Species)]
    L.ora_neutral_renew.argtypes = [C.c_void_p, C.POINTER(_Species)]
    L.ora_interp_point.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double, C.c_double, _dp]
    L.ora_species_set_profile.argtypes = [C.c_void_p, C.c_int, _dp, C.c_double]
    L.ora_set_subcycling.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double]
    L.ora_exp_fac_max.restype = C.c_double
    L.ora_exp_fac_max.argtypes = [C.c_void_p]
    L.ora_subcyc_step.restype = C.c_int
    L.ora_subcyc_step.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, _dp]
    L.ora_clamp_exp_fac.argtypes = [C.POINTER(_Species), C.c_double]
    L.ora_last_subcycles.argtypes = [C.c_void_p, C.POINTER(C.c_int), _dp, _dp]
    L.ora_beam_qdeposit.argtypes = [C.c_void_p, C.c_int, _dp]
    L.ora_beam_push.argtypes = [C.c_void_p, C.c_int]
    L.ora_species_inject.argtypes = [C.c_void_p, C.c_int]
    L.ora_species_init_q.argtypes = [C.c_void_p, C.c_int]
    L.ora_last_iters.argtypes = [C.c_void_p]
    L.ora_last_residuals.argtypes = [C.c_void_p, _dp, _dp]
    L.ora_philox4x32.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32)]
    L.ora_normal4.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, _dp]
    L.ora_timing.argtypes = [C.c_int]
    L.ora_timers.argtypes = [_dp]
    return L


def timers(L=None):
    """seconds per QPAD timer event since timing(True) (this process)"""
    out = np.zeros(len(TIMER_NAMES))
    (L or lib()).ora_timers(ptr(out))
    return dict(zip(TIMER_NAMES, out.tolist()))


def timing(enable, L=None):
    (L or lib()).ora_timing(int(enable))


def ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def make_grid(nr, nxi, max_mode, dr, dxi, relax_fac=None, bnd=BND_OPEN, iter_max=10,
              reltol=1e-3, abstol=1e-3, dt=10.0):
    if relax_fac is None:
        relax_fac = 1.0e-3 * (dr / 0.02) ** 2   # sim_fields_class.f03:137
    return Grid(nr, nxi, max_mode, dr, dxi, relax_fac, bnd, iter_max, reltol, abstol, dt)


def species_cfg(ppc1, ppc2, num_theta, qm=-1.0, mass=1.0, density=1.0, uth=(0.0, 0.0, 0.0),
                neutralized=True, den_min=1e-10, seed=0, push_type=PUSH2_ROBUST):
    return SpeciesCfg(ppc1, ppc2, num_theta, qm, mass, density, (C.c_double * 3)(*uth),
                      int(neutralized), den_min, seed, push_type)


def beam_cfg(total_num, total_charge, center, sigma, range_lo, range_hi, qm=-1.0, mass=1.0,
             gamma=20000.0, uth=(0.0, 0.0, 0.0), quiet=False, evol=True, seed=1):
    d3 = C.c_double * 3
    return BeamCfg(total_num, total_charge, qm, mass, gamma, d3(*center), d3(*sigma),
                   d3(*range_lo), d3(*range_hi), d3(*uth), int(quiet), int(evol), seed)


def file_beam_place(raw, beam_center, file_center, lconv, qconv, nr, dr, nxi, dxi):
    """Particles of a "file" beam in box coordinates, one particle at a time
    as inject_fdist3d_file does (fdist3d_file_class.f03:300-330): centre on
    file_center, scale by lconv, shift to beam_center (x3 mirrored), keep
    0 <= r < nr*dr and 0 <= xi < nxi*dxi in file order, q scaled by qconv.
    beam_center[2] is relative to the box start (init_fdist3d_file :84-85)."""
    import math
    keys = ("x1", "x2", "x3", "p1", "p2", "p3", "q")
    out = {k: [] for k in keys}
    n = len(np.asarray(raw["x1"]).ravel())
    cols = {k: np.asarray(raw[k], dtype=np.float64).ravel() for k in keys}
    for i in range(n):
        x = [(float(cols[k][i]) - file_center[d]) * lconv for d, k in enumerate(("x1", "x2", "x3"))]
        x = [beam_center[0] + x[0], beam_center[1] + x[1], beam_center[2] - x[2]]
        r = math.sqrt(x[0] ** 2 + x[1] ** 2)
        if 0.0 <= r < nr * dr and 0.0 <= x[2] < nxi * dxi:
            for d, k in enumerate(("x1", "x2", "x3")):
                out[k].append(x[d])
            for k in ("p1", "p2", "p3"):
                out[k].append(float(cols[k][i]))
            out["q"].append(float(cols["q"][i]) * qconv)
    return {k: np.array(v, dtype=np.float64) for k, v in out.items()}


def matrix(kind, mode, nr, dr, relax_fac, bnd=BND_OPEN, L=None):
    a, b, c = (np.zeros(nr) for _ in range(3))
    (L or lib()).ora_matrix(kind, mode, nr, dr, relax_fac, bnd, ptr(a), ptr(b), ptr(c))
    return a, b, c


def thomas(a, b, c, d):
    n = len(b)
    lo, inv, cp = (np.zeros(n) for _ in range(3))
    lib().ora_thomas_factor(n, ptr(np.ascontiguousarray(a)), ptr(np.ascontiguousarray(b)),
                            ptr(np.ascontiguousarray(c)), ptr(lo), ptr(inv), ptr(cp))
    x = np.array(d, dtype=np.float64)
    lib().ora_thomas_solve(n, ptr(lo), ptr(inv), ptr(cp), ptr(x))
    return x


def philox4x32(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().ora_philox4x32(c, k, o)
    return list(o)


def normal4(seed, stream, sub):
    out = np.zeros(4)
    lib().ora_normal4(seed, stream, sub, ptr(out))
    return out


class Oracle:
    """One oracle simulation instance (simulation_class restated).  `L`: the
    library to run (default: the built oracle)."""

    def __init__(self, grid, L=None):
        self.L = L or lib()
        self.grid = grid
        self.h = self.L.ora_create(C.byref(grid))
        self.nm = 2 * grid.max_mode + 1
        self.ld = grid.nr + 2
        self.nspecies = 0
        self.nbeams = 0

    def close(self):
        if self.h:
            self.L.ora_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- construction
    def add_species(self, cfg):
        r = self.L.ora_add_species(self.h, C.byref(cfg))
        self.nspecies += 1
        return r

    def add_neutral(self, cfg):
        r = self.L.ora_add_neutral(self.h, C.byref(cfg))
        assert r >= 0, "neutral refused"
        self.nspecies += 1
        return r

    def neutral_levels(self, isp):
        """multi_ion of neutral isp as [num_theta, nr, multi_max + 2] (levels
        1..L, the neutral residue, the discrete total ion charge)"""
        sp = self.species_raw(isp).contents
        nth, nr, L = sp.cfg.num_theta, self.grid.nr, sp.multi_max
        n = nth * nr * (L + 2)
        return np.ctypeslib.as_array(sp.multi_ion, shape=(n,)).reshape(nth, nr, L + 2)

    def neutral_ppc_add(self, isp):
        sp = self.species_raw(isp).contents
        n = sp.cfg.num_theta * self.grid.nr
        return np.ctypeslib.as_array(sp.ppc_add, shape=(n,)).reshape(sp.cfg.num_theta, self.grid.nr)

    def neutral_rho_hist(self, isp):
        """rho_ion history [nxi + 2, 1, nm, ld] (neutral%cbq)"""
        sp = self.species_raw(isp).contents
        ns = self.grid.nxi + 2
        n = ns * self.nm * self.ld
        return np.ctypeslib.as_array(sp.rho_hist, shape=(n,)).reshape(ns, 1, self.nm, self.ld)

    def add_beam(self, cfg, npmax=0):
        r = self.L.ora_add_beam(self.h, C.byref(cfg), npmax)
        self.nbeams += 1
        return r

    def add_beam_std(self, cfg):
        r = self.L.ora_add_beam_std(self.h, C.byref(cfg))
        self.nbeams += 1
        return r

    def add_beam_std_table(self, cfg, den, k0, nk, ni):
        """the lattice beam with a tabulated density over the box k0, nk, ni"""
        d = np.ascontiguousarray(den, dtype=np.float64)
        r = self.L.ora_add_beam_std_table(self.h, C.byref(cfg), ptr(d), int(k0), int(nk), int(ni))
        self.nbeams += 1
        return r

    def set_beam(self, ib, x1, x2, x3, p1, p2, p3, q):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x1, x2, x3, p1, p2, p3, q)]
        rc = self.L.ora_beam_set(self.h, ib, len(arrs[0]), *[ptr(a) for a in arrs])
        assert rc == 0

    # -- stepping
    def step_begin(self):
        self.L.ora_step3d_begin(self.h)

    def slice(self, j):
        n = C.c_int(0)
        self.L.ora_slice(self.h, j, C.byref(n))
        return n.value

    def step_end(self, i=1):
        self.L.ora_step3d_end(self.h, i)

    def run(self, nstep3d, nslices=0):
        self.L.ora_run(self.h, nstep3d, nslices)

    # -- state access (views)
    def field(self, fid):
        p = self.L.ora_field(self.h, fid)
        n = NCOMP[fid] * self.nm * self.ld
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(NCOMP[fid], self.nm, self.ld)

    def hist(self, fid):
        p = self.L.ora_hist(self.h, fid)
        ns = self.grid.nxi + 2
        n = ns * NCOMP[fid] * self.nm * self.ld
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(ns, NCOMP[fid], self.nm, self.ld)

    # -- laser (PGC)
    def set_laser(self, ar_hist, ai_hist):
        """The laser envelope history (real and imaginary part of the complex
        envelope), [nxi + 3, nm, ld] each: slices -1 .. nxi + 1."""
        shape = (self.grid.nxi + 3, self.nm, self.ld)
        a = np.ascontiguousarray(ar_hist, dtype=np.float64)
        b = np.ascontiguousarray(ai_hist, dtype=np.float64)
        assert a.shape == shape and b.shape == shape, (a.shape, shape)
        assert self.L.ora_set_laser(self.h, ptr(a), ptr(b)) == 0

    def laser(self, which):
        """View of a laser array (LASER_*); None before set_laser or a PGC
        species"""
        p = self.L.ora_laser_array(self.h, which)
        if not p:
            return None
        nm, ld, nxi = self.nm, self.ld, self.grid.nxi
        shape = {LASER_AR: (1, nm, ld), LASER_AI: (1, nm, ld), LASER_AR_GRAD: (3, nm, ld),
                 LASER_AI_GRAD: (3, nm, ld), LASER_CHI: (1, nm, ld), LASER_H_CHI: (nxi + 2, 1, nm, ld),
                 LASER_HIST_AR: (nxi + 3, nm, ld), LASER_HIST_AI: (nxi + 3, nm, ld)}[which]
        return np.ctypeslib.as_array(p, shape=(int(np.prod(shape)),)).reshape(shape)

    def laser_slice(self, j):
        self.L.ora_laser_slice(self.h, j)

    def set_profile(self, isp, den_perp, den_lon):
        """the profile of species isp's next injections (den_perp None: uniform)"""
        d = None if den_perp is None else np.ascontiguousarray(den_perp, dtype=np.float64)
        assert self.L.ora_species_set_profile(self.h, isp, None if d is None else ptr(d), float(den_lon)) == 0

    def renew_species(self, isp):
        """species2d%renew: inject again with the current profile, charge set-up"""
        self.L.ora_species_inject(self.h, isp)
        self.L.ora_species_init_q(self.h, isp)

    def species_raw(self, isp):
        return self.L.ora_get_species(self.h, isp)

    def species(self, isp):
        s = self.species_raw(isp).contents
        n = int(s.npp)
        out = {}
        for name in ("x1", "x2", "p1", "p2", "p3", "q", "gamma", "psi"):
            out[name] = np.ctypeslib.as_array(getattr(s, name), shape=(int(s.npmax),))[:n]
        out["id"] = np.ctypeslib.as_array(s.id, shape=(int(s.npmax),))[:n]
        return out

    def set_species(self, isp, arrs):
        """Overwrite species isp's particles (used to feed identical inputs)."""
        s = self.species_raw(isp).contents
        n = len(arrs["x1"])
        assert n <= s.npmax
        for name in ("x1", "x2", "p1", "p2", "p3", "q", "gamma", "psi"):
            dst = np.ctypeslib.as_array(getattr(s, name), shape=(int(s.npmax),))
            src = arrs.get(name)
            dst[:n] = src if src is not None else 0.0
        ids = np.ctypeslib.as_array(s.id, shape=(int(s.npmax),))
        ids[:n] = arrs.get("id", np.arange(n))
        self.species_raw(isp).contents.npp = n

    def species_q_hist(self, isp):
        """species isp's charge history (species2d%q written by cbq), [nxi + 2, 1, nm, ld]"""
        s = self.species_raw(isp).contents
        ns = self.grid.nxi + 2
        n = ns * self.nm * self.ld
        return np.ctypeslib.as_array(s.q_hist, shape=(n,)).reshape(ns, 1, self.nm, self.ld)

    def species_q_own(self, isp):
        s = self.species_raw(isp).contents
        n = self.nm * self.ld
        return (np.ctypeslib.as_array(s.q_own, shape=(n,)).reshape(1, self.nm, self.ld),
                np.ctypeslib.as_array(s.qn, shape=(n,)).reshape(1, self.nm, self.ld))

    def beam(self, ib):
        b = self.L.ora_get_beam(self.h, ib).contents
        n = int(b.npp)
        out = {}
        for name in ("x1", "x2", "x3", "p1", "p2", "p3", "q"):
            out[name] = np.ctypeslib.as_array(getattr(b, name), shape=(int(b.npmax),))[:n]
        out["id"] = np.ctypeslib.as_array(b.id, shape=(int(b.npmax),))[:n]
        return out

    def last_iters(self):
        return self.L.ora_last_iters(self.h)

    # -- sub-cycling (proj_subcyc)
    def set_subcycling(self, exp_fac_max, exp_fac_clamped, dt_2d_min, on=True):
        self.L.ora_set_subcycling(self.h, int(on), exp_fac_max, exp_fac_clamped, dt_2d_min)

    def exp_fac_max(self):
        return self.L.ora_exp_fac_max(self.h)

    def last_subcycles(self):
        """(sub-cycles, their 2D step, the expansion factor) of the last slice"""
        n, d, f = C.c_int(0), np.zeros(1), np.zeros(1)
        self.L.ora_last_subcycles(self.h, C.byref(n), ptr(d), ptr(f))
        return n.value, float(d[0]), float(f[0])

    def last_residuals(self):
        """(relative, absolute) residual of the last slice's final 'compare'"""
        r, a = np.zeros(1), np.zeros(1)
        self.L.ora_last_residuals(self.h, ptr(r), ptr(a))
        return float(r[0]), float(a[0])

    # -- building blocks on this instance's grid
    def new_field(self, ncomp):
        return np.zeros((ncomp, self.nm, self.ld))

    def solve(self, name, *arrays):
        fn = getattr(self.L, "ora_solve_" + name)
        fn(self.h, *[ptr(a) for a in arrays])


def subcyc_step(exp_fac, exp_fac_max, dt, dt_min, L=None):
    """get_subcyc_step: (sub-cycles, their 2D step)"""
    d = np.zeros(1)
    n = (L or lib()).ora_subcyc_step(exp_fac, exp_fac_max, dt, dt_min, ptr(d))
    return n, float(d[0])
