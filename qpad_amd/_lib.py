"""ctypes binding of libqpad_hip.so (C ABI in include/qpad_hip.h).

The library is the only compute path: importing this module without the
built .so raises, there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqpad_hip.so")

F_PSI, F_E, F_E_SPE, F_B, F_B_SPE, F_B_BEAM, F_CU, F_ACU, F_AMU, F_DCU, F_Q_SPE, F_Q_BEAM = range(12)
F_AR, F_AI, F_AR_GRAD, F_AI_GRAD, F_CHI = range(12, 17)
NFIELD = 17
NCOMP = [1, 3, 3, 3, 3, 3, 3, 2, 3, 2, 1, 1, 1, 1, 3, 3, 1]
FIELD_IDS = {"psi": F_PSI, "e": F_E, "e_spe": F_E_SPE, "b": F_B, "b_spe": F_B_SPE,
             "b_beam": F_B_BEAM, "cu": F_CU, "acu": F_ACU, "amu": F_AMU, "dcu": F_DCU,
             "q_spe": F_Q_SPE, "q_beam": F_Q_BEAM, "ar": F_AR, "ai": F_AI, "ar_grad": F_AR_GRAD,
             "ai_grad": F_AI_GRAD, "chi": F_CHI}
BND_ZERO, BND_OPEN = 2, 3
PUSH2_STD, PUSH2_ROBUST, PUSH2_STD_PGC, PUSH2_ROBUST_PGC = 0, 1, 4, 5
# status codes (include/qpad_hip.h)
OK, E_INVALID, E_HIP, E_NOMEM, E_STATE = 0, -1, -2, -3, -4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_fp = C.POINTER(C.c_float)


class QpadError(RuntimeError):
    """A call into libqpad_hip failed (QPAD's write_err equivalent)."""


class Grid(C.Structure):
    _fields_ = [("nr", C.c_int32), ("nxi", C.c_int32), ("max_mode", C.c_int32),
                ("dr", C.c_double), ("dxi", C.c_double), ("relax_fac", C.c_double),
                ("bnd", C.c_int32), ("iter_max", C.c_int32), ("iter_reltol", C.c_double),
                ("iter_abstol", C.c_double), ("dt", C.c_double)]


class SpeciesCfg(C.Structure):
    _fields_ = [("ppc1", C.c_int32), ("ppc2", C.c_int32), ("num_theta", C.c_int32),
                ("qm", C.c_double), ("mass", C.c_double), ("density", C.c_double),
                ("uth", C.c_double * 3), ("neutralized", C.c_int32), ("den_min", C.c_double),
                ("seed", C.c_uint64), ("push_type", C.c_int32)]


class SeedNetCfg(C.Structure):
    _fields_ = [("width", C.c_int32), ("layers", C.c_int32), ("lr", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("out_scale", C.c_float), ("seed", C.c_uint64)]


class SeedNetInfo(C.Structure):
    _fields_ = [("layers", C.c_int32), ("cin", C.c_int32), ("cout", C.c_int32),
                ("npos", C.c_int32), ("cin_p", C.c_int32 * 8), ("cout_p", C.c_int32 * 8),
                ("nparams", C.c_int64), ("steps", C.c_int64)]


class NeutralCfg(C.Structure):
    _fields_ = [("ppc1", C.c_int32), ("ppc2", C.c_int32), ("num_theta", C.c_int32),
                ("qm", C.c_double), ("mass", C.c_double), ("density", C.c_double),
                ("den_min", C.c_double), ("element", C.c_int32), ("ion_max", C.c_int32),
                ("push_type", C.c_int32), ("n0", C.c_double)]


class LaunchShape(C.Structure):
    _fields_ = [("dep_cells", C.c_int32), ("dep_mfma", C.c_int32)]


class BeamCfg(C.Structure):
    _fields_ = [("total_num", C.c_int64), ("total_charge", C.c_double), ("qm", C.c_double),
                ("mass", C.c_double), ("gamma", C.c_double), ("center", C.c_double * 3),
                ("sigma", C.c_double * 3), ("range_lo", C.c_double * 3),
                ("range_hi", C.c_double * 3), ("uth", C.c_double * 3), ("quiet", C.c_int32),
                ("evol", C.c_int32), ("seed", C.c_uint64)]


PROF_UNIFORM, PROF_GAUSSIAN, PROF_PIECEWISE = 0, 1, 2
GEOM_CARTESIAN, GEOM_CYLINDRICAL = 0, 1
PW_MAX, OFFSET_MAX = 32, 8


class BeamStdCfg(C.Structure):
    _fields_ = [("ppc", C.c_int32 * 3), ("num_theta", C.c_int32), ("geometry", C.c_int32),
                ("profile", C.c_int32 * 3), ("qm", C.c_double), ("mass", C.c_double),
                ("density", C.c_double), ("gamma", C.c_double), ("den_min", C.c_double),
                ("gauss_center", C.c_double * 3), ("gauss_sigma", C.c_double * 3),
                ("pw_n", C.c_int32 * 3), ("pw_x", (C.c_double * PW_MAX) * 3),
                ("pw_fx", (C.c_double * PW_MAX) * 3), ("range_lo", C.c_double * 3),
                ("range_hi", C.c_double * 3), ("uth", C.c_double * 3), ("alpha", C.c_double * 2),
                ("twiss_ctr", C.c_double * 2), ("twiss_beta", C.c_double * 2),
                ("offset_n", C.c_int32 * 2), ("offset_xi0", C.c_double * 2),
                ("offset_c", (C.c_double * OFFSET_MAX) * 2), ("quiet", C.c_int32), ("evol", C.c_int32),
                ("seed", C.c_uint64)]


# name -> (restype, argtypes); every entry point returns int status
_CTX = C.c_void_p
SIGNATURES = {
    "qpad_last_error": (C.c_char_p, []),
    "qpad_abi_version": (C.c_int, []),
    "qpad_create": (C.c_int, [C.POINTER(Grid), C.POINTER(C.c_void_p)]),
    "qpad_destroy": (C.c_int, [_CTX]),
    "qpad_get_grid": (C.c_int, [_CTX, C.POINTER(Grid)]),
    "qpad_species_new": (C.c_int, [_CTX, C.POINTER(SpeciesCfg), C.POINTER(C.c_int32)]),
    "qpad_species_renew": (C.c_int, [_CTX, C.c_int32]),
    "qpad_species_qdp": (C.c_int, [_CTX, C.c_int32]),
    "qpad_species_qdp_noalloc": (C.c_int, [_CTX, C.c_int32]),
    "qpad_species_charge_history": (C.c_int, [_CTX, C.c_int32]),
    "qpad_species_cbq": (C.c_int, [_CTX, C.c_int32, C.c_int32]),
    "qpad_get_species_charge": (C.c_int, [_CTX, C.c_int32, C.c_int32, _dp]),
    "qpad_species_amjdp": (C.c_int, [_CTX, C.c_int32, C.c_double]),
    "qpad_species_push": (C.c_int, [_CTX, C.c_int32, C.c_double]),
    "qpad_species_push_clamped": (C.c_int, [_CTX, C.c_int32, C.c_double, C.c_double]),
    "qpad_set_subcycling": (C.c_int, [_CTX, C.c_int32, C.c_double, C.c_double, C.c_double]),
    "qpad_exp_fac_max": (C.c_int, [_CTX, _dp]),
    "qpad_species_set_profile": (C.c_int, [_CTX, C.c_int32, _dp, C.c_double]),
    "qpad_beam_lattice_box": (C.c_int, [_CTX, C.POINTER(BeamStdCfg), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32)]),
    "qpad_beam_new_std_density": (C.c_int, [_CTX, C.POINTER(BeamStdCfg), _dp, C.c_int32, C.c_int32, C.c_int32,
                                            C.POINTER(C.c_int32)]),
    "qpad_last_subcycles": (C.c_int, [_CTX, C.POINTER(C.c_int32), _dp, _dp]),
    "qpad_species_interp_psi": (C.c_int, [_CTX, C.c_int32]),
    "qpad_set_laser": (C.c_int, [_CTX, _dp, _dp]),
    "qpad_neutral_new": (C.c_int, [_CTX, C.POINTER(NeutralCfg), C.POINTER(C.c_int32)]),
    "qpad_neutral_ion_deposit": (C.c_int, [_CTX, C.c_int32]),
    "qpad_neutral_update": (C.c_int, [_CTX, C.c_int32]),
    "qpad_get_neutral_levels": (C.c_int, [_CTX, C.c_int32, _dp]),
    "qpad_get_ion_density": (C.c_int, [_CTX, C.c_int32, C.c_int32, _dp]),
    "qpad_neutral_adk": (C.c_int, [C.c_int32, C.c_int32, _dp]),
    "qpad_laser_slice": (C.c_int, [_CTX, C.c_int32]),
    "qpad_deposit_chi": (C.c_int, [_CTX, C.c_int32]),
    "qpad_step_iterations": (C.c_int, [_CTX, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "qpad_field_solve_psi": (C.c_int, [_CTX]),
    "qpad_field_solve_ez": (C.c_int, [_CTX, C.c_int32, C.c_int32]),
    "qpad_field_solve_et": (C.c_int, [_CTX, C.c_int32, C.c_int32]),
    "qpad_field_solve_bz": (C.c_int, [_CTX, C.c_int32, C.c_int32]),
    "qpad_field_solve_bt_beam": (C.c_int, [_CTX, C.c_int32]),
    "qpad_field_solve_djdxi": (C.c_int, [_CTX]),
    "qpad_field_solve_bt_iter": (C.c_int, [_CTX]),
    "qpad_field_add_b": (C.c_int, [_CTX]),
    "qpad_convergence_tester": (C.c_int, [_CTX, C.c_int32, _dp, _dp]),
    "qpad_beam_qdp": (C.c_int, [_CTX]),
    "qpad_beam_push": (C.c_int, [_CTX]),
    "qpad_beam_new": (C.c_int, [_CTX, C.POINTER(BeamCfg), C.POINTER(C.c_int32)]),
    "qpad_beam_new_std": (C.c_int, [_CTX, C.POINTER(BeamStdCfg), C.POINTER(C.c_int32)]),
    "qpad_beam_new_arrays": (C.c_int, [_CTX, C.c_double, C.c_double, C.c_int32, C.c_int64] + [_dp] * 7
                             + [C.POINTER(C.c_int32)]),
    "qpad_step3d_begin": (C.c_int, [_CTX]),
    "qpad_slice": (C.c_int, [_CTX, C.c_int32, C.POINTER(C.c_int32)]),
    "qpad_step3d_end": (C.c_int, [_CTX]),
    "qpad_step3d": (C.c_int, [_CTX, C.POINTER(C.c_int32)]),
    "qpad_iteration_policy": (C.c_int, [_CTX, C.c_int32, C.c_int32]),
    "qpad_get_launch_shape": (C.c_int, [_CTX, C.POINTER(LaunchShape)]),
    "qpad_set_launch_shape": (C.c_int, [_CTX, C.POINTER(LaunchShape)]),
    "qpad_tune_launch": (C.c_int, [_CTX, C.c_int32, _dp]),
    "qpad_get_field": (C.c_int, [_CTX, C.c_int32, _dp]),
    "qpad_set_field": (C.c_int, [_CTX, C.c_int32, _dp]),
    "qpad_get_hist": (C.c_int, [_CTX, C.c_int32, C.c_int32, _dp]),
    "qpad_set_hist": (C.c_int, [_CTX, C.c_int32, C.c_int32, _dp]),
    "qpad_species_count": (C.c_int, [_CTX, C.c_int32, C.POINTER(C.c_int64)]),
    "qpad_get_species": (C.c_int, [_CTX, C.c_int32] + [_dp] * 8 + [_ip]),
    "qpad_set_species": (C.c_int, [_CTX, C.c_int32, C.c_int64] + [_dp] * 8 + [_ip]),
    "qpad_get_species_qn": (C.c_int, [_CTX, C.c_int32, _dp]),
    "qpad_set_species_qn": (C.c_int, [_CTX, C.c_int32, _dp]),
    "qpad_beam_count": (C.c_int, [_CTX, C.c_int32, C.POINTER(C.c_int64)]),
    "qpad_get_beam": (C.c_int, [_CTX, C.c_int32] + [_dp] * 7),
    "qpad_set_beam": (C.c_int, [_CTX, C.c_int32, C.c_int64] + [_dp] * 7),
    "qpad_kernel_time": (C.c_int, [_CTX, C.c_int32, _dp, C.POINTER(C.c_int64)]),
    "qpad_kernel_time_reset": (C.c_int, [_CTX, C.c_int32]),
    "qpad_seednet_create": (C.c_int, [_CTX, C.POINTER(SeedNetCfg)]),
    "qpad_seednet_get_info": (C.c_int, [_CTX, C.POINTER(SeedNetInfo)]),
    "qpad_seednet_mode": (C.c_int, [_CTX, C.c_int32, C.c_int32]),
    "qpad_seednet_get_params": (C.c_int, [_CTX, _fp, C.c_int64]),
    "qpad_seednet_set_params": (C.c_int, [_CTX, _fp, C.c_int64]),
    "qpad_seednet_losses": (C.c_int, [_CTX, C.c_int64, C.c_int64, _fp]),
    "qpad_seednet_seed": (C.c_int, [_CTX, C.c_int32]),
    "qpad_seednet_learn": (C.c_int, [_CTX, C.c_int32]),
    "qpad_seednet_eval": (C.c_int, [_CTX, _fp, _fp, C.c_int32, _fp, _fp, _fp, C.c_int32]),
    "qpad_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]),
    "qpad_batch_destroy": (C.c_int, [C.c_void_p]),
    "qpad_batch_mode": (C.c_int, [C.c_void_p, C.c_int32]),
    "qpad_batch_step3d": (C.c_int, [C.c_void_p, _ip]),
    "qpad_batch_tune": (C.c_int, [C.c_void_p, C.c_int32]),
}

_lib = None


def load(path=None):
    """Load libqpad_hip.so and attach the C signatures (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise QpadError(f"{path} is missing: build it with `python -m qpad_amd.build` "
                        "(there is no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load().qpad_last_error()
        raise QpadError(f"libqpad_hip error {rc}: {msg.decode() if msg else ''}")


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "need contiguous float64"
    return a.ctypes.data_as(_dp)


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], "need contiguous float32"
    return a.ctypes.data_as(_fp)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], "need contiguous int32"
    return a.ctypes.data_as(_ip)
