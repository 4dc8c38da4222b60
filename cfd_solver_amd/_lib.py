"""ctypes binding of libcfd_gpu.so (the gfx950 HIP library, C ABI in include/cfd_gpu.h).

There is no CPU fallback: if the library is missing or cannot be loaded the
import of any stage raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcfd_gpu.so")

CFD_OK = 0
ERRORS = {-1: "CFD_E_ARG", -2: "CFD_E_OOM", -3: "CFD_E_NAN", -4: "CFD_E_NOTCONV",
          -5: "CFD_E_HIP", -6: "CFD_E_STATE", -7: "CFD_E_FULL"}

CFD_FP64 = 0
CFD_MIXED = 1


class CfdError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__(f"{fn} returned {ERRORS.get(code, code)}")
        self.code = code


dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i, _d, _p = C.c_int, C.c_double, C.c_void_p


class Params(C.Structure):
    _fields_ = [("Re", _d), ("Ri", _d), ("Pr", _d), ("dt", _d), ("alpha_u", _d), ("alpha_p", _d),
                ("alpha_t", _d), ("lid_u", _d), ("t_hot", _d), ("t_cold", _d), ("mom_sweeps", _i),
                ("p_rtol", _d), ("p_maxit", _i), ("mg_levels", _i), ("mg_pre", _i), ("mg_post", _i),
                ("mg_coarse", _i), ("mg_omega", _d), ("precision", _i), ("graph", _i), ("mg_tune", _i),
                ("wall_upwind", _i), ("hyb_blend", _d), ("fas_levels", _i), ("fas_pre", _i), ("fas_post", _i),
                ("fas_coarse", _i), ("fas_min_n", _i), ("fas_coarse_scheme", _i),
                ("fas_small", _i), ("fas_coarse_pmaxit", _i)]


class Stats(C.Structure):
    _fields_ = [("mass_l1", _d), ("energy_l1", _d), ("mom_l1", _d), ("dmax", _d), ("nu_hot", _d),
                ("nu_cold", _d), ("p_iters", _i), ("p_relres", _d), ("mass_res", _d),
                ("energy_res", _d), ("iter", _i)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RefParams(C.Structure):
    _fields_ = [("Re", _d), ("Pr", _d), ("Ri", _d), ("dt", _d), ("uinf", _d), ("vinf", _d),
                ("speed_amp", _d), ("accn_amp", _d), ("F", _d), ("time", _d), ("mom_rtol", _d),
                ("p_rtol", _d), ("mom_maxit", _i), ("p_maxit", _i), ("time_accurate", _i),
                ("alpha_deg", _d), ("ar", _d)]


class RefStats(C.Structure):
    _fields_ = [("dmax", _d), ("circ", _d), ("iters", _i * 6), ("relres", _d * 6), ("cl", _d), ("cd", _d),
                ("cm", _d), ("nu", _d), ("cl_pr", _d), ("cd_pr", _d), ("cl_vor", _d), ("cd_vor", _d)]


# every symbol include/cfd_gpu.h declares, with its ctypes signature
SIGNATURES = {
    "cfd_version": ([], C.c_char_p),
    "cfd_device_count": ([C.POINTER(_i)], _i),
    "cfd_ref_assemble_mom": ([_i, _i, _d, _d, dp, dp, dp, dp, dp, _d, _d, _d, dp, dp, dp, dp], _i),
    "cfd_ref_assemble_p9": ([_i, _i, _d, _d, dp, dp, dp, dp, dp], _i),
    "cfd_ref_rhs": ([_i, _i, _d, _d] + [dp] * 14 + [_d] * 4 + [dp] * 7, _i),
    "cfd_gauss": ([_i, _i] + [dp] * 25 + [dp, dp, _d, _i, C.POINTER(_i), C.POINTER(_d)], _i),
    "cfd_ref_face_div": ([_i, _i, _d, _d] + [dp] * 7 + [_d, dp], _i),
    "cfd_sip9p": ([_i, _i] + [dp] * 9 + [dp, dp, _d, _i, C.POINTER(_i), C.POINTER(_d)], _i),
    "cfd_sip9p_noalloc": ([_i, _i] + [dp] * 9 + [dp, dp, _d, _i, C.POINTER(_i), C.POINTER(_d)], _i),
    "cfd_ref_pcor_bc": ([_i, _i, dp, dp, dp, _d, _d], _i),
    "cfd_ref_correct": ([_i, _i, _d, _d] + [dp] * 7 + [_d, dp, dp, dp], _i),
    "cfd_ref_outflow_up": ([_i, _i, dp, dp, dp, dp, dp, dp, _d, _d], _i),
    "cfd_ref_farfield": ([_i, _i, _d, dp, dp] + [dp] * 5 + [dp] * 7 + [_d] * 6 + [_p], _i),
    "cfd_ref_exit_pressure": ([_i, _i, _d, _d] + [dp] * 6 + [dp] * 9 + [dp, dp] + [_d] * 5, _i),
    "cfd_ref_wall_pressure": ([_i, _i, _d, _d] + [dp] * 4 + [dp] * 9 + [_p, _p, _d, _d], _i),
    "cfd_ref_dilation": ([_i, _i, _d, _d] + [dp] * 6 + [dp, C.POINTER(_d)], _i),
    "cfd_ref_nusselt": ([_i, _i, _d, _d, dp, dp, dp, dp, _d, C.POINTER(_d), _p], _i),
    "cfd_ref_case_create": ([_i, _i, _d, _d] + [dp] * 10 + [_p, _p, dp, dp, C.POINTER(RefParams),
                            C.POINTER(_p)], _i),
    "cfd_ref_case_set_state": ([_p] + [dp] * 6, _i),
    "cfd_ref_case_get_state": ([_p] + [dp] * 6, _i),
    "cfd_ref_case_step": ([_p, _i, C.POINTER(RefStats)], _i),
    "cfd_ref_case_destroy": ([_p], None),
    "cfd_ref_case_time": ([_p, C.POINTER(_d)], _i),
    "cfd_ref_case_get_post": ([_p, dp, dp], _i),
    "cfd_ref_stream_function": ([_i, _i, _d] + [dp] * 5 + [dp], _i),
    "cfd_ref_vorticity": ([_i, _i, _d, _d] + [dp] * 6 + [dp], _i),
    "cfd_ref_forces": ([_i, _i, _d, _d] + [dp] * 8 + [_d, _d, _d, dp], _i),
    "cfd_params_default": ([C.POINTER(Params)], None),
    "cfd_fas_coarsest": ([_i, _i, C.POINTER(Params)], _i),
    "cfd_case_create": ([_i, _i, _d, _d] + [dp] * 10 + [_p, _p, C.POINTER(Params), C.POINTER(_p)], _i),
    "cfd_case_destroy": ([_p], _i),
    "cfd_case_set_fields": ([_p, dp, dp, dp, dp], _i),
    "cfd_case_get_fields": ([_p, dp, dp, dp, dp], _i),
    "cfd_case_apply_bcs": ([_p], _i),
    "cfd_case_outer": ([_p, _i, C.POINTER(Stats)], _i),
    "cfd_case_solve": ([_p, _d, _i, C.POINTER(Stats)], _i),
    "cfd_case_bytes": ([_p, C.POINTER(C.c_size_t)], _i),
    "cfd_case_residuals": ([_p, dp, dp, dp, dp], _i),
    "cfd_case_p_apply": ([_p, dp, dp, dp], _i),
    "cfd_case_p_solve": ([_p, _p, _p, _p, _d, _i, C.POINTER(_i), C.POINTER(_d)], _i),
    "cfd_case_time_kernel": ([_p, _i, _i, C.POINTER(_d)], _i),
    "cfd_case_profile": ([_p, _i], _i),
    "cfd_case_profile_read": ([_p, C.POINTER(_d), C.POINTER(C.c_long), _i], _i),
    "cfd_batch_create": ([_i, _i, _d, _d, _i, C.POINTER(Params), C.POINTER(_p)], _i),
    "cfd_batch_destroy": ([_p], _i),
    "cfd_batch_set_case": ([_p, _i] + [dp] * 10 + [_p, _p, C.POINTER(Params)], _i),
    "cfd_batch_release": ([_p, _i], _i),
    "cfd_batch_outer": ([_p, _i, _p], _i),
    "cfd_batch_launch": ([_p], _i),
    "cfd_batch_finish": ([_p, _p], _i),
    "cfd_batch_get_fields": ([_p, _i, dp, dp, dp, dp], _i),
    "cfd_batch_set_fields": ([_p, _i, dp, dp, dp, dp], _i),
    "cfd_batch_bytes": ([_p, C.POINTER(C.c_size_t)], _i),
    "cfd_batch_bytes_needed": ([_i, _i, _d, _d, _i, C.POINTER(Params), C.POINTER(C.c_size_t)], _i),
    "cfd_pack_create": ([_i, _i, _d, _d, _i, C.POINTER(_i), C.POINTER(Params), C.POINTER(_p)], _i),
    "cfd_pack_destroy": ([_p], _i),
    "cfd_pack_bytes_needed": ([_i, _i, _d, _d, _i, C.POINTER(_i), C.POINTER(Params), C.POINTER(C.c_size_t)], _i),
    "cfd_pack_bytes": ([_p, C.POINTER(C.c_size_t)], _i),
    "cfd_pack_capacity": ([_p, C.POINTER(_i)], _i),
    "cfd_pack_add_case": ([_p, _i, _i] + [dp] * 10 + [_p, _p, C.POINTER(Params), C.POINTER(_i)], _i),
    "cfd_pack_release": ([_p, _i], _i),
    "cfd_pack_outer": ([_p, _i, _p], _i),
    "cfd_pack_launch": ([_p], _i),
    "cfd_pack_finish": ([_p, _p], _i),
    "cfd_pack_get_fields": ([_p, _i, dp, dp, dp, dp], _i),
    "cfd_server_create": ([C.POINTER(Params), C.c_size_t, _i, C.POINTER(_p)], _i),
    "cfd_server_create_packed": ([C.POINTER(Params), _i, _i, _d, _d, _i, C.POINTER(_i), C.POINTER(_p)], _i),
    "cfd_server_destroy": ([_p], _i),
    "cfd_server_submit": ([_p, _i, _i, _d, _d] + [dp] * 10 + [_p, _p, C.POINTER(Params), _d, _i,
                          C.POINTER(C.c_long)], _i),
    "cfd_server_step": ([_p, C.POINTER(_i), C.POINTER(_i)], _i),
    "cfd_server_poll": ([_p, C.c_long, C.POINTER(_i), C.POINTER(Stats), C.POINTER(_d), _p, _p, _p, _p], _i),
    "cfd_server_info": ([_p, C.POINTER(_i), C.POINTER(C.c_size_t), C.POINTER(C.c_long)], _i),
}

_LIB = None


def load():
    """Load libcfd_gpu.so (built in place by cfd_solver_amd.build); raise if absent."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m cfd_solver_amd.build` "
                               "(the HIP library is required; there is no CPU fallback)")
        lib = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = lib
    return _LIB


def check(fn, code):
    if code != CFD_OK:
        raise CfdError(fn, code)


def call(name, *args):
    code = getattr(load(), name)(*args)
    check(name, code)
    return code


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
