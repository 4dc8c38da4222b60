"""ctypes bindings of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or the timed CPU baseline.  The product
package (cfd_solver_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i, _d = C.c_int, C.c_double


def build(force=False):
    """Compile liboracle.so in place (gcc, seconds)."""
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        L.or_assemble_mom_ref.argtypes = [_i, _i, _d, _d, dp, dp, dp, dp, dp, _d, _d, _d, dp, dp, dp, dp]
        L.or_assemble_p9_ref.argtypes = [_i, _i, _d, _d, dp, dp, dp, dp, dp]
        L.or_rhs_ref.argtypes = [_i, _i, _d, _d] + [dp] * 12 + [dp, dp] + [_d] * 4 + [dp] * 7
        L.or_gauss_ref.argtypes = [_i, _i, dp, _i, dp, dp, _d, _i, C.POINTER(_d)]
        L.or_gauss_ref.restype = _i
        L.or_face_div_ref.argtypes = [_i, _i, _d, _d] + [dp] * 7 + [_d, dp]
        L.or_sip9p_ref.argtypes = [_i, _i, dp, dp, dp, _d, _i, _d, C.POINTER(_d)]
        L.or_sip9p_ref.restype = _i
        L.or_pcor_bc_ref.argtypes = [_i, _i, dp, dp, dp, _d, _d]
        L.or_correct_ref.argtypes = [_i, _i, _d, _d] + [dp] * 7 + [_d, dp, dp, dp]
        L.or_wall_p_ref.argtypes = [_i, _i, _d, _d] + [dp] * 4 + [dp] * 9 + [dp, dp, _d, _d]
        L.or_dil_ref.argtypes = [_i, _i, _d, _d] + [dp] * 6 + [dp]
        L.or_dil_ref.restype = _d
        L.or_nusselt_ref.argtypes = [_i, _i, _d, _d, dp, dp, dp, dp, _d, C.c_void_p]
        L.or_nusselt_ref.restype = _d
        L.or_cav_create.argtypes = [_i, _i, _d, _d] + [dp] * 12 + [C.c_void_p]
        L.or_cav_create.restype = C.c_void_p
        L.or_cav_destroy.argtypes = [C.c_void_p]
        L.or_cav_set_fields.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.or_cav_get_fields.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.or_cav_apply_bcs.argtypes = [C.c_void_p]
        L.or_cav_outer.argtypes = [C.c_void_p, C.c_void_p]
        L.or_cav_outer.restype = _i
        L.or_cav_lo_diag.argtypes = [C.c_void_p, dp]
        L.or_cav_p_apply.argtypes = [C.c_void_p, dp, dp, dp]
        L.or_cav_p_solve.argtypes = [C.c_void_p, dp, dp, dp, _d, _i, C.POINTER(_d)]
        L.or_cav_p_solve.restype = _i
        L.or_cav_residuals.argtypes = [C.c_void_p, dp, dp, dp, dp]
        _LIB = L
    return _LIB


class CavParams(C.Structure):
    _fields_ = [("Re", _d), ("Ri", _d), ("Pr", _d), ("dt", _d), ("alpha_u", _d),
                ("alpha_p", _d), ("alpha_t", _d), ("lid_u", _d), ("t_hot", _d),
                ("t_cold", _d), ("mom_sweeps", _i), ("p_rtol", _d), ("p_maxit", _i),
                ("mg_levels", _i), ("mg_pre", _i), ("mg_post", _i), ("mg_coarse", _i),
                ("mg_omega", _d)]


class CavStats(C.Structure):
    _fields_ = [("mass_l1", _d), ("energy_l1", _d), ("mom_l1", _d), ("dmax", _d),
                ("nu_hot", _d), ("nu_cold", _d), ("p_iters", _i), ("p_relres", _d)]


def cav_params(**kw):
    base = dict(Re=100.0, Ri=1.0, Pr=0.71, dt=1.0e4, alpha_u=0.7, alpha_p=0.3, alpha_t=1.0,
                lid_u=1.0, t_hot=1.0, t_cold=0.0, mom_sweeps=8, p_rtol=1e-2, p_maxit=4,
                mg_levels=12, mg_pre=2, mg_post=2, mg_coarse=10, mg_omega=0.8)
    base.update(kw)
    return CavParams(**base)


class CavityOracle:
    """The cavity SIMPLE iteration on the CPU (FP64, serial)."""

    def __init__(self, grid, params: CavParams):
        self.g = grid
        self.n0, self.n1 = grid.n0, grid.n1
        self.params = params
        a = grid
        self._keep = [np.ascontiguousarray(v, dtype=np.float64) for v in
                      (a.x, a.y, a.xix, a.xiy, a.etx, a.ety, a.alph, a.beta, a.gamma, a.ajac, a.p1, a.q1)]
        self.h = lib().or_cav_create(self.n0, self.n1, a.dxi, a.deta, *self._keep, C.byref(params))
        z = np.zeros((self.n1, self.n0))
        self.set_fields(z, z, z, z)
        lib().or_cav_apply_bcs(self.h)

    def set_fields(self, u, v, T, p):
        f = [np.ascontiguousarray(x, dtype=np.float64) for x in (u, v, T, p)]
        lib().or_cav_set_fields(self.h, *f)

    def apply_bcs(self):
        lib().or_cav_apply_bcs(self.h)

    def fields(self):
        out = [np.zeros((self.n1, self.n0)) for _ in range(4)]
        lib().or_cav_get_fields(self.h, *out)
        return dict(u=out[0], v=out[1], T=out[2], p=out[3])

    def step(self):
        st = CavStats()
        lib().or_cav_outer(self.h, C.byref(st))
        return st

    def residuals(self):
        o = [np.zeros((self.n1, self.n0)) for _ in range(4)]
        lib().or_cav_residuals(self.h, *o)
        return o

    def p_apply(self, d, x):
        y = np.zeros((self.n1, self.n0))
        lib().or_cav_p_apply(self.h, np.ascontiguousarray(d), np.ascontiguousarray(x), y)
        return y

    def p_solve(self, d, q, rtol=1e-10, maxit=200, x0=None):
        x = np.zeros((self.n1, self.n0)) if x0 is None else np.array(x0, dtype=np.float64)
        rr = _d()
        it = lib().or_cav_p_solve(self.h, np.ascontiguousarray(d), np.ascontiguousarray(q), x,
                                  rtol, maxit, C.byref(rr))
        return x, it, rr.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().or_cav_destroy(self.h)
                self.h = None
        except Exception:
            pass
