"""Device-resident steady mixed-convection cavity (SIMPLE outer iteration on gfx950)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Params, Stats, call, f64
from .grid import Grid, cavity_grid


def fas_hierarchy(n0: int, n1: int, params: Params) -> list[tuple[int, int, int]]:
    """[(n0, n1, SIMPLE iterations per outer iteration)] of every level of the FAS
    hierarchy a case of n0 x n1 nodes gets (one level without coarsening: one SIMPLE
    iteration per outer iteration)."""
    coarsest = _lib.load().cfd_fas_coarsest(n0, n1, C.byref(params))
    levels = [(n0, n1)]
    while min(levels[-1]) > coarsest:
        a, b = levels[-1]
        levels.append(((a + 1) // 2, (b + 1) // 2))
    if len(levels) == 1:
        return [(n0, n1, 1)]
    its = [params.fas_pre + params.fas_post] * (len(levels) - 1) + [params.fas_coarse]
    return [(a, b, k) for (a, b), k in zip(levels, its)]


def cells_per_outer(n0: int, n1: int, params: Params) -> int:
    """Interior nodes x SIMPLE iterations of every level in one outer iteration."""
    return sum((a - 2) * (b - 2) * k for a, b, k in fas_hierarchy(n0, n1, params))


def default_params(**kw) -> Params:
    p = Params()
    _lib.load().cfd_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class CavityCase:
    """One cavity case held in HBM; fields start at rest with the wall BCs applied."""

    def __init__(self, grid: Grid, params: Params | None = None, **kw):
        self.grid = grid
        self.params = params if params is not None else default_params(**kw)
        self.n0, self.n1 = grid.n0, grid.n1
        g = grid
        self._h = C.c_void_p()
        self._keep = [f64(a) for a in (g.x, g.y, g.xix, g.xiy, g.etx, g.ety, g.alph, g.beta, g.gamma,
                                       g.ajac, g.p1, g.q1)]
        call("cfd_case_create", self.n0, self.n1, g.dxi, g.deta, *self._keep[:10],
             self._keep[10].ctypes.data, self._keep[11].ctypes.data, C.byref(self.params),
             C.byref(self._h))

    @classmethod
    def from_config(cls, n, skew=0.05, beta=None, **kw):
        return cls(cavity_grid(n, skew=skew, beta=beta), **kw)

    def close(self):
        if self._h and self._h.value:
            _lib.load().cfd_case_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fields(self, u, v, T, p):
        call("cfd_case_set_fields", self._h, f64(u), f64(v), f64(T), f64(p))

    def fields(self):
        out = [np.zeros((self.n1, self.n0)) for _ in range(4)]
        call("cfd_case_get_fields", self._h, *out)
        return dict(u=out[0], v=out[1], T=out[2], p=out[3])

    def apply_bcs(self):
        call("cfd_case_apply_bcs", self._h)

    def step(self, n=1) -> Stats:
        st = Stats()
        call("cfd_case_outer", self._h, int(n), C.byref(st))
        return st

    def solve(self, tol=1e-8, max_outer=100000) -> tuple[Stats, bool]:
        st = Stats()
        code = _lib.load().cfd_case_solve(self._h, tol, int(max_outer), C.byref(st))
        if code not in (0, -4):
            _lib.check("cfd_case_solve", code)
        return st, code == 0

    def nbytes(self):
        b = C.c_size_t()
        call("cfd_case_bytes", self._h, C.byref(b))
        return b.value

    def residuals(self):
        o = [np.zeros((self.n1, self.n0)) for _ in range(4)]
        call("cfd_case_residuals", self._h, *o)
        return o

    def p_apply(self, d, x):
        y = np.zeros((self.n1, self.n0))
        call("cfd_case_p_apply", self._h, f64(d), f64(x), y)
        return y

    def p_solve(self, d, q, rtol=1e-10, maxit=200, want_x=True):
        """Solve L x = q; d or q None keeps the device-resident values (d from the
        last residual evaluation); want_x False leaves x on the device."""
        x = np.zeros((self.n1, self.n0)) if want_x else None
        it, rr = C.c_int(), C.c_double()
        keep = [f64(a) if a is not None else None for a in (d, q)]
        call("cfd_case_p_solve", self._h, None if d is None else keep[0].ctypes.data,
             None if q is None else keep[1].ctypes.data, None if x is None else x.ctypes.data, rtol, maxit,
             C.byref(it), C.byref(rr))
        return x, it.value, rr.value

    KERNELS = ("p_apply", "jacobi", "residual", "face_div", "mg_down", "mg_up")

    def profile(self, enable=True):
        """True: time hot kernels live (recaptures the graph); False: stop; 2: reset counters."""
        mode = 2 if enable == 2 and enable is not True else (1 if enable else 0)
        call("cfd_case_profile", self._h, mode)

    def profile_read(self):
        n = len(self.KERNELS)
        ms = (C.c_double * n)()
        cnt = (C.c_long * n)()
        call("cfd_case_profile_read", self._h, ms, cnt, n)
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.KERNELS)}

    def time_kernel(self, kernel, reps=20):
        ms = C.c_double()
        call("cfd_case_time_kernel", self._h, int(kernel), int(reps), C.byref(ms))
        return ms.value


def _slot_params(base: Params, **phys) -> Params:
    p = Params()
    C.memmove(C.byref(p), C.byref(base), C.sizeof(Params))
    for k, v in phys.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class CavityBatch:
    """`slots` cavity cases of one grid shape advanced together on one stream.

    Batch-wide settings (predictor sweeps, p' solver, multigrid, precision,
    graph) come from `params`; each slot's physics from the keywords of
    `set_case` (Re, Ri, Pr, dt, alpha_u, alpha_p, alpha_t, lid_u, t_hot, t_cold).
    """

    def __init__(self, n0: int, n1: int, dxi: float, deta: float, slots: int, params: Params | None = None, **kw):
        self.n0, self.n1, self.slots = n0, n1, slots
        self.params = params if params is not None else default_params(**kw)
        self._h = C.c_void_p()
        call("cfd_batch_create", n0, n1, dxi, deta, slots, C.byref(self.params), C.byref(self._h))
        self._stats = (Stats * slots)()

    def close(self):
        if self._h and self._h.value:
            _lib.load().cfd_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_case(self, slot: int, grid: Grid, **phys):
        if (grid.n0, grid.n1) != (self.n0, self.n1):
            raise ValueError(f"grid {grid.n0}x{grid.n1} does not fit a {self.n0}x{self.n1} batch")
        g = grid
        arrs = [f64(a) for a in (g.x, g.y, g.xix, g.xiy, g.etx, g.ety, g.alph, g.beta, g.gamma, g.ajac,
                                 g.p1, g.q1)]
        sp = _slot_params(self.params, **phys)
        call("cfd_batch_set_case", self._h, int(slot), *arrs[:10], arrs[10].ctypes.data, arrs[11].ctypes.data,
             C.byref(sp))

    def release(self, slot: int):
        call("cfd_batch_release", self._h, int(slot))

    def step(self, n: int = 1):
        """n outer iterations of every active slot; returns the per-slot Stats array."""
        call("cfd_batch_outer", self._h, int(n), C.cast(self._stats, C.c_void_p))
        return self._stats

    def launch(self):
        call("cfd_batch_launch", self._h)

    def finish(self):
        call("cfd_batch_finish", self._h, C.cast(self._stats, C.c_void_p))
        return self._stats

    def fields(self, slot: int):
        out = [np.zeros((self.n1, self.n0)) for _ in range(4)]
        call("cfd_batch_get_fields", self._h, int(slot), *out)
        return dict(u=out[0], v=out[1], T=out[2], p=out[3])

    def set_fields(self, slot: int, u, v, T, p):
        call("cfd_batch_set_fields", self._h, int(slot), f64(u), f64(v), f64(T), f64(p))

    def nbytes(self):
        b = C.c_size_t()
        call("cfd_batch_bytes", self._h, C.byref(b))
        return b.value


class CavityPack:
    """Cavity cases of uneven size sharing every launch (cfd_pack_*).

    Slab k holds slots[k] grids of ((n0 - 1) >> k) + 1 x ((n1 - 1) >> k) + 1 nodes; a
    case of slab k's size takes a slot there and one in each slab down to its
    coarsest multigrid level.  Pack-wide settings come from `params`; each case's
    physics from the keywords of `add_case`.
    """

    def __init__(self, n0: int, n1: int, dxi: float, deta: float, slots, params: Params | None = None, **kw):
        self.n0, self.n1 = n0, n1
        self.params = params if params is not None else default_params(**kw)
        self.slots = list(slots)
        arr = (C.c_int * len(self.slots))(*self.slots)
        self._h = C.c_void_p()
        call("cfd_pack_create", n0, n1, dxi, deta, len(self.slots), arr, C.byref(self.params), C.byref(self._h))
        cap = C.c_int()
        call("cfd_pack_capacity", self._h, C.byref(cap))
        self.capacity = cap.value
        self._stats = (Stats * self.capacity)()
        self._shape = {}

    @staticmethod
    def bytes_needed(n0, n1, dxi, deta, slots, params: Params) -> int:
        arr = (C.c_int * len(slots))(*slots)
        b = C.c_size_t()
        call("cfd_pack_bytes_needed", n0, n1, dxi, deta, len(slots), arr, C.byref(params), C.byref(b))
        return b.value

    def close(self):
        if self._h and self._h.value:
            _lib.load().cfd_pack_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_case(self, grid: Grid, **phys) -> int:
        """Load a case; returns its id (raises CfdError CFD_E_FULL when its slabs are full)."""
        g = grid
        arrs = [f64(a) for a in (g.x, g.y, g.xix, g.xiy, g.etx, g.ety, g.alph, g.beta, g.gamma, g.ajac,
                                 g.p1, g.q1)]
        sp = _slot_params(self.params, **phys)
        cid = C.c_int()
        call("cfd_pack_add_case", self._h, g.n0, g.n1, *arrs[:10], arrs[10].ctypes.data, arrs[11].ctypes.data,
             C.byref(sp), C.byref(cid))
        self._shape[cid.value] = (g.n0, g.n1)
        return cid.value

    def release(self, cid: int):
        call("cfd_pack_release", self._h, int(cid))

    def step(self, n: int = 1):
        """n outer iterations of every case; returns the Stats array indexed by case id."""
        call("cfd_pack_outer", self._h, int(n), C.cast(self._stats, C.c_void_p))
        return self._stats

    def launch(self):
        call("cfd_pack_launch", self._h)

    def finish(self):
        call("cfd_pack_finish", self._h, C.cast(self._stats, C.c_void_p))
        return self._stats

    def fields(self, cid: int):
        n0, n1 = self._shape[cid]
        out = [np.zeros((n1, n0)) for _ in range(4)]
        call("cfd_pack_get_fields", self._h, int(cid), *out)
        return dict(u=out[0], v=out[1], T=out[2], p=out[3])

    def nbytes(self):
        b = C.c_size_t()
        call("cfd_pack_bytes", self._h, C.byref(b))
        return b.value
