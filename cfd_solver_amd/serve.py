"""Continuous-batching server of cavity solves (wrapper of the cfd_server_* C ABI)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Params, Stats, call, f64
from .cavity import _slot_params, default_params
from .grid import Grid

STATE = {0: "queued", 1: "running", 2: "done", 3: "notconv", 4: "failed"}


class CavityServer:
    """Queue solves with submit(), advance all running solves one outer iteration
    per step(), collect finished ones with poll()."""

    def __init__(self, params: Params | None = None, hbm_budget: int = 64 << 30, slots_per_batch: int = 16,
                 pack=None, **kw):
        """pack: None (size-class batches under hbm_budget) or (n0, n1, dxi, deta, slots):
        one pack of those slabs (cfd_server_create_packed)."""
        self.params = params if params is not None else default_params(**kw)
        self._h = C.c_void_p()
        if pack is None:
            call("cfd_server_create", C.byref(self.params), C.c_size_t(int(hbm_budget)), int(slots_per_batch),
                 C.byref(self._h))
        else:
            n0, n1, dxi, deta, slots = pack
            arr = (C.c_int * len(slots))(*slots)
            call("cfd_server_create_packed", C.byref(self.params), n0, n1, dxi, deta, len(slots), arr,
                 C.byref(self._h))
        self._shape = {}

    def close(self):
        if self._h and self._h.value:
            _lib.load().cfd_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, grid: Grid, tol: float = 1e-8, max_outer: int = 100000, **phys) -> int:
        g = grid
        arrs = [f64(a) for a in (g.x, g.y, g.xix, g.xiy, g.etx, g.ety, g.alph, g.beta, g.gamma, g.ajac,
                                 g.p1, g.q1)]
        sp = _slot_params(self.params, **phys)
        t = C.c_long()
        call("cfd_server_submit", self._h, g.n0, g.n1, g.dxi, g.deta, *arrs[:10], arrs[10].ctypes.data,
             arrs[11].ctypes.data, C.byref(sp), float(tol), int(max_outer), C.byref(t))
        self._shape[t.value] = (g.n1, g.n0)
        return t.value

    def step(self) -> tuple[int, int]:
        run, q = C.c_int(), C.c_int()
        call("cfd_server_step", self._h, C.byref(run), C.byref(q))
        return run.value, q.value

    def poll(self, ticket: int, fields: bool = False):
        """None while queued/running; else dict(state, stats, latency, [fields])."""
        state = C.c_int()
        st = Stats()
        lat = C.c_double()
        shape = self._shape[ticket]
        out = [np.zeros(shape) for _ in range(4)] if fields else [None] * 4
        ptrs = [o.ctypes.data if o is not None else None for o in out]
        call("cfd_server_poll", self._h, C.c_long(ticket), C.byref(state), C.byref(st), C.byref(lat), *ptrs)
        if state.value in (0, 1):
            return None
        del self._shape[ticket]
        res = dict(state=STATE[state.value], stats=st, latency=lat.value)
        if fields:
            res["fields"] = dict(u=out[0], v=out[1], T=out[2], p=out[3])
        return res

    def info(self):
        nb, by, fin = C.c_int(), C.c_size_t(), C.c_long()
        call("cfd_server_info", self._h, C.byref(nb), C.byref(by), C.byref(fin))
        return dict(batches=nb.value, bytes=by.value, finished=fin.value)
