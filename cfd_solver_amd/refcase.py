"""The reference solver's O-grid time step, resident on the GPU (cfd_ref_case_*).

`RefCase` holds one O-grid case in HBM: the grid and metric arrays of INP.DAT
(solver.cpp:516-553), the operators the reference's constructor assembles
(solver.cpp:751-1039) and the state u, v, T, p, up, vp.  `step()` runs
solver.cpp:1057-1852 -- every stage of the reference's time step in its order --
with the Gauss-Seidel and SIP solves replaced by BiCGStab to the tolerances in
the parameters (same linear systems, same fixed points).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import RefParams, RefStats, call, f64


def ref_params(**kw) -> RefParams:
    """The reference's compiled-in constants (solver.cpp:159-173) plus solver tolerances.
    Given F alone, the oscillation amplitudes follow the reference's definitions
    (solver.cpp:163-164): speed_amp = (pi/12) 2 pi F, accn_amp = 2 pi F speed_amp."""
    a = 82.0 * math.pi / 180.0
    base = dict(Re=1000.0, Pr=0.71, Ri=0.0, dt=1e-4, uinf=math.sin(a), vinf=math.cos(a), speed_amp=0.0,
                accn_amp=0.0, F=0.0, time=0.0, mom_rtol=1e-12, p_rtol=1e-12, mom_maxit=500, p_maxit=2000,
                time_accurate=0, alpha_deg=82.0, ar=1.0)
    F = kw.get("F", 0.0)
    if F and "speed_amp" not in kw:
        pi = math.acos(-1.0)
        base["speed_amp"] = pi / 12.0 * 2.0 * pi * F
        base["accn_amp"] = 2.0 * pi * F * base["speed_amp"]
    base.update(kw)
    return RefParams(**base)


class RefCase:
    def __init__(self, grid, params: RefParams | None = None, **kw):
        """grid: an O-grid with n0, n1, dxi, deta, x, y, xix, xiy, etx, ety, alph, beta,
        gamma, ajac, xnox, xnoy (and optionally p1, q1) -- e.g. grid.ogrid_ellipse()."""
        _lib.load()
        self.n0, self.n1 = int(grid.n0), int(grid.n1)
        self.params = params if params is not None else ref_params(**kw)
        g = grid
        self._keep = [f64(a) for a in (g.x, g.y, g.xix, g.xiy, g.etx, g.ety, g.alph, g.beta, g.gamma, g.ajac)]
        opt = [f64(getattr(g, k)) if getattr(g, k, None) is not None else None for k in ("p1", "q1")]
        self._keep += [a for a in opt if a is not None]
        nox, noy = f64(g.xnox), f64(g.xnoy)
        self._keep += [nox, noy]
        h = C.c_void_p()
        call("cfd_ref_case_create", self.n0, self.n1, float(g.dxi), float(g.deta), *self._keep[:10],
             opt[0].ctypes.data if opt[0] is not None else None, opt[1].ctypes.data if opt[1] is not None else None,
             nox, noy, C.byref(self.params), C.byref(h))
        self.h = h
        self.grid = grid

    def set_state(self, u, v, T, p, up, vp):
        call("cfd_ref_case_set_state", self.h, *[f64(a) for a in (u, v, T, p, up, vp)])

    def state(self):
        out = [np.zeros((self.n1, self.n0)) for _ in range(6)]
        call("cfd_ref_case_get_state", self.h, *out)
        return dict(zip(("u", "v", "T", "p", "up", "vp"), out))

    def step(self, nsteps=1):
        st = RefStats()
        call("cfd_ref_case_step", self.h, int(nsteps), C.byref(st))
        self.last = dict(dmax=st.dmax, circ=st.circ, iters=list(st.iters), relres=list(st.relres), cl=st.cl,
                         cd=st.cd, cm=st.cm, nu=st.nu, cl_pr=st.cl_pr, cd_pr=st.cd_pr, cl_vor=st.cl_vor,
                         cd_vor=st.cd_vor)
        return self.last

    def post(self):
        """Stream function and vorticity of the last step."""
        si, vo = np.zeros((self.n1, self.n0)), np.zeros((self.n1, self.n0))
        call("cfd_ref_case_get_post", self.h, si, vo)
        return si, vo

    def write_outputs(self, directory, loop, time=None):
        """The reference's output files of the last step (solver.cpp:1984-2039) in
        `directory`: spt100.dat and SURF_DIST.dat rewritten, the COEFF_HIS lines
        appended.  time: printed in COEFF_HIS (default: the case's time, an int 0
        when it is held fixed, as the reference prints it)."""
        import os
        from . import refio
        s, (si, vo) = self.state(), self.post()
        g, st = self.grid, self.last
        if time is None:
            time = self.time() if self.params.time_accurate else int(self.params.time)
        refio.write_tecplot(os.path.join(directory, "spt100.dat"), g.x, g.y, s["u"], s["v"], s["T"], s["p"], si, vo)
        refio.append_coeff_his(os.path.join(directory, "COEFF_HIS.dat"), time, st["cl"], st["cd"], st["cm"], st["nu"])
        refio.append_coeff_pr_vor(os.path.join(directory, "COEFF_HIS_pr_vor.dat"), time, st["cl_pr"], st["cd_pr"],
                                  st["cl_vor"], st["cd_vor"])
        refio.write_surf_dist(os.path.join(directory, "SURF_DIST.dat"), g.dxi, g.deta, s["p"], vo, s["T"], g.etx,
                              g.ety)

    def save_restart(self, path, loop):
        from . import refio
        s, (si, _) = self.state(), self.post()
        refio.write_restart(path, loop, self.time(), self.last["dmax"] if getattr(self, "last", None) else 0.0,
                            dict(x=self.grid.x, y=self.grid.y, si=si, **s))

    @classmethod
    def from_restart(cls, grid, path, **kw):
        """A case continuing from a restart file: its state, and its time as the
        parameters' time."""
        from . import refio
        r = refio.read_restart(path)
        if (r["n0"], r["n1"]) != (int(grid.n0), int(grid.n1)):
            raise ValueError(f"restart holds a {r['n0']}x{r['n1']} grid")
        c = cls(grid, ref_params(**dict(kw, time=r["time"])))
        c.set_state(r["u"], r["v"], r["T"], r["p"], r["up"], r["vp"])
        return c

    def time(self):
        t = C.c_double()
        call("cfd_ref_case_time", self.h, C.byref(t))
        return t.value

    def close(self):
        if getattr(self, "h", None):
            _lib.load().cfd_ref_case_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def initial_state(grid, params: RefParams):
    """The reference's fields for a fresh start (restart == 0, solver.cpp:612-720): the
    free stream everywhere, T = 0 and p = 0; the wall row moving with the cylinder
    (-a y, a x) at T = 1 (a = speed_amp); the outer row keeps the free stream where
    uinf*xnox + vinf*xnoy >= 0 and copies row n1-2 elsewhere (the same values here)."""
    n1, n0 = np.shape(grid.x)
    u = np.full((n1, n0), params.uinf)
    v = np.full((n1, n0), params.vinf)
    T = np.zeros((n1, n0))
    u[0] = -params.speed_amp * np.asarray(grid.y)[0]
    v[0] = params.speed_amp * np.asarray(grid.x)[0]
    T[0] = 1.0
    return dict(u=u, v=v, T=T, p=np.zeros((n1, n0)), up=u.copy(), vp=v.copy())
