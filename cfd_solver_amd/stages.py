"""Host mirror of the reference solver's hot-path stages, one method per stage.

Each method takes the arrays the reference's `Solver` members hold, in the
contiguous (n1, n0) layout, and runs the corresponding gfx950 kernel through
the C ABI (include/cfd_gpu.h).  Method names follow the reference:

  assemble_mom   solver.cpp:751-939      assemble_p9  solver.cpp:943-1039
  rhs            solver.cpp:1062-1236    gauss        solver.cpp:2279-2445
  face_div       solver.cpp:1389-1448    sip9p        solver.cpp:2066-2277
  pcor_bc        solver.cpp:1474-1497    correct      solver.cpp:1504-1541
  wall_pressure  solver.cpp:1667-1728    dilation     solver.cpp:1806-1852
  nusselt        solver.cpp:1951-1982, 2031-2038
  outflow_up     solver.cpp:1351-1381    farfield     solver.cpp:1543-1655
  exit_pressure  solver.cpp:1731-1782    stream_function solver.cpp:1784-1800
  vorticity      solver.cpp:1802-1887    forces       solver.cpp:1891-1982

Like the reference, `sip9p` and `gauss` update `phi` in place and take the
right-hand side `q`; unlike it they return (iterations, relative residual)
and stop at a 2-norm tolerance relative to the starting residual.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, f64

# order of the 25 planes of a momentum operator: plane k = (dj+2)*5 + (di+2)
K25 = {(di, dj): (dj + 2) * 5 + (di + 2) for dj in range(-2, 3) for di in range(-2, 3)}
K9 = {(di, dj): (dj + 1) * 3 + (di + 1) for dj in range(-1, 2) for di in range(-1, 2)}
# the reference's names of the planes (aup, aue, ... / ap, ae, ...)
NAMES25 = {"p": (0, 0), "e": (1, 0), "w": (-1, 0), "n": (0, 1), "s": (0, -1), "ne": (1, 1),
           "nw": (-1, 1), "se": (1, -1), "sw": (-1, -1), "ee": (2, 0), "ww": (-2, 0), "nn": (0, 2),
           "ss": (0, -2), "nnee": (2, 2), "nnww": (-2, 2), "ssee": (2, -2), "ssww": (-2, -2),
           "nne": (1, 2), "nnw": (-1, 2), "sse": (1, -2), "ssw": (-1, -2), "nee": (2, 1),
           "nww": (-2, 1), "see": (2, -1), "sww": (-2, -1)}
# argument order of Solver::gauss (solver.cpp:2279-2285) and Solver::sip9p (2066-2068)
GAUSS_ORDER = ["p", "e", "s", "n", "w", "se", "sw", "ne", "nw", "ss", "ssee", "ssww", "sse", "ssw",
               "see", "sww", "nn", "nnee", "nnww", "nne", "nnw", "nee", "nww", "ee", "ww"]
SIP_ORDER = ["p", "e", "s", "n", "w", "se", "sw", "ne", "nw"]


class RefStages:
    """Stage entry points for one O-grid of n0 x n1 nodes with spacings (dxi, deta)."""

    def __init__(self, n0, n1, dxi, deta):
        _lib.load()
        self.n0, self.n1, self.dxi, self.deta = int(n0), int(n1), float(dxi), float(deta)

    def _z(self, *shape):
        return np.zeros(shape if shape else (self.n1, self.n0))

    def assemble_mom(self, alph, beta, gamma, p1, q1, dt, Re, Pr):
        AU, AT = self._z(25, self.n1, self.n0), self._z(25, self.n1, self.n0)
        BU, BT = self._z(6, self.n0), self._z(6, self.n0)
        call("cfd_ref_assemble_mom", self.n0, self.n1, self.dxi, self.deta, f64(alph), f64(beta),
             f64(gamma), f64(p1), f64(q1), dt, Re, Pr, AU, AT, BU, BT)
        return AU, AT, BU, BT

    def assemble_p9(self, xix, xiy, etx, ety):
        A9 = self._z(9, self.n1, self.n0)
        call("cfd_ref_assemble_p9", self.n0, self.n1, self.dxi, self.deta, f64(xix), f64(xiy),
             f64(etx), f64(ety), A9)
        return A9

    def rhs(self, u, v, T, up, vp, p, xix, xiy, etx, ety, alph, gamma, BU, BT, Re, Pr, Ri, dt):
        out = [self._z() for _ in range(7)]
        call("cfd_ref_rhs", self.n0, self.n1, self.dxi, self.deta, f64(u), f64(v), f64(T), f64(up),
             f64(vp), f64(p), f64(xix), f64(xiy), f64(etx), f64(ety), f64(alph), f64(gamma),
             f64(BU), f64(BT), Re, Pr, Ri, dt, *out)
        return dict(zip(["uxi", "uet", "qu", "qv", "qt", "qup", "qvp"], out))

    def gauss(self, A25, phi, q, rtol=1e-12, maxit=500, nw_for_nnw=False):
        """A25: (25, n1, n0) planes; nw_for_nnw reproduces the temperature call of
        solver.cpp:1289 (the (-1,+1) plane passed in the (-1,+2) slot)."""
        planes = {nm: f64(A25[K25[o]]) for nm, o in NAMES25.items()}
        if nw_for_nnw:
            planes["nnw"] = planes["nw"]
        it, rr = C.c_int(), C.c_double()
        phi_ = f64(phi).copy()
        call("cfd_gauss", self.n0, self.n1, *[planes[nm] for nm in GAUSS_ORDER], phi_, f64(q),
             rtol, maxit, C.byref(it), C.byref(rr))
        phi[...] = phi_
        return it.value, rr.value

    def sip9p(self, A9, phi, q, rtol=1e-12, maxit=500):
        names9 = {"p": (0, 0), "e": (1, 0), "w": (-1, 0), "n": (0, 1), "s": (0, -1), "ne": (1, 1),
                  "nw": (-1, 1), "se": (1, -1), "sw": (-1, -1)}
        planes = {nm: f64(A9[K9[o]]) for nm, o in names9.items()}
        it, rr = C.c_int(), C.c_double()
        phi_ = f64(phi).copy()
        call("cfd_sip9p", self.n0, self.n1, *[planes[nm] for nm in SIP_ORDER], phi_, f64(q), rtol,
             maxit, C.byref(it), C.byref(rr))
        phi[...] = phi_
        return it.value, rr.value

    def face_div(self, up, vp, p, xix, xiy, etx, ety, dt):
        q = self._z()
        call("cfd_ref_face_div", self.n0, self.n1, self.dxi, self.deta, f64(up), f64(vp), f64(p),
             f64(xix), f64(xiy), f64(etx), f64(ety), dt, q)
        return q

    def pcor_bc(self, pcor, xnox, xnoy, uinf, vinf):
        pc = f64(pcor).copy()
        call("cfd_ref_pcor_bc", self.n0, self.n1, pc, f64(xnox), f64(xnoy), uinf, vinf)
        return pc

    def correct(self, us, vs, pcor, xix, xiy, etx, ety, dt, u, v, p):
        u_, v_, p_ = f64(u).copy(), f64(v).copy(), f64(p).copy()
        call("cfd_ref_correct", self.n0, self.n1, self.dxi, self.deta, f64(us), f64(vs), f64(pcor),
             f64(xix), f64(xiy), f64(etx), f64(ety), dt, u_, v_, p_)
        return u_, v_, p_

    def wall_pressure(self, u, v, T, p, xix, xiy, etx, ety, alph, beta, gamma, q1, ajac, Re, Ri,
                      alc_u=None, alc_v=None):
        p_ = f64(p).copy()
        au = f64(alc_u) if alc_u is not None else None
        av = f64(alc_v) if alc_v is not None else None
        call("cfd_ref_wall_pressure", self.n0, self.n1, self.dxi, self.deta, f64(u), f64(v), f64(T),
             p_, f64(xix), f64(xiy), f64(etx), f64(ety), f64(alph), f64(beta), f64(gamma), f64(q1),
             f64(ajac), au.ctypes.data if au is not None else None,
             av.ctypes.data if av is not None else None, Re, Ri)
        return p_

    def outflow_up(self, up, vp, u, v, xnox, xnoy, uinf, vinf):
        up_, vp_ = f64(up).copy(), f64(vp).copy()
        call("cfd_ref_outflow_up", self.n0, self.n1, up_, vp_, f64(u), f64(v), f64(xnox), f64(xnoy), uinf, vinf)
        return up_, vp_

    def farfield(self, x, y, u, v, T, up, vp, Told, uet, etx, ety, ajac, xnox, xnoy, uinf, vinf, dt,
                 speed_amp=0.0, F=0.0, time=0.0):
        """Returns (u, v, T, up, vp, circulation)."""
        o = [f64(a).copy() for a in (u, v, T, up, vp)]
        circ = C.c_double()
        call("cfd_ref_farfield", self.n0, self.n1, self.deta, f64(x), f64(y), *o, f64(Told), f64(uet), f64(etx),
             f64(ety), f64(ajac), f64(xnox), f64(xnoy), uinf, vinf, dt, speed_amp, F, time, C.byref(circ))
        return (*o, circ.value)

    def exit_pressure(self, u, v, T, uold, vold, p, xix, xiy, etx, ety, alph, beta, gamma, q1, ajac, xnox, xnoy,
                      uinf, vinf, dt, Re, Ri):
        p_ = f64(p).copy()
        call("cfd_ref_exit_pressure", self.n0, self.n1, self.dxi, self.deta, f64(u), f64(v), f64(T), f64(uold),
             f64(vold), p_, f64(xix), f64(xiy), f64(etx), f64(ety), f64(alph), f64(beta), f64(gamma), f64(q1),
             f64(ajac), f64(xnox), f64(xnoy), uinf, vinf, dt, Re, Ri)
        return p_

    def stream_function(self, u, v, xix, xiy, ajac):
        si = self._z()
        call("cfd_ref_stream_function", self.n0, self.n1, self.deta, f64(u), f64(v), f64(xix), f64(xiy), f64(ajac),
             si)
        return si

    def vorticity(self, u, v, xix, xiy, etx, ety):
        vo = self._z()
        call("cfd_ref_vorticity", self.n0, self.n1, self.dxi, self.deta, f64(u), f64(v), f64(xix), f64(xiy),
             f64(etx), f64(ety), vo)
        return vo

    def forces(self, x, y, p, vort, T, ajac, etx, ety, Re, alpha_deg=82.0, ar=1.0):
        """cl, cd, cm, CL_pr, CD_pr, CL_vor, CD_vor, Nu."""
        out = np.zeros(8)
        call("cfd_ref_forces", self.n0, self.n1, self.dxi, self.deta, f64(x), f64(y), f64(p), f64(vort), f64(T),
             f64(ajac), f64(etx), f64(ety), Re, alpha_deg, ar, out)
        return out

    def dilation(self, u, v, xix, xiy, etx, ety):
        dil = self._z()
        dm = C.c_double()
        call("cfd_ref_dilation", self.n0, self.n1, self.dxi, self.deta, f64(u), f64(v), f64(xix),
             f64(xiy), f64(etx), f64(ety), dil, C.byref(dm))
        return dil, dm.value

    def nusselt(self, T, ajac, etx, ety, ar):
        nu = C.c_double()
        prof = np.zeros(self.n0)
        call("cfd_ref_nusselt", self.n0, self.n1, self.dxi, self.deta, f64(T), f64(ajac), f64(etx),
             f64(ety), ar, C.byref(nu), prof.ctypes.data)
        return nu.value, prof
