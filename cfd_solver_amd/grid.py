"""Body-fitted grids, metric arrays and the INP.DAT text format.

Arrays are numpy float64 of shape (n1, n0): row j (eta index), column i (xi
index), so element (i, j) sits at flat offset j*n0 + i -- the layout of every
device buffer in this package.

The metric set is the one the reference reads from INP.DAT
(solver.cpp:516-546):

  xix = xi_x,  xiy = xi_y,  etx = eta_x,  ety = eta_y
  alph  = xi_x^2 + xi_y^2
  gamma = eta_x^2 + eta_y^2
  beta  = -(xi_x*eta_x + xi_y*eta_y)
  ajac  = x_xi*y_eta - x_eta*y_xi

(conventions fixed by their use at solver.cpp:1113-1114, 1571, 1725, 1795).
The reference forces the Laplacian control functions P = lap(xi) and
Q = lap(eta) to zero (solver.cpp:548-553); `metrics(..., control=True)`
computes them from the grid instead, which the curvilinear diffusion operator
needs to stay consistent on stretched grids.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "Grid",
    "cavity_grid",
    "ogrid_ellipse",
    "metrics",
    "write_inp_dat",
    "read_inp_dat",
    "round_pow2p1",
]


@dataclass
class Grid:
    n0: int
    n1: int
    dxi: float
    deta: float
    periodic_xi: bool
    x: np.ndarray
    y: np.ndarray
    xix: np.ndarray
    xiy: np.ndarray
    etx: np.ndarray
    ety: np.ndarray
    alph: np.ndarray
    beta: np.ndarray
    gamma: np.ndarray
    ajac: np.ndarray
    p1: np.ndarray
    q1: np.ndarray
    # INP.DAT header extras (solver.cpp:517-518) and boundary normals (solver.cpp:544-546)
    p_grid: float = 0.0
    a_grid: float = 0.0
    ar: float = 1.0
    xnix: np.ndarray = field(default=None)
    xniy: np.ndarray = field(default=None)
    xnox: np.ndarray = field(default=None)
    xnoy: np.ndarray = field(default=None)

    def arrays(self):
        return dict(x=self.x, y=self.y, xix=self.xix, xiy=self.xiy, etx=self.etx,
                    ety=self.ety, alph=self.alph, beta=self.beta, gamma=self.gamma,
                    ajac=self.ajac, p1=self.p1, q1=self.q1)


def _d1(f, h, axis, periodic):
    """Second-order first derivative along `axis` (1 = xi, 0 = eta)."""
    f = np.asarray(f, dtype=np.float64)
    out = np.empty_like(f)
    if axis == 1 and periodic:
        # node n0-1 duplicates node 0; the period is n0-1
        core = f[:, :-1]
        out[:, :-1] = (np.roll(core, -1, axis=1) - np.roll(core, 1, axis=1)) / (2.0 * h)
        out[:, -1] = out[:, 0]
        return out
    fs = np.moveaxis(f, axis, 0)
    os_ = np.moveaxis(out, axis, 0)
    os_[1:-1] = (fs[2:] - fs[:-2]) / (2.0 * h)
    os_[0] = (-3.0 * fs[0] + 4.0 * fs[1] - fs[2]) / (2.0 * h)
    os_[-1] = (3.0 * fs[-1] - 4.0 * fs[-2] + fs[-3]) / (2.0 * h)
    return out


def _d2(f, h, axis, periodic):
    """Second-order second derivative (interior only; boundaries one-sided)."""
    f = np.asarray(f, dtype=np.float64)
    out = np.empty_like(f)
    if axis == 1 and periodic:
        core = f[:, :-1]
        out[:, :-1] = (np.roll(core, -1, axis=1) - 2.0 * core + np.roll(core, 1, axis=1)) / (h * h)
        out[:, -1] = out[:, 0]
        return out
    fs = np.moveaxis(f, axis, 0)
    os_ = np.moveaxis(out, axis, 0)
    os_[1:-1] = (fs[2:] - 2.0 * fs[1:-1] + fs[:-2]) / (h * h)
    os_[0] = (2.0 * fs[0] - 5.0 * fs[1] + 4.0 * fs[2] - fs[3]) / (h * h)
    os_[-1] = (2.0 * fs[-1] - 5.0 * fs[-2] + 4.0 * fs[-3] - fs[-4]) / (h * h)
    return out


def metrics(x, y, dxi, deta, periodic_xi=False, control=False):
    """Metric arrays of a grid by second-order finite differences.

    Returns a dict with xix, xiy, etx, ety, alph, beta, gamma, ajac, p1, q1.
    With control=True, p1 = lap(xi) and q1 = lap(eta) from the identity
    lap(x) = lap(y) = 0 written in (xi, eta); otherwise both are zero, as the
    reference sets them (solver.cpp:548-553).
    """
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    xx = _d1(x, dxi, 1, periodic_xi)
    xe = _d1(x, deta, 0, False)
    yx = _d1(y, dxi, 1, periodic_xi)
    ye = _d1(y, deta, 0, False)
    jac = xx * ye - xe * yx
    xix = ye / jac
    xiy = -xe / jac
    etx = -yx / jac
    ety = xx / jac
    alph = xix * xix + xiy * xiy
    gamma = etx * etx + ety * ety
    beta = -(xix * etx + xiy * ety)
    p1 = np.zeros_like(x)
    q1 = np.zeros_like(x)
    if control:
        xxx = _d2(x, dxi, 1, periodic_xi)
        yxx = _d2(y, dxi, 1, periodic_xi)
        xee = _d2(x, deta, 0, False)
        yee = _d2(y, deta, 0, False)
        xxe = _d1(xe, dxi, 1, periodic_xi)
        yxe = _d1(ye, dxi, 1, periodic_xi)
        dx = alph * xxx - 2.0 * beta * xxe + gamma * xee
        dy = alph * yxx - 2.0 * beta * yxe + gamma * yee
        p1 = -(xix * dx + xiy * dy)
        q1 = -(etx * dx + ety * dy)
        # only interior rows/columns enter an operator; keep the border at zero
        p1[0, :] = p1[-1, :] = 0.0
        q1[0, :] = q1[-1, :] = 0.0
        if not periodic_xi:
            p1[:, 0] = p1[:, -1] = 0.0
            q1[:, 0] = q1[:, -1] = 0.0
    return dict(xix=xix, xiy=xiy, etx=etx, ety=ety, alph=alph, beta=beta,
                gamma=gamma, ajac=jac, p1=p1, q1=q1)


def _stretch(s, beta):
    if beta is None or beta == 0.0:
        return s
    return 0.5 * (1.0 + np.tanh(beta * (2.0 * s - 1.0)) / math.tanh(beta))


def cavity_grid(n0, n1=None, skew=0.05, beta=None, control=True):
    """Unit-square cavity grid with a sinusoidal skew and optional tanh clustering.

    X = s(xi), Y = s(eta) with s(t) = (1 + tanh(beta(2t-1))/tanh(beta))/2
    (identity when beta is None), then
      x = X + A sin(pi X) sin(pi Y),  y = Y + A sin(pi X) sin(pi Y).
    xi and eta are uniform on [0, 1], so dxi = 1/(n0-1), deta = 1/(n1-1).
    For the untransformed map the Jacobian is exactly 1 + A*pi*sin(pi(xi+eta)).
    """
    n1 = n0 if n1 is None else n1
    xi = np.linspace(0.0, 1.0, n0)
    et = np.linspace(0.0, 1.0, n1)
    X = _stretch(xi, beta)[None, :] * np.ones((n1, 1))
    Y = _stretch(et, beta)[:, None] * np.ones((1, n0))
    bump = skew * np.sin(math.pi * X) * np.sin(math.pi * Y)
    x = X + bump
    y = Y + bump
    # exact boundary values (sin(pi*1) is not exactly zero in floating point)
    x[:, 0] = 0.0
    x[:, -1] = 1.0
    y[0, :] = 0.0
    y[-1, :] = 1.0
    x[0, :] = X[0, :]
    x[-1, :] = X[-1, :]
    y[:, 0] = Y[:, 0]
    y[:, -1] = Y[:, -1]
    dxi = 1.0 / (n0 - 1)
    deta = 1.0 / (n1 - 1)
    m = metrics(x, y, dxi, deta, periodic_xi=False, control=control)
    z = np.zeros(n0)
    return Grid(n0=n0, n1=n1, dxi=dxi, deta=deta, periodic_xi=False, x=x, y=y,
                xnix=z.copy(), xniy=z.copy(), xnox=z.copy(), xnoy=z.copy(), **m)


def ogrid_ellipse(n0, n1, ar=1.0, r_out=10.0, stretch=2.0):
    """O-grid around an ellipse of semi-axes (1, 1/ar): the reference's topology.

    xi runs once around the body (node n0-1 duplicates node 0, as the reference
    assumes at solver.cpp:883-937), eta runs from the wall (j=0) to the far
    field (j=n1-1).  Radial spacing is geometric-like via exp stretching.
    Boundary normals xni*/xno* are the unit outward normals of the body and of
    the outer circle (the arrays INP.DAT carries, solver.cpp:544-546).
    """
    dxi = 1.0 / (n0 - 1)
    deta = 1.0 / (n1 - 1)
    xi = np.arange(n0) * dxi
    et = np.arange(n1) * deta
    th = -2.0 * math.pi * xi  # clockwise so that the Jacobian is positive
    a_in, b_in = 1.0, 1.0 / ar
    s = (np.exp(stretch * et) - 1.0) / (math.exp(stretch) - 1.0)
    xin = a_in * np.cos(th)
    yin = b_in * np.sin(th)
    xout = r_out * np.cos(th)
    yout = r_out * np.sin(th)
    x = xin[None, :] + s[:, None] * (xout - xin)[None, :]
    y = yin[None, :] + s[:, None] * (yout - yin)[None, :]
    x[:, -1] = x[:, 0]
    y[:, -1] = y[:, 0]
    m = metrics(x, y, dxi, deta, periodic_xi=True, control=False)
    if np.any(m["ajac"] <= 0):
        raise RuntimeError("O-grid Jacobian is not positive")
    # unit normals: body (pointing into the body, i.e. -grad eta) and far field (+grad eta)
    g0 = np.sqrt(m["gamma"][0])
    g1 = np.sqrt(m["gamma"][-1])
    xnix = -m["etx"][0] / g0
    xniy = -m["ety"][0] / g0
    xnox = m["etx"][-1] / g1
    xnoy = m["ety"][-1] / g1
    return Grid(n0=n0, n1=n1, dxi=dxi, deta=deta, periodic_xi=True, x=x, y=y,
                p_grid=0.0, a_grid=0.0, ar=ar, xnix=xnix, xniy=xniy, xnox=xnox,
                xnoy=xnoy, **m)


def _fmt(v):
    return repr(float(v))


def write_inp_dat(path, g: Grid):
    """Write a grid in the INP.DAT record layout the reference reads (solver.cpp:516-546).

    Values are printed with repr() so that a text round trip is exact.
    """
    n0, n1 = g.n0, g.n1
    lines = [f"{n0} {n1} {_fmt(g.dxi)} {_fmt(g.deta)}",
             f"{_fmt(g.p_grid)} {_fmt(g.a_grid)} {_fmt(g.ar)}",
             "0 0 0 0"]
    for j in range(n1):
        for i in range(n0):
            lines.append(f"{_fmt(i * g.dxi)} {_fmt(j * g.deta)} {_fmt(g.x[j, i])} {_fmt(g.y[j, i])}")
    for j in range(n1):
        for i in range(n0):
            lines.append(f"{_fmt(g.xix[j, i])} {_fmt(g.xiy[j, i])} {_fmt(g.etx[j, i])} {_fmt(g.ety[j, i])}")
    for j in range(n1):
        for i in range(n0):
            lines.append(f"{_fmt(g.alph[j, i])} {_fmt(g.beta[j, i])} {_fmt(g.gamma[j, i])}")
    for j in range(n1):
        for i in range(n0):
            lines.append(_fmt(g.ajac[j, i]))
    for i in range(n0):
        lines.append(f"{_fmt(g.xnix[i])} {_fmt(g.xniy[i])} {_fmt(g.xnox[i])} {_fmt(g.xnoy[i])}")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def read_inp_dat(path, periodic_xi=True):
    """Read an INP.DAT file (solver.cpp:516-553): free-format, whitespace separated.

    p1 and q1 are set to zero as the reference does (solver.cpp:548-553).
    """
    with open(path) as fh:
        tok = fh.read().split()
    pos = 0

    def take(k):
        nonlocal pos
        v = tok[pos:pos + k]
        if len(v) < k:
            raise ValueError(f"INP.DAT truncated at token {pos}")
        pos += k
        return v

    n0, n1 = int(take(1)[0]), int(take(1)[0])
    if n0 < 5 or n1 < 5:
        raise ValueError(f"INP.DAT grid {n0}x{n1} too small")
    dxi, deta = float(take(1)[0]), float(take(1)[0])
    p_grid, a_grid, ar = (float(t) for t in take(3))
    take(4)
    N = n0 * n1
    rec = np.array(take(4 * N), dtype=np.float64).reshape(n1, n0, 4)
    x, y = rec[..., 2].copy(), rec[..., 3].copy()
    rec = np.array(take(4 * N), dtype=np.float64).reshape(n1, n0, 4)
    xix, xiy, etx, ety = (rec[..., k].copy() for k in range(4))
    rec = np.array(take(3 * N), dtype=np.float64).reshape(n1, n0, 3)
    alph, beta, gamma = (rec[..., k].copy() for k in range(3))
    ajac = np.array(take(N), dtype=np.float64).reshape(n1, n0)
    rec = np.array(take(4 * n0), dtype=np.float64).reshape(n0, 4)
    z = np.zeros((n1, n0))
    return Grid(n0=n0, n1=n1, dxi=dxi, deta=deta, periodic_xi=periodic_xi, x=x, y=y,
                xix=xix, xiy=xiy, etx=etx, ety=ety, alph=alph, beta=beta, gamma=gamma,
                ajac=ajac, p1=z.copy(), q1=z.copy(), p_grid=p_grid, a_grid=a_grid, ar=ar,
                xnix=rec[:, 0].copy(), xniy=rec[:, 1].copy(), xnox=rec[:, 2].copy(),
                xnoy=rec[:, 3].copy())


def round_pow2p1(n):
    """Round a node count to the nearest 2^k+1 in log space (65, 129, 257, ...)."""
    k = round(math.log2(max(n - 1, 1)))
    lo, hi = 2 ** k + 1, 2 ** (k + 1) + 1
    cands = [2 ** (k - 1) + 1, lo, hi]
    return min(cands, key=lambda c: abs(math.log(c - 1) - math.log(max(n - 1, 1))))
