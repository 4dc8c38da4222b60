"""The reference solver's output and restart files (solver.cpp:1984-2039), written
from host arrays in the (n1, n0) layout of this package.

  write_tecplot     spt100.dat: "zone", I=, J=, then per node x y (%.9f) and
                    u v T p si vort (%.13e), a blank line after each eta line
  append_coeff_his  COEFF_HIS.dat: time cl cd cm Nu (%.8f; the reference's
                    time is an int and prints as one)
  append_coeff_pr_vor  COEFF_HIS_pr_vor.dat: time CL_pr CD_pr CL_vor CD_vor
  write_surf_dist   SURF_DIST.dat: xi, wall p, wall vorticity, dT/dn in the
                    stream's default format (6 significant digits)
  write_restart / read_restart
                    a working counterpart of spa100.dat: the reference writes
                    the pointers of its arrays instead of their contents
                    (solver.cpp:2004-2012) and reads time as an int.  Here a
                    header (magic, n0, n1, loop, time as a double, dmax) is
                    followed by the reference's record order -- x, y, si, u, v,
                    T, p, each as the reference's a[i][j] (j fastest) -- and the
                    predictor velocities up, vp, which an exact continuation
                    needs.
"""
from __future__ import annotations

import struct

import numpy as np

RESTART_MAGIC = b"CFDRST01"


def _g6(x: float) -> str:
    """A double as an ostream with default flags and precision prints it."""
    return format(float(x), ".6g")


def write_tecplot(path, x, y, u, v, T, p, si, vort):
    n1, n0 = np.shape(x)
    arrs = [np.asarray(a, dtype=np.float64) for a in (x, y, u, v, T, p, si, vort)]
    out = [f"zone\nI={n0}\nJ={n1}\n"]
    for j in range(n1):
        rows = [a[j] for a in arrs]
        for i in range(n0):
            out.append("%.9f %.9f %.13e %.13e %.13e %.13e %.13e %.13e\n"
                       % tuple(r[i] for r in rows))
        out.append("\n")
    with open(path, "w") as f:
        f.write("".join(out))


def _time_str(time) -> str:
    return str(time) if isinstance(time, (int, np.integer)) else "%.8f" % time


def append_coeff_his(path, time, cl, cd, cm, nu):
    with open(path, "a") as f:
        f.write("%s %.8f %.8f %.8f %.8f\n" % (_time_str(time), cl, cd, cm, nu))


def append_coeff_pr_vor(path, time, cl_pr, cd_pr, cl_vor, cd_vor):
    with open(path, "a") as f:
        f.write("%s %.8f %.8f %.8f %.8f\n" % (_time_str(time), cl_pr, cd_pr, cl_vor, cd_vor))


def write_surf_dist(path, dxi, deta, p, vort, T, etx, ety):
    p, vort, T, etx, ety = (np.asarray(a, dtype=np.float64) for a in (p, vort, T, etx, ety))
    n0 = p.shape[1]
    out = []
    for i in range(n0):
        d = -(4 * T[1, i] - 3 * T[0, i] - T[2, i]) / (2 * deta)
        d = d * np.sqrt(etx[0, i] * etx[0, i] + ety[0, i] * ety[0, i])
        out.append(f"{_g6(i * dxi)} {_g6(p[0, i])} {_g6(vort[0, i])} {_g6(d)}\n")
    with open(path, "w") as f:
        f.write("".join(out))


_HEAD = struct.Struct("<8siiiidd")  # magic, n0, n1, loop, reserved, time, dmax
_FIELDS = ("x", "y", "si", "u", "v", "T", "p", "up", "vp")


def write_restart(path, loop, time, dmax, fields):
    """fields: x, y, si, u, v, T, p, up, vp as (n1, n0) arrays."""
    n1, n0 = np.shape(fields["x"])
    with open(path, "wb") as f:
        f.write(_HEAD.pack(RESTART_MAGIC, n0, n1, int(loop), 0, float(time), float(dmax)))
        for k in _FIELDS:
            a = np.asarray(fields[k], dtype="<f8")
            if a.shape != (n1, n0):
                raise ValueError(f"{k}: shape {a.shape}, expected {(n1, n0)}")
            f.write(np.ascontiguousarray(a.T).tobytes())  # a[i][j], j fastest


def read_restart(path):
    with open(path, "rb") as f:
        head = f.read(_HEAD.size)
        if len(head) != _HEAD.size:
            raise ValueError("truncated restart header")
        magic, n0, n1, loop, _, time, dmax = _HEAD.unpack(head)
        if magic != RESTART_MAGIC or n0 < 1 or n1 < 1:
            raise ValueError("not a restart file of this format")
        out = {"loop": loop, "time": time, "dmax": dmax, "n0": n0, "n1": n1}
        for k in _FIELDS:
            buf = f.read(8 * n0 * n1)
            if len(buf) != 8 * n0 * n1:
                raise ValueError(f"truncated restart record {k}")
            out[k] = np.frombuffer(buf, dtype="<f8").reshape(n0, n1).T.copy()
        if f.read(1):
            raise ValueError("trailing bytes after the restart records")
    return out
