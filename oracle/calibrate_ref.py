"""Timing calibration of the CPU baseline against the compiled reference -- TEST
INFRASTRUCTURE ONLY; runs only on a machine that holds the reference checkout
(the reference's sources are read from there and compiled by the oracle's Makefile).

The reference (solver.cpp, compiled from where it lies by `make -C oracle ref` at
-O0 as its own Makefile builds it, and `make -C oracle ref-o3` at -O3) and the
oracle's reference-faithful step (cfd_oracle_ref.c: the same stages, SIP and
Gauss-Seidel iterate for iterate, gcc -O3, one thread) advance the same O-grid
INP.DAT (an ellipse of aspect ratio 2, outer radius 20, written by
cfd_solver_amd.grid) from the reference's initial state; the median wall time of
`steps` steps after step 0 is printed for each.

    python oracle/calibrate_ref.py [n0 n1 steps]
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main():
    n0, n1, steps = (int(a) for a in (sys.argv[1:4] if len(sys.argv) >= 4 else (350, 570, 5)))
    from cfd_solver_amd.grid import ogrid_ellipse, write_inp_dat
    from cfd_solver_amd.refcase import initial_state, ref_params
    from oracle.oracle import build, lib, ref_step
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "ref-o3"], check=True)
    g = ogrid_ellipse(n0, n1, ar=2.0, r_out=20.0)
    res = {}
    with tempfile.TemporaryDirectory() as work:
        write_inp_dat(os.path.join(work, "INP.DAT"), g)
        out = os.path.join(work, "out")
        os.makedirs(out)
        for tag, exe in (("reference -O0", "ref_harness"), ("reference -O3", "ref_harness_o3")):
            p = subprocess.run([os.path.join(ROOT, "oracle", "_ref", exe), work, out, str(steps), "--time", "1"],
                               check=True, capture_output=True, text=True)
            ms = [float(x) for x in p.stderr.split()]
            res[tag] = float(np.median(ms))
    build()
    L = lib()
    prm = ref_params()
    AU, AT = np.zeros((25, n1, n0)), np.zeros((25, n1, n0))
    BU, BT, A9 = np.zeros((6, n0)), np.zeros((6, n0)), np.zeros((9, n1, n0))
    z = np.zeros((n1, n0))
    L.or_assemble_mom_ref(n0, n1, g.dxi, g.deta, g.alph, g.beta, g.gamma, z, z, prm.dt, prm.Re, prm.Pr,
                          AU, AT, BU, BT)
    L.or_assemble_p9_ref(n0, n1, g.dxi, g.deta, g.xix, g.xiy, g.etx, g.ety, A9)
    gd = {k: getattr(g, k) for k in ("x", "y", "xix", "xiy", "etx", "ety", "alph", "beta", "gamma", "ajac",
                                     "xnox", "xnoy")}
    ops = dict(AU=AU, AT=AT, BU=BU, BT=BT, A9=A9)
    pd = dict(dxi=g.dxi, deta=g.deta, dt=prm.dt, Re=prm.Re, Pr=prm.Pr, Ri=prm.Ri, uinf=prm.uinf, vinf=prm.vinf)
    st = dict(initial_state(g, prm))
    o = ref_step(gd, ops, st, pd)  # step 0
    st = {k: o[k] for k in ("u", "v", "T", "p", "up", "vp")}
    ms = []
    for _ in range(steps):
        t0 = time.perf_counter()
        o = ref_step(gd, ops, st, pd)
        ms.append((time.perf_counter() - t0) * 1e3)
        st = {k: o[k] for k in ("u", "v", "T", "p", "up", "vp")}
    res["oracle reference-faithful step (gcc -O3)"] = float(np.median(ms))
    print(f"{n0} x {n1} O-grid, median ms per step over {steps} steps after step 0:")
    for k, v in res.items():
        print(f"  {k:42s} {v:10.1f}")


if __name__ == "__main__":
    main()
