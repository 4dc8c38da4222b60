"""Regenerate tests/golden/ref_ogrid_33x17.npz from the compiled reference solver.

Runs only where the reference checkout is present (it is never read on the GPU
box): builds oracle/_ref/ref_harness (solver.cpp compiled from where it lies),
writes a 33x17 O-grid INP.DAT around an ellipse of aspect ratio 2 with this
package's grid generator, lets the reference run three time steps and packs
its arrays into the fixture.

Provenance of the values: Re = 1000, Pr = 0.71, dt = 1e-4, alpha = 82 deg are
the reference's compiled-in constants (solver.cpp:159-182); the harness sets
Ri = 1 after step 0 so that steps 1-2 exercise the buoyancy terms (the
operators do not depend on Ri).  The converged solves ("*_conv") are the
reference's own sip9p/gauss called repeatedly on step 2's systems.

    python tests/golden/make_ref_golden.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cfd_solver_amd.grid import ogrid_ellipse, write_inp_dat  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "ref_ogrid_33x17.npz")
N0, N1 = 33, 17

PLANES25 = {"p": (0, 0), "e": (1, 0), "w": (-1, 0), "n": (0, 1), "s": (0, -1), "ne": (1, 1),
            "nw": (-1, 1), "se": (1, -1), "sw": (-1, -1), "ee": (2, 0), "ww": (-2, 0), "nn": (0, 2),
            "ss": (0, -2), "nnee": (2, 2), "nnww": (-2, 2), "ssee": (2, -2), "ssww": (-2, -2),
            "nne": (1, 2), "nnw": (-1, 2), "sse": (1, -2), "ssw": (-1, -2), "nee": (2, 1),
            "nww": (-2, 1), "see": (2, -1), "sww": (-2, -1)}
PLANES9 = {"ap": (0, 0), "ae": (1, 0), "aw": (-1, 0), "an": (0, 1), "as": (0, -1), "ane": (1, 1),
           "anw": (-1, 1), "ase": (1, -1), "asw": (-1, -1)}
STATE = ["u", "v", "T", "up", "vp", "us", "vs", "uold", "vold", "Told", "p", "pcor", "q", "qu",
         "qv", "qt", "qup", "qvp", "dil", "uxi", "uet", "si", "vort"]


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    g = ogrid_ellipse(N0, N1, ar=2.0, r_out=8.0)
    with tempfile.TemporaryDirectory() as work:
        write_inp_dat(os.path.join(work, "INP.DAT"), g)
        out = os.path.join(work, "out")
        os.makedirs(out)
        subprocess.run([harness, work, out, "2", "--ri", "1.0", "--solve-step", "2"], check=True)

        def rd(name, shape=(N1, N0)):
            a = np.fromfile(os.path.join(out, name + ".bin"), dtype=np.float64)
            return a.reshape(shape) if shape is not None else a

        d = {}
        sc = rd("setup_scalars", None)
        d["meta"] = np.array([N0, N1, g.dxi, g.deta, sc[0], sc[1], sc[2], 1.0, sc[4], sc[5], sc[6]])
        for k in ["x", "y", "xix", "xiy", "etx", "ety", "alph", "beta", "gamma", "ajac"]:
            d["grid_" + k] = rd("setup_" + k)
        d["grid_xnox"] = rd("setup_xnox", None)
        d["grid_xnoy"] = rd("setup_xnoy", None)
        A9 = np.zeros((9, N1, N0))
        for nm, (di, dj) in PLANES9.items():
            A9[(dj + 1) * 3 + (di + 1)] = rd("setup_" + nm)
        d["A9"] = A9
        for pre in ("au", "at"):
            A = np.zeros((25, N1, N0))
            for nm, (di, dj) in PLANES25.items():
                A[(dj + 2) * 5 + (di + 2)] = rd("setup_" + pre + nm)
            d["A" + pre[1].upper()] = A
        d["BU"] = np.stack([rd("setup_" + k, None) for k in ["bus", "buse", "busw", "bun", "bune", "bunw"]])
        d["BT"] = np.stack([rd("setup_" + k, None) for k in ["bts", "btse", "btsw", "btn", "btne", "btnw"]])
        for k in ["p", "up", "vp", "T"]:
            d["s1_" + k] = rd("s1_" + k)
        for k in STATE:
            d["s2_" + k] = rd("s2_" + k)
        d["s2_dmax"] = rd("s2_dmax", None)
        d["s2_Ri"] = rd("s2_Ri", None)
        for k in ["pcor_conv", "pcor_one", "u_conv", "T_conv"]:
            d["s2_" + k] = rd("s2_" + k)
        for k in ["u", "v", "T", "p", "si", "vort"]:
            d["s0_" + k] = rd("s0_" + k)
        his = open(os.path.join(out, "s0_COEFF_HIS.dat")).read().split()
        d["s0_nu_8dp"] = np.array([float(his[4])])
        # COEFF_HIS.dat: time, CL, CD, CM, Nu; COEFF_HIS_pr_vor.dat: time, CL_pr, CD_pr,
        # CL_vor, CD_vor (8 decimals each); SURF_DIST.dat: xi, p, vorticity, dT/dn at
        # the wall (6 significant digits); spt100.dat, the Tecplot field file, as bytes
        d["s0_coeff_his_8dp"] = np.array([float(v) for v in his[:5]])
        prv = open(os.path.join(out, "s0_COEFF_HIS_pr_vor.dat")).read().split()
        d["s0_coeff_pr_vor_8dp"] = np.array([float(v) for v in prv[:5]])
        surf = np.loadtxt(os.path.join(out, "s0_SURF_DIST.dat"))
        d["s0_dthdn_6sig"] = surf[:, 3].copy()
        d["s0_surf_dist_6sig"] = surf.copy()
        for name in ("spt100", "COEFF_HIS", "COEFF_HIS_pr_vor", "SURF_DIST"):
            with open(os.path.join(out, "s0_" + name + ".dat"), "rb") as f:
                d["s0_" + name.lower() + "_bytes"] = np.frombuffer(f.read(), dtype=np.uint8).copy()
    np.savez_compressed(OUT, **d)
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
