"""Write tests/golden/cavity_129_re400.npz: the oracle's converged cavity at 129 x 129.

The configuration-2 case (skew 0.1, tanh clustering 1.5, Re 400, Ri 1, Pr 0.71,
hot bottom, cold lid, adiabatic sides) on a 129 x 129 grid, solved by the CPU
oracle (oracle/cfd_oracle_cavity.c, FP64 FAS cycles with SIMPLE smoothing) until
the mass and energy residual ratios are below 1e-11.  Holds u, v, T, p, the wall
Nusselt numbers and the cycle count; the GPU test compares its own converged
fields with these (1e-9 relative L2 in FP64, 1e-6 in the mixed path).

    python tests/golden/make_cavity_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cfd_solver_amd.grid import cavity_grid  # noqa: E402
from oracle.oracle import CavityOracle, cav_params  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "cavity_129_re400.npz")
N, SKEW, BETA = 129, 0.1, 1.5
PHYS = dict(Re=400.0, Ri=1.0, Pr=0.71)
TOL = 1e-11


def main():
    g = cavity_grid(N, skew=SKEW, beta=BETA)
    c = CavityOracle(g, cav_params(**PHYS))
    for k in range(1, 500):
        st = c.step()
        if st.mass_res < TOL and st.energy_res < TOL:
            break
    f = c.fields()
    np.savez_compressed(OUT, u=f["u"], v=f["v"], T=f["T"], p=f["p"], nu_hot=st.nu_hot, nu_cold=st.nu_cold,
                        cycles=k, mass_res=st.mass_res, energy_res=st.energy_res, n=N, skew=SKEW, beta=BETA,
                        **{k_: v for k_, v in PHYS.items()})
    print(f"{OUT}: {k} cycles, mass {st.mass_res:.2e}, energy {st.energy_res:.2e}, Nu {st.nu_hot:.12f}")


if __name__ == "__main__":
    main()
