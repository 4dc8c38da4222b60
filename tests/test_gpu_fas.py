"""FAS multigrid outer cycles on the GPU against the CPU oracle (FP64: 1e-9 relative L2)."""
import numpy as np
import pytest

from cfd_solver_amd.grid import cavity_grid
from tests.conftest import rel_l2
from tests.test_gpu_cavity import pair

pytestmark = pytest.mark.gpu


def test_fas_cycles_match_oracle():
    """Config-2 physics (Re 400, Ri 1, Pr 0.71, skew 0.1, tanh 1.5) at 65x65: levels 65,
    33, 17; eight cycles, each with equal p' iteration counts in its last fine SIMPLE
    iteration, and fields, residual norms and Nusselt numbers equal to the oracle's."""
    print("\n[gpu] FAS 65x65 (3 levels), 8 cycles vs oracle", flush=True)
    gpu, cpu, g = pair(65, skew=0.1, beta=1.5, Re=400.0, Ri=1.0, Pr=0.71)
    assert cpu.fas_depth() == 3
    for k in range(8):
        sg = gpu.step()
        sc = cpu.step()
        assert sg.p_iters == sc.p_iters, k
        assert abs(sg.mass_res - sc.mass_res) <= 1e-7 * sc.mass_res, (k, sg.mass_res, sc.mass_res)
        assert abs(sg.energy_res - sc.energy_res) <= 1e-7 * sc.energy_res, k
    fg, fc = gpu.fields(), cpu.fields()
    for k in ("u", "v", "T", "p"):
        assert rel_l2(fg[k], fc[k]) < 1e-9, k
    assert abs(sg.nu_hot - sc.nu_hot) < 1e-9 * abs(sc.nu_hot)


def test_fas_converges_to_1e8_at_129():
    """129x129 config-2 case to R_m, R_T <= 1e-8 on the GPU in FP64: at most 80 cycles
    (the oracle needs 54 at this size, 56 at 257)."""
    print("\n[gpu] FAS 129x129 to 1e-8", flush=True)
    from cfd_solver_amd.cavity import CavityCase
    c = CavityCase(cavity_grid(129, skew=0.1, beta=1.5), Re=400.0, Ri=1.0)
    st, ok = c.solve(tol=1e-8, max_outer=80)
    print(f"[gpu] cycles {st.iter} mass {st.mass_res:.2e} energy {st.energy_res:.2e} Nu {st.nu_hot:.10f}", flush=True)
    assert ok, (st.iter, st.mass_res, st.energy_res)
