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
    gpu, cpu, g = pair(65, skew=0.1, beta=1.5, Re=400.0, Ri=1.0, Pr=0.71, fas_min_n=17)
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


@pytest.mark.parametrize("pmaxit", [-1, 1])
def test_fas_coarse_p_solves_match_oracle(pmaxit):
    """129x129 with levels 129, 65, 33, 17: the 65x65 level runs the stage kernels, so its
    p' solve is the multi-kernel one-V-cycle path (fas_coarse_pmaxit -1, the default) or
    one BiCGStab iteration (1); five cycles equal the oracle's in both (fields to 1e-9,
    residual ratios to 1e-7, the same fine-level p' iteration counts)."""
    print(f"\n[gpu] FAS 129x129 fas_coarse_pmaxit {pmaxit}, 5 cycles vs oracle", flush=True)
    gpu, cpu, g = pair(129, skew=0.1, beta=1.5, Re=400.0, Ri=1.0, Pr=0.71, fas_min_n=17,
                       fas_coarse_pmaxit=pmaxit)
    assert cpu.fas_depth() == 4
    for k in range(5):
        sg = gpu.step()
        sc = cpu.step()
        assert sg.p_iters == sc.p_iters, k
        assert abs(sg.mass_res - sc.mass_res) <= 1e-7 * sc.mass_res, (k, sg.mass_res, sc.mass_res)
    fg, fc = gpu.fields(), cpu.fields()
    for k in ("u", "v", "T", "p"):
        assert rel_l2(fg[k], fc[k]) < 1e-9, k


@pytest.mark.parametrize("prec", [0, 1])
def test_fas_nonsquare_matches_oracle(prec):
    """A 129 x 65 grid (levels 129 x 65, 65 x 33, 33 x 17): five cycles against the
    oracle, FP64 fields to 1e-9 and mixed to 1e-6, with the one-V-cycle p' solves on
    the non-square coarse levels."""
    print(f"\n[gpu] FAS 129x65, precision {prec}, 5 cycles vs oracle", flush=True)
    gpu, cpu, g = pair(129, n1=65, skew=0.1, beta=1.5, Re=400.0, Ri=1.0, Pr=0.71, fas_min_n=17, precision=prec)
    assert cpu.fas_depth() == 3
    for k in range(5):
        sg = gpu.step()
        sc = cpu.step()
        if prec == 0:
            assert sg.p_iters == sc.p_iters, k
    fg, fc = gpu.fields(), cpu.fields()
    for k in ("u", "v", "T", "p"):
        assert rel_l2(fg[k], fc[k]) < (1e-9 if prec == 0 else 1e-6), (k, rel_l2(fg[k], fc[k]))


def test_fas_converges_to_1e8_at_129():
    """129x129 config-2 case to R_m, R_T <= 1e-8 on the GPU in FP64 within 100 cycles
    (the oracle needs 64-68 from 65 to 257 nodes per side)."""
    print("\n[gpu] FAS 129x129 to 1e-8", flush=True)
    from cfd_solver_amd.cavity import CavityCase
    c = CavityCase(cavity_grid(129, skew=0.1, beta=1.5), Re=400.0, Ri=1.0)
    st, ok = c.solve(tol=1e-8, max_outer=100)
    print(f"[gpu] cycles {st.iter} mass {st.mass_res:.2e} energy {st.energy_res:.2e} Nu {st.nu_hot:.10f}", flush=True)
    assert ok, (st.iter, st.mass_res, st.energy_res)


@pytest.mark.parametrize("prec,tol", [(0, 1e-9), (1, 1e-6)])
def test_converged_129_matches_oracle_fixture(prec, tol):
    """Config-2 physics at 129x129 converged to residual ratios below 1e-11 on the GPU,
    FP64 and mixed, against the oracle's converged fields (tests/golden/cavity_129_re400.npz,
    written by tests/golden/make_cavity_golden.py): relative L2 of u, v, T, p (means
    removed) and Nu below 1e-9 (FP64) / 1e-6 (mixed); per-field errors printed."""
    import os
    from cfd_solver_amd.cavity import CavityCase
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "cavity_129_re400.npz"))
    n = int(ref["n"])
    print(f"\n[gpu] 129x129 to 1e-11, precision {prec}, vs the oracle fixture", flush=True)
    c = CavityCase(cavity_grid(n, skew=float(ref["skew"]), beta=float(ref["beta"])), Re=float(ref["Re"]),
                   Ri=float(ref["Ri"]), Pr=float(ref["Pr"]), precision=prec)
    for k in range(300):
        st = c.step()
        if st.mass_res < 1e-11 and st.energy_res < 1e-11:
            break
    assert st.mass_res < 1e-11 and st.energy_res < 1e-11, (k, st.mass_res, st.energy_res)
    f = c.fields()
    errs = {q: rel_l2(f[q], ref[q]) for q in ("u", "v", "T")}
    pg, pc = f["p"] - f["p"][1:-1, 1:-1].mean(), ref["p"] - ref["p"][1:-1, 1:-1].mean()
    errs["p"] = rel_l2(pg, pc)
    errs["Nu"] = abs(st.nu_hot - float(ref["nu_hot"])) / abs(float(ref["nu_hot"]))
    print(f"[gpu] cycles {st.iter}; per-field rel L2 vs oracle:", {q: f"{v:.2e}" for q, v in errs.items()}, flush=True)
    for q, v in errs.items():
        assert v < tol, (q, v)


def test_high_re_high_pr_cycles_match_oracle():
    """Config-4 extreme physics, Re 1000, Pr 7, Ri 1 at 65x65 (heated from below, Ra 7e6:
    no steady state is expected, the residual stays O(1)): six FAS cycles of the GPU
    equal the oracle's, fields to 1e-9 and the same p' iteration counts, and stay finite."""
    print("\n[gpu] FAS 65x65 Re 1000 Pr 7 Ri 1, 6 cycles vs oracle", flush=True)
    gpu, cpu, g = pair(65, skew=0.15, beta=1.5, Re=1000.0, Ri=1.0, Pr=7.0)
    for k in range(6):
        sg = gpu.step()
        sc = cpu.step()
        assert sg.p_iters == sc.p_iters, k
        assert np.isfinite(sg.mass_res) and np.isfinite(sg.energy_res)
    fg, fc = gpu.fields(), cpu.fields()
    for q in ("u", "v", "T", "p"):
        assert rel_l2(fg[q], fc[q]) < 1e-9, q


@pytest.mark.parametrize("prec", [0, 1])
def test_small_level_kernel_equals_stage_kernels(prec):
    """The coarse 33x33 and 17x17 levels of a 65x65 hierarchy run in one workgroup per
    case (fas_small 1, FP64) or in the stage kernels (fas_small 0): after five cycles
    the fields agree to 1e-9 (FP64) / 1e-6 (mixed, whose stage kernels run the coarse
    levels in FP32), and the small-level run matches the oracle as closely as the
    stage kernels do (FP64)."""
    import time
    from cfd_solver_amd.cavity import CavityCase
    print(f"\n[gpu] 65x65 FAS (33, 17 coarse), fas_small 1 vs 0, precision {prec}", flush=True)
    g = cavity_grid(65, skew=0.1, beta=1.5)
    kw = dict(Re=400.0, Ri=1.0, Pr=0.71, fas_min_n=17, precision=prec)
    a, b = CavityCase(g, fas_small=1, **kw), CavityCase(g, fas_small=0, **kw)
    ta = tb = 0.0
    for k in range(5):
        t0 = time.perf_counter()
        sa = a.step()
        ta += time.perf_counter() - t0
        t0 = time.perf_counter()
        sb = b.step()
        tb += time.perf_counter() - t0
        assert np.isfinite(sa.mass_res) and np.isfinite(sb.mass_res)
    print(f"[gpu] ms per cycle: small kernel {ta / 5 * 1e3:.2f}, stage kernels {tb / 5 * 1e3:.2f}; "
          f"mass {sa.mass_res:.3e} / {sb.mass_res:.3e}", flush=True)
    fa, fb = a.fields(), b.fields()
    tol = 1e-9 if prec == 0 else 1e-6
    for q in ("u", "v", "T", "p"):
        assert rel_l2(fa[q], fb[q]) < tol, (q, rel_l2(fa[q], fb[q]))
    assert abs(sa.mass_res - sb.mass_res) <= 1e-4 * sb.mass_res
