#!/usr/bin/env python3
"""Benchmark: the steady mixed-convection cavity's outer iterations on one MI355X.

Headline (value): 1025 x 1025 nodes, skew amplitude 0.1 with tanh wall clustering
beta = 1.5, Re = 400, Ri = 1, Pr = 0.71 (configuration 2), mixed path (FP32
predictor sweeps and p' inner solves, FP64 residuals, refinement, accumulation and
reductions), the FP64 path beside it.  One step = one outer iteration = one FAS
multigrid cycle over the whole case (SIMPLE iterations as the smoother on grids
1025, 513, ..., 17: two before and two after the coarse-grid correction on each
level, 40 on the coarsest), fields resident in HBM; the K timed steps are one
call of the library's outer loop.  A cell-iteration is one interior node through
one SIMPLE iteration on any level, so value = sum over levels of interior nodes x
SIMPLE iterations per cycle x cycles / s.  Then each precision solves the case
from rest to mass and energy residual ratios of 1e-8 (time to solution, cycles)
and the converged mixed fields are compared with the FP64 ones per field.

    python bench.py [--steps K] [--warmup W] [--n 1025] [--no-cpu]
                    [--sections main,batch,serve,piso,ogrid]
                    (pack: the packed 256-case mix alone; none: the timed cycles only)

Prints one JSON line.  roofline: the fine-level kernel with the largest device
time per step, timed live with HIP events on the case's stream; its algorithmic
bytes per launch are (arrays read + arrays written) x nodes x element size; a
whole-cycle estimate (the per-node byte model summed over every level's SIMPLE
iterations, divided by the step time) stands beside it.  cpu_baseline times the
CPU oracle (the same FAS cycles in serial FP64 C) on this host for a bounded
number of cycles of the same case.  Further objects in config:
  solve     1025^2 from rest to 1e-8, FP64 and mixed: cycles, seconds, per-field
            relative L2 (u, v, T, p, Nu) of mixed against FP64;
  batched   the 256-case mix (65..513 nodes per side, Re 50..1000, Ri 0.1..10, Pr
            0.71/7) as batches of one size and multigrid depth, every slot to 1e-8
            or its cycle cap: done / not converged / failed counts, time;
  served    Poisson-arriving solves of 65..1025 nodes per side with the same
            physics draws, continuous batching to 1e-8: solves/s, latency p50/p99,
            beside the same requests solved one at a time on the plain FP64 path;
  p_isolate the p' solve alone at 2049 x 2049 to 1e-10, FP64 vs mixed, with the
            algorithmic HBM rate of its iterations;
  ref_ogrid the reference's own O-grid time step resident on the GPU at the
            reference's largest array size (350 x 570), with the CPU oracle's
            reference-faithful step (SIP + Gauss-Seidel, serial) beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SIMPLE cell-iters/s at 1025² FP64 vs mixed; served solves/s + p99 at 1e-8 res"
HBM_PEAK_GBS = 8000.0
PHYS = dict(Re=400.0, Ri=1.0, Pr=0.71)


# algorithmic bytes per node per launch: (arrays read + arrays written) x element size
def kernel_bytes(name, mixed):
    w = 4 if mixed else 8  # predictor and p' inner solve precision
    return {
        "p_apply": 7 * w,          # x, 4 face coefficients, r0 (or s) -> y; dot product in the kernel
        "mg_down": 7.25 * w,       # b, 1/diag, 4 face coefficients -> x, restricted residual (1/4 node)
        # finest up leg fused with BiCGStab's operator application: x, b, 1/diag, 4 faces,
        # coarse x (1/4 node), r0 (every other launch) -> z, y = L z
        "mg_up": 9.75 * w,
        "jacobi": 16 * w,          # four sweeps: R_u,R_v,R_T, U,V, alph,beta,gamma,P,Q, du,dv,dT -> du,dv,dT
        "residual": 13 * 8 + 8 + 5 * w,  # u,v,T,p, 4 metrics, alph..Q (FP64) -> d (FP64), R_u,R_v,R_T,U,V
        "face_div": 10 * 8 + 2 * w,  # u,v,p,d, 4 metrics, J (FP64), du,dv -> q (FP64)
    }[name]


def simple_bytes_per_node(mixed, p_iters):
    """Algorithmic bytes per node of one SIMPLE iteration: residual, two four-sweep
    predictor launches, face divergence, face coefficients of every p' level (4/3 of
    the fine level's), correction, and per BiCGStab iteration two V-cycles (fine down
    and fused up-and-apply legs, coarser levels a third more) plus the vector updates
    (p: 4 words, s: 3, x and r: 8); the mixed path adds one FP64 defect/p' pass.
    p_iters < 0: the p' solve is one V-cycle (start pass q -> x, r, r0; the down legs
    and plain up legs, which skip the operator application's 1.5 words)."""
    w = 4 if mixed else 8
    fixed = (kernel_bytes("residual", mixed) + 2 * kernel_bytes("jacobi", mixed) + kernel_bytes("face_div", mixed)
             + (10 * 8 + 4 * w) * 4 / 3 + 13 * 8 + 3 * w + (13 * 8 if mixed else 0))
    if p_iters < 0:
        return fixed + 8 + 3 * w + (kernel_bytes("mg_down", mixed) + kernel_bytes("mg_up", mixed) - 1.5 * w) * 4 / 3
    per_it = 2 * (kernel_bytes("mg_down", mixed) + kernel_bytes("mg_up", mixed)) * 4 / 3 + 15 * w
    return fixed + per_it * max(p_iters, 1)


def hbm_traffic(mixed, name, n0, n1):
    """HBM bytes per launch of kernel `name` at 1025 x 1025 from the committed counter
    summary profiles/hbm_traffic_cavity1025.json (two rocprofv3 runs, FETCH_SIZE then
    WRITE_SIZE, each scaled by its reading for a known-size copy), or None."""
    if (n0, n1) != (1025, 1025):
        return None
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic_cavity1025.json")) as f:
            return json.load(f)["mixed" if mixed else "fp64"].get(name)
    except (OSError, KeyError, ValueError):
        return None


def copy_peak(mib=2048, reps=10):
    """Device-to-device copy rate of this GPU (hipMemcpyAsync of `mib` MiB, best of
    `reps` timed copies with HIP events), as read + write GB/s: the HBM rate a
    streaming kernel can reach here, beside the 8 TB/s data-sheet peak."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = C.c_size_t(mib << 20)
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), n) != 0 or hip.hipMalloc(C.byref(b), n) != 0:
        return None
    e0, e1 = C.c_void_p(), C.c_void_p()
    hip.hipEventCreate(C.byref(e0))
    hip.hipEventCreate(C.byref(e1))
    hip.hipMemset(a, 1, n)
    best = None
    for _ in range(reps + 1):
        hip.hipEventRecord(e0, None)
        hip.hipMemcpyAsync(b, a, n, 3, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        best = ms.value if best is None else min(best, ms.value)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return round(2 * n.value / (best * 1e-3) / 1e9, 1)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    # the default window, cycles 26-55 of a solve that converges in 59, lies late in
    # the solve (its per-cycle time matches the early cycles' within 1%)
    ap.add_argument("--warmup", type=int, default=25)
    ap.add_argument("--n", type=int, default=1025)
    ap.add_argument("--cpu-cycles", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--sections", default="main,batch,serve,piso,ogrid")
    ap.add_argument("--serve-rate", type=float, default=6.0, help="served arrivals per second")
    ap.add_argument("--serve-count", type=int, default=48)
    ap.add_argument("--pack-graph", type=int, default=1, help="1: replay the pack's cycle as a captured graph; "
                    "0: launch it eagerly (same kernels and results)")
    ap.add_argument("--copy-check", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (the known-size kernel that scales "
                         "rocprof's FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def host_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def cpu_baseline(grid, n_cycles):
    """The CPU oracle (serial FP64 C, gcc -O3) on the same case: one untimed cycle, then
    n_cycles cycles timed one by one (median); cell-iterations counted as on the GPU."""
    from cfd_solver_amd.cavity import cells_per_outer, default_params
    from oracle.oracle import CavityOracle, build, cav_params
    build()
    c = CavityOracle(grid, cav_params(**PHYS))
    c.step()
    runs = []
    for _ in range(n_cycles):
        t0 = time.perf_counter()
        c.step()
        runs.append(time.perf_counter() - t0)
    dt = float(np.median(runs))
    cells = cells_per_outer(grid.n0, grid.n1, default_params(**PHYS))
    return {"value": cells / dt, "unit": "cell-iter/s", "cores": 1, "kind": "port",
            "sample": f"median of {n_cycles} FAS cycles timed one by one (after one untimed) of the same "
                      f"{grid.n0}x{grid.n1} case, serial FP64 C oracle (gcc -O3), 1 of {os.cpu_count()} host "
                      f"threads ({host_model()})", "cycle_seconds": [round(r, 3) for r in runs]}


def run_gpu(grid, mixed, steps, warmup, prof_steps=3):
    """Timed region: `steps` outer iterations (FAS cycles) after `warmup`.  Then
    `prof_steps` cycles with every fine-level hot-kernel launch bracketed by HIP
    events on the case's stream (same kernels, same data)."""
    from cfd_solver_amd._lib import CFD_FP64, CFD_MIXED
    from cfd_solver_amd.cavity import CavityCase, cells_per_outer, fas_hierarchy
    case = CavityCase(grid, precision=CFD_MIXED if mixed else CFD_FP64, **PHYS)
    for _ in range(max(warmup, 1)):
        case.step()
    t0 = time.perf_counter()
    st = case.step(steps)  # the K outer iterations in one call of the C driver (cfd_case_outer)
    elapsed = time.perf_counter() - t0
    case.profile(True)
    case.step()
    case.profile(2)
    t1 = time.perf_counter()
    for _ in range(prof_steps):
        case.step()
    prof_ms = (time.perf_counter() - t1) / prof_steps * 1e3
    prof = case.profile_read()
    case.profile(False)
    n0, n1 = grid.n0, grid.n1
    cells = cells_per_outer(n0, n1, case.params)
    value = cells * steps / elapsed
    ms_step = elapsed / steps * 1e3
    name, (ms_tot, launches) = max(prof.items(), key=lambda kv: kv[1][0])
    if launches <= 0 or ms_tot <= 0.0:
        raise RuntimeError(f"live kernel timing recorded nothing: {prof}")
    avg_ms = ms_tot / launches
    bytes_launch = int(kernel_bytes(name, mixed) * n0 * n1)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    levels = fas_hierarchy(n0, n1, case.params)
    # algorithmic bytes of the stage-kernel levels (a level of at most 33 x 33 nodes runs
    # from LDS in one workgroup): the finest level's p' solve takes the BiCGStab
    # iterations of the last statistics, the coarse levels' at most fas_coarse_pmaxit
    # (< 0: one V-cycle)
    pc = case.params.fas_coarse_pmaxit if case.params.fas_coarse_pmaxit != 0 else case.params.p_maxit
    cycle_bytes = sum(a * b * k * simple_bytes_per_node(mixed, max(st.p_iters, 1) if i == 0 else pc)
                      for i, (a, b, k) in enumerate(levels) if a > 33)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": hbm_traffic(mixed, name, n0, n1), "kernel": name,
            "avg_launch_ms": round(avg_ms, 5), "launches": launches, "bytes_per_launch": bytes_launch,
            "timed_in": f"{prof_steps} cycles right after the timed steps (fine level only)",
            "profiled_ms_per_step": round(prof_ms, 4),
            "per_kernel_ms_per_step": {k: round(v[0] / prof_steps, 4) for k, v in prof.items()},
            "per_kernel_GBs": {k: round(kernel_bytes(k, mixed) * n0 * n1 / (v[0] / v[1] * 1e-3) / 1e9, 1)
                               for k, v in prof.items() if v[1] > 0 and v[0] > 0},
            "whole_cycle": {"bytes_model": round(cycle_bytes), "GBs": round(cycle_bytes / (ms_step * 1e-3) / 1e9, 1),
                            "frac": round(cycle_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                            "fine_level_share_of_cells": round((n0 - 2) * (n1 - 2) * levels[0][2] / cells, 4)}}
    info = {"levels": [list(l) for l in levels], "cells_per_step": cells, "mass_res_last": st.mass_res,
            "p_iters_last": st.p_iters, "timed_cycles": f"{warmup + 1}..{warmup + steps} of a solve from rest"}
    case.close()
    return value, ms_step, roof, info


def solve_to_tolerance(grid, mixed, tol=1e-8, cap=300):
    """Fresh case from rest to mass and energy residual ratios <= tol: cycles, seconds
    (the first cycle, which times the multigrid block sizes, included), fields."""
    from cfd_solver_amd._lib import CFD_FP64, CFD_MIXED
    from cfd_solver_amd.cavity import CavityCase
    case = CavityCase(grid, precision=CFD_MIXED if mixed else CFD_FP64, **PHYS)
    t0 = time.perf_counter()
    for k in range(1, cap + 1):
        st = case.step()
        if st.mass_res <= tol and st.energy_res <= tol:
            break
    el = time.perf_counter() - t0
    f = case.fields()
    case.close()
    return {"cycles": k, "seconds": el, "converged": bool(st.mass_res <= tol and st.energy_res <= tol),
            "mass_res": st.mass_res, "energy_res": st.energy_res, "nu_hot": st.nu_hot, "nu_cold": st.nu_cold}, f


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_solve(grid):
    """FP64 and mixed solves of one case from rest to 1e-8, and the per-field relative
    L2 of the mixed fields against the FP64 ones."""
    out = {}
    fields = {}
    for prec in ("fp64", "mixed"):
        out[prec], fields[prec] = solve_to_tolerance(grid, prec == "mixed")
        print(f"[bench] solve {prec}: {out[prec]['cycles']} cycles {out[prec]['seconds']:.3f} s", file=sys.stderr,
              flush=True)
    a, b = fields["mixed"], fields["fp64"]
    err = {k: rel_l2(a[k], b[k]) for k in ("u", "v", "T")}
    pa, pb = a["p"] - a["p"][1:-1, 1:-1].mean(), b["p"] - b["p"][1:-1, 1:-1].mean()
    err["p"] = rel_l2(pa, pb)
    err["Nu"] = abs(out["mixed"]["nu_hot"] - out["fp64"]["nu_hot"]) / abs(out["fp64"]["nu_hot"])
    out["mixed_vs_fp64_rel_l2"] = err
    out["mixed_speedup_time_to_solution"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    out["tolerance"] = 1e-8
    return out


def run_batched(mixed, cap=150, tol=1e-8):
    """The 256-case mix as batches of one grid size and multigrid depth, on their own
    streams (graph replay), every case to tol or its cap."""
    from collections import defaultdict
    from cfd_solver_amd._lib import CFD_FP64, CFD_MIXED, load
    import ctypes as C
    from cfd_solver_amd.cavity import CavityBatch, cells_per_outer, default_params
    from cfd_solver_amd.grid import cavity_grid
    from cfd_solver_amd.workloads import batched_cases
    cases = batched_cases()
    L = load()
    groups = defaultdict(list)
    for idx, (n, skew, phys) in enumerate(cases):
        # the coarsest multigrid level a case gets depends on Re, Ri and Pr
        p = default_params(precision=CFD_MIXED if mixed else CFD_FP64, graph=1, **phys)
        groups[(n, L.cfd_fas_coarsest(n, n, C.byref(p)))].append((idx, skew, phys))
    batches = []
    for (n, depth), members in sorted(groups.items()):
        prm = default_params(precision=CFD_MIXED if mixed else CFD_FP64, graph=1, fas_min_n=depth)
        g0 = cavity_grid(n)
        b = CavityBatch(n, n, g0.dxi, g0.deta, len(members), prm)
        for z, (_, skew, phys) in enumerate(members):
            b.set_case(z, cavity_grid(n, skew=skew), **phys)
        batches.append(dict(b=b, n=n, depth=depth, live=set(range(len(members))), members=members,
                            cells=[cells_per_outer(n, n, default_params(fas_min_n=depth, **ph))
                                   for _, _, ph in members]))
    states = {}
    cycles = {}
    cell_its = 0.0
    t0 = time.perf_counter()
    step = 0
    while any(B["live"] for B in batches):
        step += 1
        live = [B for B in batches if B["live"]]
        for B in live:
            B["b"].launch()
        for B in live:
            st = B["b"].finish()
            for z in list(B["live"]):
                s = st[z]
                cell_its += B["cells"][z]
                idx = B["members"][z][0]
                fin = None
                if not np.isfinite(s.mass_l1) or not np.isfinite(s.energy_l1):
                    fin = "failed"
                elif s.mass_res <= tol and s.energy_res <= tol:
                    fin = "done"
                elif s.iter >= cap:
                    fin = "notconv"
                if fin:
                    states[idx] = fin
                    cycles[idx] = s.iter
                    B["live"].discard(z)
                    B["b"].release(z)
    el = time.perf_counter() - t0
    nbytes = sum(B["b"].nbytes() for B in batches)
    for B in batches:
        B["b"].close()
    counts = {k: sum(1 for v in states.values() if v == k) for k in ("done", "notconv", "failed")}
    done_cycles = [cycles[i] for i, v in states.items() if v == "done"]
    return {"value": cell_its / el, "unit": "cell-iter/s", "seconds": el, "cases": len(cases), "states": counts,
            "cycles_done_median": float(np.median(done_cycles)) if done_cycles else None,
            "cycle_cap": cap, "tolerance": tol, "launch_rounds": step, "hbm_bytes": nbytes,
            "batches": {f"{n}/{d}": len(v) for (n, d), v in sorted(groups.items())},
            "precision": "mixed" if mixed else "fp64",
            "workload": "256 cases, n log-uniform in [65, 513] (2^k+1), A ~ U[0, 0.15], Re ~ U[50, 1000], "
                        "Ri log-uniform [0.1, 10], Pr in {0.71, 7}, seed 4; batches keyed by size / coarsest "
                        "multigrid level (\"n/depth\"), each case to mass and energy residual 1e-8 or 150 cycles; "
                        "cases heated from below at these Rayleigh numbers often have no steady state"}


def pack_slots(n_top, chains):
    """Slots per slab of a pack whose largest slab is n_top x n_top, for cases given
    as (n, coarsest) chains: each case takes one slot of every slab from n down to
    its coarsest level."""
    sizes = []
    n = n_top
    while n >= 5:
        sizes.append(n)
        if not n & 1:
            break
        n = (n + 1) // 2
    lo = min(c for _, c in chains)
    sizes = [m for m in sizes if m >= lo]
    return sizes, [sum(1 for n, c in chains if c <= m <= n) for m in sizes]


def run_packed(mixed, cap=150, tol=1e-8, graph=1):
    """The 256-case mix in one pack (every launch of a cycle serves every case that
    has a grid of that size), graph replay, every case to tol or its cap."""
    import ctypes as C
    from cfd_solver_amd._lib import CFD_FP64, CFD_MIXED, load
    from cfd_solver_amd.cavity import CavityPack, cells_per_outer, default_params
    from cfd_solver_amd.grid import cavity_grid
    from cfd_solver_amd.workloads import batched_cases
    cases = batched_cases()
    L = load()
    prm = default_params(precision=CFD_MIXED if mixed else CFD_FP64, graph=graph)
    from cfd_solver_amd.cavity import fas_hierarchy
    chains = []
    for n, skew, phys in cases:
        p = default_params(precision=prm.precision, **phys)
        chains.append((n, L.cfd_fas_coarsest(n, n, C.byref(p))))
    n_top = max(n for n, _ in chains)
    pc = prm.fas_coarse_pmaxit if prm.fas_coarse_pmaxit != 0 else prm.p_maxit
    sizes, slots = pack_slots(n_top, chains)
    g0 = cavity_grid(n_top)
    pk = CavityPack(n_top, n_top, g0.dxi, g0.deta, slots, prm)
    ids, cells = {}, {}
    mix_bytes = 0.0  # algorithmic bytes of one cycle of every case (stage-kernel levels)
    for idx, (n, skew, phys) in enumerate(cases):
        cid = pk.add_case(cavity_grid(n, skew=skew), **phys)
        ids[cid] = idx
        cells[cid] = cells_per_outer(n, n, default_params(**phys))
        for i, (a, b, k) in enumerate(fas_hierarchy(n, n, default_params(**phys))):
            if a > 33:  # levels of at most 33 x 33 nodes run from LDS in one workgroup
                mix_bytes += a * b * k * simple_bytes_per_node(mixed, prm.p_maxit if i == 0 else pc)
    live = set(ids)
    states, cycles = {}, {}
    cell_its = 0.0
    step = 0
    round_ms = []
    t0 = time.perf_counter()
    while live:
        step += 1
        tr = time.perf_counter()
        pk.launch()
        st = pk.finish()
        round_ms.append((time.perf_counter() - tr) * 1e3)
        for cid in list(live):
            s = st[cid]
            cell_its += cells[cid]
            fin = None
            if not np.isfinite(s.mass_l1) or not np.isfinite(s.energy_l1):
                fin = "failed"
            elif s.mass_res <= tol and s.energy_res <= tol:
                fin = "done"
            elif s.iter >= cap:
                fin = "notconv"
            if fin:
                states[ids[cid]] = fin
                cycles[ids[cid]] = s.iter
                live.discard(cid)
                pk.release(cid)
    el = time.perf_counter() - t0
    nbytes = pk.nbytes()
    pk.close()
    counts = {k: sum(1 for v in states.values() if v == k) for k in ("done", "notconv", "failed")}
    done_cycles = [cycles[i] for i, v in states.items() if v == "done"]
    return {"value": cell_its / el, "unit": "cell-iter/s", "seconds": el, "cases": len(cases), "states": counts,
            "cycles_done_median": float(np.median(done_cycles)) if done_cycles else None,
            "cycle_cap": cap, "tolerance": tol, "launch_rounds": step, "hbm_bytes": nbytes,
            "slabs": {str(m): k for m, k in zip(sizes, slots)},
            "full_mix_cycle_ms": float(np.median(round_ms[1:4])),
            "full_mix_roofline": {
                "bytes_model": round(mix_bytes), "GBs": round(mix_bytes / (np.median(round_ms[1:4]) * 1e-3) / 1e9, 1),
                "frac": round(mix_bytes / (np.median(round_ms[1:4]) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "note": "cycles 2-4 (every case live); algorithmic bytes of the SIMPLE iterations on levels above "
                        "33 x 33 (the p' solve at its cap, 2 BiCGStab iterations, on a case's own grid and one "
                        "V-cycle on its coarse levels) over the cycle's wall time"},
            "precision": "mixed" if mixed else "fp64",
            "workload": "the same 256 cases as one pack: slabs of 513, 257, ..., 17 nodes per side, each case "
                        "holding a slot in every slab from its own size down to its coarsest multigrid level; "
                        "one FAS cycle of every live case per launch round (one captured graph), each case to "
                        "mass and energy residual 1e-8 or 150 cycles"}


def run_served(mixed, lam, count, kmin=6, kmax=10, tol=1e-8, slots=8, max_outer=150, packed=False, solo=True):
    """Poisson-arriving solves to tol, served by continuous batching (size-class
    batches, or one pack when packed), then the same requests solved one at a time
    on the plain FP64 path."""
    from cfd_solver_amd._lib import CFD_FP64, CFD_MIXED
    from cfd_solver_amd.cavity import CavityCase, default_params
    from cfd_solver_amd.grid import cavity_grid
    from cfd_solver_amd.serve import CavityServer
    from cfd_solver_amd.workloads import served_cases
    arrivals = served_cases(count, lam, kmin=kmin, kmax=kmax, re=(50.0, 1000.0), ri=(0.1, 10.0), pr=(0.71, 7.0))
    grids = [cavity_grid(n, skew=skew) for _, n, skew, _ in arrivals]
    prm = default_params(precision=CFD_MIXED if mixed else CFD_FP64, graph=1)
    if packed:
        # slabs 2^kmax + 1 down to 17: `slots` cases of each size at once, every larger
        # case's chain through the smaller slabs on top
        top = 2 ** kmax + 1
        g0 = cavity_grid(top)
        nsl = kmax - 3
        pslots = [slots * min(k + 1, kmax - kmin + 1) for k in range(nsl)]
        srv = CavityServer(prm, pack=(top, top, g0.dxi, g0.deta, pslots))
    else:
        srv = CavityServer(prm, hbm_budget=96 << 30, slots_per_batch=slots)
    # warm a graph of every size class outside the timed run
    warm = {srv.submit(cavity_grid(2 ** k + 1), tol=tol, max_outer=2) for k in range(kmin, kmax + 1)}
    while warm:
        srv.step()
        warm = {t for t in warm if srv.poll(t) is None}
    done, lat, its, states = 0, [], [], {}
    pending = {}
    nxt = 0
    busy = 0.0
    t0 = time.perf_counter()
    while done < count:
        now = time.perf_counter() - t0
        while nxt < count and arrivals[nxt][0] <= now:
            _, n, skew, phys = arrivals[nxt]
            pending[srv.submit(grids[nxt], tol=tol, max_outer=max_outer, **phys)] = nxt
            nxt += 1
        if not pending:
            time.sleep(max(0.0, min(arrivals[nxt][0] - now, 0.01)))
            continue
        ts = time.perf_counter()
        srv.step()
        busy += time.perf_counter() - ts
        for t in list(pending):
            r = srv.poll(t)
            if r is not None:
                del pending[t]
                done += 1
                lat.append(r["latency"])
                its.append(r["stats"].iter)
                states[r["state"]] = states.get(r["state"], 0) + 1
    el = time.perf_counter() - t0
    hbm = srv.info()["bytes"]
    srv.close()
    lat = np.sort(np.array(lat))
    if not solo:
        return {"solves_per_s": count / el, "latency_p50_s": float(np.percentile(lat, 50)),
                "latency_p99_s": float(np.percentile(lat, 99)), "mean_outer_iterations": float(np.mean(its)),
                "requests": count, "arrival_rate_per_s": lam, "busy_fraction": busy / el, "states": states,
                "elapsed_s": el, "hbm_bytes": hbm, "pack_slots": pslots if packed else None}
    # the same requests, one at a time, plain FP64 path (no batching, no graphs), each
    # to the same tolerance or cap, back to back
    t1 = time.perf_counter()
    solo_states = {}
    for k, (_, n, skew, phys) in enumerate(arrivals):
        c = CavityCase(grids[k], precision=CFD_FP64, **phys)
        st, ok = c.solve(tol=tol, max_outer=max_outer)
        s = "done" if ok else ("failed" if not np.isfinite(st.mass_l1) else "notconv")
        solo_states[s] = solo_states.get(s, 0) + 1
        c.close()
    solo = time.perf_counter() - t1
    sizes = {}
    for _, n, _, _ in arrivals:
        sizes[str(n)] = sizes.get(str(n), 0) + 1
    return {"solves_per_s": count / el, "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)), "mean_outer_iterations": float(np.mean(its)),
            "requests": count, "arrival_rate_per_s": lam, "busy_fraction": busy / el, "states": states,
            "elapsed_s": el, "tolerance": tol, "precision": "mixed" if mixed else "fp64", "sizes": sizes,
            "one_at_a_time_fp64": {"solves_per_s": count / solo, "seconds": solo, "states": solo_states},
            "hbm_bytes": hbm,
            "workload": f"Poisson arrivals (seed 5) at {lam}/s, n log-uniform over 2^k+1, k in [{kmin}, {kmax}], "
                        "A ~ U[0, 0.15], Re ~ U[50, 1000], Ri log-uniform [0.1, 10], Pr in {0.71, 7}; each solve "
                        f"to mass and energy residual {tol:g} or {max_outer} cycles; {slots} slots per batch class"}


def p_isolate_bytes_per_node(mixed, refinements):
    """Algorithmic bytes per node of one BiCGStab iteration of the isolated p' solve
    (as simple_bytes_per_node's per-iteration part) and of one FP64 refinement step."""
    w = 4 if mixed else 8
    per_it = 2 * (kernel_bytes("mg_down", mixed) + kernel_bytes("mg_up", mixed)) * 4 / 3 + 15 * w
    return per_it, (13 * 8 if mixed else 0) * refinements


def run_p_isolate(n=2049, reps=3):
    """p' solve alone to 1e-10 on an n x n grid (skew 0.15), FP64 vs mixed, device-resident."""
    from cfd_solver_amd._lib import CFD_FP64, CFD_MIXED
    from cfd_solver_amd.cavity import CavityCase
    from cfd_solver_amd.grid import cavity_grid
    from cfd_solver_amd.workloads import p_isolate
    g = cavity_grid(n, skew=0.15)
    u, v, q = p_isolate(g)
    out = {"grid": [n, n], "rtol": 1e-10,
           "workload": "u, v = 8 Fourier modes each (seed 20250101) setting d; q ~ N(0, 1) minus its mean"}
    for name, prec in (("fp64", CFD_FP64), ("mixed", CFD_MIXED)):
        c = CavityCase(g, precision=prec)
        c.set_fields(u, v, np.zeros_like(u), np.zeros_like(u))
        c.residuals()
        _, its, rr = c.p_solve(None, q, rtol=1e-10, maxit=400, want_x=False)
        t0 = time.perf_counter()
        for _ in range(reps):
            _, its, rr = c.p_solve(None, None, rtol=1e-10, maxit=400, want_x=False)
        ms = (time.perf_counter() - t0) / reps * 1e3
        per_it, refine = p_isolate_bytes_per_node(name == "mixed", 3 if name == "mixed" else 0)
        nbytes = (per_it * its + refine) * n * n
        out[name] = {"ms_per_solve": ms, "iterations": its, "relres": rr,
                     "algorithmic_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
                     "frac_of_peak": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        c.close()
    out["mixed_speedup"] = out["fp64"]["ms_per_solve"] / out["mixed"]["ms_per_solve"]
    out["bytes_model"] = ("per BiCGStab iteration and node: two V-cycles (fine down leg 7.25 words, fused up leg "
                          "and operator 9.75 words, coarser levels a third more) and 15 words of vector updates; "
                          "mixed: FP32 words plus 13 FP64 words per refinement step (3 at 1e-10)")
    return out


def run_ref_ogrid(n0=350, n1=570, steps=10, warmup=2, cpu_steps=3, cpu_runs=3, no_cpu=False):
    """The reference solver's own O-grid time step (solver.cpp:1057-1852), resident on the
    GPU, from the reference's initial state on an n0 x n1 O-grid (the reference's array
    bounds, solver.cpp:13-14) around an ellipse of aspect ratio 2 (outer radius 20), the
    reference's constants (Re 1000, dt 1e-4).  Every inner solve stops at 0.75e-2 of its
    first residual (2-norm; the reference's Gauss-Seidel and SIP use the L1 norm); the
    CPU oracle's reference-faithful step (SIP + Gauss-Seidel as solver.cpp:2066-2445,
    serial) is timed beside it on the same grid, median of cpu_runs runs."""
    from cfd_solver_amd.grid import ogrid_ellipse
    from cfd_solver_amd.refcase import RefCase, initial_state, ref_params
    g = ogrid_ellipse(n0, n1, ar=2.0, r_out=20.0)
    prm = ref_params(mom_rtol=0.75e-2, p_rtol=0.75e-2, mom_maxit=200, p_maxit=2000)
    st0 = initial_state(g, prm)
    c = RefCase(g, prm)
    c.set_state(**st0)
    c.step(warmup)
    t0 = time.perf_counter()
    s = c.step(steps)
    ms = (time.perf_counter() - t0) / steps * 1e3
    c.close()
    cells = (n0 - 1) * (n1 - 2)
    out = {"grid": [n0, n1], "ms_per_step": ms, "steps": steps, "last_iters": s["iters"], "dmax": s["dmax"],
           "cell_steps_per_s": cells / (ms * 1e-3),
           "workload": "reference O-grid time step, ellipse ar 2, r_out 20, Re 1000, dt 1e-4, "
                       "inner solves to 0.75e-2 relative"}
    if not no_cpu:
        from oracle.oracle import build, lib, ref_step
        build()
        L = lib()
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
        st = dict(st0)
        for _ in range(warmup):
            o = ref_step(gd, ops, st, pd)
            st = {k: o[k] for k in ("u", "v", "T", "p", "up", "vp")}
        runs = []
        for _ in range(cpu_runs):
            t0 = time.perf_counter()
            for _ in range(cpu_steps):
                o = ref_step(gd, ops, st, pd)
                st = {k: o[k] for k in ("u", "v", "T", "p", "up", "vp")}
            runs.append((time.perf_counter() - t0) / cpu_steps * 1e3)
        cms = float(np.median(runs))
        out["cpu_oracle_ms_per_step"] = cms
        out["cpu_cell_steps_per_s"] = cells / (cms * 1e-3)
        out["cpu_sample"] = (f"median of {cpu_runs} runs of {cpu_steps} steps after {warmup}, SIP + Gauss-Seidel "
                             f"as in the reference, 1 of {os.cpu_count()} host threads ({host_model()})")
        out["gpu_over_cpu"] = cms / ms
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); multi-GPU sharding is not "
              "built in this version", file=sys.stderr)
        sys.exit(2)
    from cfd_solver_amd.grid import cavity_grid
    sections = set(a.sections.split(","))
    if a.copy_check > 0:
        import torch
        n = a.copy_check << 18  # doubles
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy check: {a.copy_check} MiB device-to-device", file=sys.stderr, flush=True)
    n = a.n
    grid = cavity_grid(n, skew=0.1, beta=1.5)
    res = {}
    for prec in (["fp64", "mixed"] if a.precision == "both" else [a.precision]):
        res[prec] = run_gpu(grid, prec == "mixed", a.steps, a.warmup)
        print(f"[bench] {prec}: {res[prec][1]:.4f} ms/step", file=sys.stderr, flush=True)
    main_prec = "mixed" if "mixed" in res else "fp64"
    value, ms_step, roof, info = res[main_prec]
    cfg = {"workload": f"steady mixed-convection cavity {n}x{n}, skew 0.1 + tanh 1.5, Re 400, Ri 1, Pr 0.71; "
                       f"one FAS cycle (SIMPLE 0.7/0.3/0.85 smoothing on every level) per step, {main_prec} path",
           "grid": [n, n], "precision": main_prec, **info}
    if "fp64" in res and "mixed" in res:
        cfg["fp64_value"] = res["fp64"][0]
        cfg["fp64_ms_per_step"] = res["fp64"][1]
        cfg["fp64_roofline"] = res["fp64"][2]
        cfg["mixed_over_fp64"] = res["mixed"][0] / res["fp64"][0]
    if "main" in sections:
        cfg["solve"] = run_solve(grid)
        if n > 513:
            cfg["solve_513"] = run_solve(cavity_grid((n + 1) // 2, skew=0.1, beta=1.5))
            print(f"[bench] 513 solves: {cfg['solve_513']['fp64']['cycles']} / {cfg['solve_513']['mixed']['cycles']} "
                  "cycles", file=sys.stderr, flush=True)
    if "batch" in sections or "pack" in sections:
        cfg["batched"] = run_packed(main_prec == "mixed", graph=a.pack_graph)
        print(f"[bench] packed: {cfg['batched']['seconds']:.2f} s {cfg['batched']['states']}", file=sys.stderr,
              flush=True)
    if "batch" in sections:
        cfg["batched_by_size"] = run_batched(main_prec == "mixed")
        print(f"[bench] batched by size: {cfg['batched_by_size']['seconds']:.2f} s "
              f"{cfg['batched_by_size']['states']}", file=sys.stderr, flush=True)
    if "serve" in sections:
        cfg["served"] = run_served(main_prec == "mixed", a.serve_rate, a.serve_count)
        print(f"[bench] served: {cfg['served']['solves_per_s']:.2f} solves/s", file=sys.stderr, flush=True)
        cfg["served"]["packed_server"] = run_served(main_prec == "mixed", a.serve_rate, a.serve_count, packed=True,
                                                    solo=False)
        print(f"[bench] served from one pack: {cfg['served']['packed_server']['solves_per_s']:.2f} solves/s",
              file=sys.stderr, flush=True)
    if "piso" in sections:
        cfg["p_isolate"] = run_p_isolate()
        print("[bench] p' isolate done", file=sys.stderr, flush=True)
    if "ogrid" in sections:
        cfg["ref_ogrid"] = run_ref_ogrid(no_cpu=a.no_cpu)
        print(f"[bench] reference O-grid step: {cfg['ref_ogrid']['ms_per_step']:.2f} ms", file=sys.stderr,
              flush=True)
    copy_gbs = copy_peak()
    if copy_gbs:
        roof["measured_copy_GBs"] = copy_gbs
        roof["frac_of_measured_copy"] = round(roof["achieved"] / copy_gbs, 4)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(grid, a.cpu_cycles)
        print(f"[bench] cpu baseline: {cpu['value']:.3e} cell-iter/s", file=sys.stderr, flush=True)
    out = {
        "metric": METRIC, "value": value, "unit": "cell-iter/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if main_prec == "mixed" else "f64", "data": "synthetic",
        "config": cfg, "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
