#!/usr/bin/env python3
"""Benchmark: SIMPLE outer iterations of the steady mixed-convection cavity on one MI355X.

Workload (default): 1025 x 1025 nodes, skew amplitude 0.1 with tanh wall
clustering beta = 1.5, Re = 400, Ri = 1, Pr = 0.71, FP64, relaxation 0.7/0.3,
8 predictor sweeps, p' BiCGStab + multigrid to 1e-2 per outer iteration.
One step = one outer iteration over the whole case (fields resident in HBM).

    python bench.py [--steps K] [--warmup W] [--n 1025] [--cpu-iters 6] [--no-cpu]

Prints one JSON line (value = cell-iterations/s, a cell = an interior node).
The roofline object is for the kernel with the largest device time per outer
iteration, timed live with HIP events on the case's stream during the timed
steps; its algorithmic bytes per launch are (arrays read + arrays written) x
interior nodes x 8 B.  cpu_baseline times the CPU oracle (a serial FP64 port of
the same iteration) on this host for a bounded number of outer iterations.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SIMPLE cell-iters/s at 1025² FP64 vs mixed; served solves/s + p99 at 1e-8 res"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per node per launch: (arrays read + arrays written) x element size
def kernel_bytes(name, mixed):
    w = 4 if mixed else 8  # predictor and p' inner solve precision
    return {
        "p_apply": 7 * w,          # x, 4 face coefficients, r0 (or s) -> y; dot product in the kernel
        "mg_down": 7.25 * w,       # b, 1/diag, 4 face coefficients -> x, restricted residual (1/4 node)
        "mg_up": 8.25 * w,         # x, b, 1/diag, 4 faces, coarse x (1/4 node) -> x
        "jacobi": 16 * w,          # R_u,R_v,R_T, U,V, alph,beta,gamma,P,Q, du,dv,dT -> du,dv,dT
        "residual": 13 * 8 + 8 + 5 * w,  # u,v,T,p, 4 metrics, alph..Q (FP64) -> d (FP64), R_u,R_v,R_T,U,V
        "face_div": 10 * 8,        # u^,v^,p,d, 4 metrics, J -> q (FP64)
    }[name]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1025)
    ap.add_argument("--cpu-iters", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    return ap.parse_args()


def cpu_baseline(grid, n_iters, kw):
    """Serial CPU oracle on the same workload for a bounded number of outer iterations."""
    from oracle.oracle import CavityOracle, build, cav_params
    build()
    c = CavityOracle(grid, cav_params(**kw))
    c.step()  # first iteration sets the normalisers; not timed
    t0 = time.perf_counter()
    for _ in range(n_iters):
        c.step()
    dt = time.perf_counter() - t0
    cells = (grid.n0 - 2) * (grid.n1 - 2)
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": cells * n_iters / dt, "unit": "cell-iter/s", "cores": 1, "kind": "port",
            "sample": f"{n_iters} outer iterations of the same {grid.n0}x{grid.n1} case, serial FP64 "
                      f"C oracle (gcc -O3), 1 of {os.cpu_count()} host threads ({model})"}


def run_gpu(grid, mixed, steps, warmup, prof_steps=10):
    """Timed region: `steps` outer iterations replayed from the captured hipGraph.
    Then `prof_steps` iterations of the same case run eagerly with every hot-kernel
    launch bracketed by HIP events on the case's stream (same kernels, same data)."""
    from cfd_solver_amd._lib import CFD_FP64, CFD_MIXED
    from cfd_solver_amd.cavity import CavityCase
    case = CavityCase(grid, Re=400.0, Ri=1.0, Pr=0.71, precision=CFD_MIXED if mixed else CFD_FP64)
    for _ in range(max(warmup, 1)):
        case.step()                      # first call captures the graph
    t0 = time.perf_counter()
    st = None
    for _ in range(steps):
        st = case.step()
    elapsed = time.perf_counter() - t0
    case.profile(True)
    case.step()
    case.profile(2)
    t1 = time.perf_counter()
    for _ in range(prof_steps):
        case.step()
    eager_ms = (time.perf_counter() - t1) / prof_steps * 1e3
    prof = case.profile_read()
    case.profile(False)
    n0, n1 = grid.n0, grid.n1
    value = (n0 - 2) * (n1 - 2) * steps / elapsed
    name, (ms_tot, launches) = max(prof.items(), key=lambda kv: kv[1][0])
    if launches <= 0 or ms_tot <= 0.0:
        raise RuntimeError(f"live kernel timing recorded nothing: {prof}")
    avg_ms = ms_tot / launches
    bytes_launch = int(kernel_bytes(name, mixed) * n0 * n1)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": name,
            "avg_launch_ms": round(avg_ms, 5), "launches": launches, "bytes_per_launch": bytes_launch,
            "timed_in": f"{prof_steps} eager outer iterations right after the graph-replayed timed steps",
            "eager_ms_per_step": round(eager_ms, 4),
            "per_kernel_ms_per_step": {k: round(v[0] / prof_steps, 4) for k, v in prof.items()},
            "per_kernel_GBs": {k: round(kernel_bytes(k, mixed) * n0 * n1 / (v[0] / v[1] * 1e-3) / 1e9, 1)
                               for k, v in prof.items() if v[1] > 0 and v[0] > 0}}
    info = {"mass_res_last": st.mass_res, "p_iters_last": st.p_iters, "nu_hot_last": st.nu_hot}
    case.close()
    return value, elapsed / steps * 1e3, roof, info


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); multi-GPU sharding is not "
              "built in this version", file=sys.stderr)
        sys.exit(2)
    from cfd_solver_amd.grid import cavity_grid
    n = a.n
    grid = cavity_grid(n, skew=0.1, beta=1.5)
    res = {}
    for prec in (["fp64", "mixed"] if a.precision == "both" else [a.precision]):
        res[prec] = run_gpu(grid, prec == "mixed", a.steps, a.warmup)
    main_prec = "mixed" if "mixed" in res else "fp64"
    value, ms_step, roof, info = res[main_prec]
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(grid, a.cpu_iters, dict(Re=400.0, Ri=1.0, Pr=0.71))
    cfg = {"workload": f"steady mixed-convection cavity {n}x{n}, skew 0.1 + tanh 1.5, Re 400, Ri 1, Pr 0.71, "
                       f"SIMPLE (0.7/0.3), one outer iteration per step, {main_prec} path",
           "grid": [n, n], "precision": main_prec, **info}
    if "fp64" in res and "mixed" in res:
        cfg["fp64_value"] = res["fp64"][0]
        cfg["fp64_ms_per_step"] = res["fp64"][1]
        cfg["fp64_roofline"] = res["fp64"][2]
        cfg["mixed_over_fp64"] = res["mixed"][0] / res["fp64"][0]
    out = {
        "metric": METRIC, "value": value, "unit": "cell-iter/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if main_prec == "mixed" else "f64", "data": "synthetic",
        "config": cfg, "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
