#!/usr/bin/env python3
"""Benchmark: SIMPLE outer iterations of the steady mixed-convection cavity on one MI355X.

Headline (value): 1025 x 1025 nodes, skew amplitude 0.1 with tanh wall clustering
beta = 1.5, Re = 400, Ri = 1, Pr = 0.71, relaxation 0.7/0.3, 8 predictor sweeps,
p' BiCGStab + multigrid to 1e-2 per outer iteration; mixed path (FP32 predictor
and p' inner solve, FP64 residuals, refinement and accumulation), with the FP64
path beside it.  One step = one outer iteration over the whole case (fields
resident in HBM), launched eagerly with the p' solver's convergence flags read
once per BiCGStab iteration (faster than replaying a captured graph, which must
unroll the iteration cap); the K timed steps are one call of the library's outer
loop, which reads the statistics back after the last.  The batched and served
runs replay captured graphs so the size classes overlap on their streams.

    python bench.py [--steps K] [--warmup W] [--n 1025] [--cpu-iters 6] [--no-cpu]
                    [--sections main,batch,serve,piso,ogrid]

Prints one JSON line (value = cell-iterations/s, a cell = an interior node).
The roofline object is for the kernel with the largest device time per outer
iteration, timed live with HIP events on the case's stream; its algorithmic
bytes per launch are (arrays read + arrays written) x nodes x element size.
cpu_baseline times the CPU oracle (a serial FP64 port of the same iteration) on
this host for a bounded number of outer iterations.  Further objects in config:
  batched   256 cases (65..513 nodes per side, mixed physics) advanced as four
            size-class batches, all slots per launch, vs one case at a time;
  served    continuous batching of Poisson-arriving solves to 1e-8: solves/s,
            latency p50/p99;
  p_isolate the p' solve alone at 2049 x 2049 to 1e-10, FP64 vs mixed;
  ref_ogrid the reference's own O-grid time step resident on the GPU (257 x 513),
            with the CPU oracle's reference-faithful step beside it.
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1025)
    ap.add_argument("--cpu-iters", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--sections", default="main,batch,serve,piso,ogrid")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="served arrivals per second")
    ap.add_argument("--serve-count", type=int, default=200)
    ap.add_argument("--copy-check", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (the known-size kernel that scales "
                         "rocprof's FETCH_SIZE / WRITE_SIZE counters)")
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
    """Timed region: `steps` outer iterations.  Then `prof_steps` iterations of the
    same case with every hot-kernel launch bracketed by HIP events on the case's
    stream (same kernels, same data)."""
    from cfd_solver_amd._lib import CFD_FP64, CFD_MIXED
    from cfd_solver_amd.cavity import CavityCase
    case = CavityCase(grid, Re=400.0, Ri=1.0, Pr=0.71, precision=CFD_MIXED if mixed else CFD_FP64)
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
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": hbm_traffic(mixed, name, n0, n1), "kernel": name,
            "avg_launch_ms": round(avg_ms, 5), "launches": launches, "bytes_per_launch": bytes_launch,
            "timed_in": f"{prof_steps} outer iterations right after the timed steps",
            "profiled_ms_per_step": round(eager_ms, 4),
            "per_kernel_ms_per_step": {k: round(v[0] / prof_steps, 4) for k, v in prof.items()},
            "per_kernel_GBs": {k: round(kernel_bytes(k, mixed) * n0 * n1 / (v[0] / v[1] * 1e-3) / 1e9, 1)
                               for k, v in prof.items() if v[1] > 0 and v[0] > 0}}
    info = {"mass_res_last": st.mass_res, "p_iters_last": st.p_iters, "nu_hot_last": st.nu_hot}
    case.close()
    return value, elapsed / steps * 1e3, roof, info


def _single_ms(n, mixed, steps=15, warmup=3):
    """ms per outer iteration of one n x n case alone."""
    from cfd_solver_amd._lib import CFD_FP64, CFD_MIXED
    from cfd_solver_amd.cavity import CavityCase
    from cfd_solver_amd.grid import cavity_grid
    c = CavityCase(cavity_grid(n, skew=0.075), precision=CFD_MIXED if mixed else CFD_FP64)
    c.step(warmup)
    t0 = time.perf_counter()
    c.step(steps)
    ms = (time.perf_counter() - t0) / steps * 1e3
    c.close()
    return ms


def run_batched(mixed, steps=10, warmup=2):
    """The 256-case mix as one batch per grid size; every step advances all cases."""
    from collections import defaultdict
    from cfd_solver_amd._lib import CFD_FP64, CFD_MIXED
    from cfd_solver_amd.cavity import CavityBatch, default_params
    from cfd_solver_amd.grid import cavity_grid
    from cfd_solver_amd.workloads import batched_cases
    cases = batched_cases()
    by_n = defaultdict(list)
    for n, skew, phys in cases:
        by_n[n].append((skew, phys))
    prm = default_params(precision=CFD_MIXED if mixed else CFD_FP64, graph=1)
    batches = []
    for n in sorted(by_n):
        g0 = cavity_grid(n)
        b = CavityBatch(n, n, g0.dxi, g0.deta, len(by_n[n]), prm)
        for z, (skew, phys) in enumerate(by_n[n]):
            b.set_case(z, cavity_grid(n, skew=skew), **phys)
        batches.append(b)
    for _ in range(warmup):
        for b in batches:
            b.launch()
        for b in batches:
            b.finish()
    t0 = time.perf_counter()
    diverged = 0
    for k in range(steps):
        for b in batches:
            b.launch()
        for b in batches:
            st = b.finish()
            if any(s.iter == 0 for s in st):
                raise RuntimeError("a batched case did not advance")
            if k == steps - 1:
                diverged += sum(1 for s in st if not np.isfinite(s.mass_l1) or not np.isfinite(s.energy_l1))
    el = time.perf_counter() - t0
    cells = sum((n - 2) ** 2 for n, _, _ in cases)
    nbytes = sum(b.nbytes() for b in batches)
    for b in batches:
        b.close()
    single = {n: _single_ms(n, mixed) for n in sorted(by_n)}
    alone_ms = sum(single[n] * len(by_n[n]) for n in by_n)
    ms = el / steps * 1e3
    return {"value": cells * steps / el, "unit": "cell-iter/s", "ms_per_step": ms, "cases": len(cases),
            "nodes": sum(n * n for n, _, _ in cases), "size_classes": {str(n): len(v) for n, v in sorted(by_n.items())},
            "hbm_bytes": nbytes, "non_finite_cases": diverged, "one_at_a_time_ms_per_step": alone_ms, "speedup_vs_one_at_a_time": alone_ms / ms,
            "single_case_ms_per_step": {str(n): round(v, 4) for n, v in single.items()},
            "precision": "mixed" if mixed else "fp64",
            "workload": "256 cases, n log-uniform in [65, 513] (2^k+1), A ~ U[0, 0.15], Re ~ U[50, 1000], "
                        "Ri log-uniform [0.1, 10], Pr in {0.71, 7}, seed 4; one step = one outer iteration of "
                        "every case"}


def run_served(mixed, lam, count, kmin=5, kmax=6, tol=1e-8, slots=32, max_outer=10000):
    """Poisson-arriving solves to tol, served by continuous batching."""
    from cfd_solver_amd._lib import CFD_FP64, CFD_MIXED
    from cfd_solver_amd.cavity import default_params
    from cfd_solver_amd.grid import cavity_grid
    from cfd_solver_amd.serve import CavityServer
    from cfd_solver_amd.workloads import served_cases
    arrivals = served_cases(count, lam, kmin=kmin, kmax=kmax)
    grids = [cavity_grid(n, skew=skew) for _, n, skew, _ in arrivals]
    srv = CavityServer(default_params(precision=CFD_MIXED if mixed else CFD_FP64, graph=1), hbm_budget=64 << 30,
                       slots_per_batch=slots)
    # warm the graphs of both size classes outside the timed run
    warm = {srv.submit(cavity_grid(2 ** k + 1), tol=tol, max_outer=3) for k in range(kmin, kmax + 1)}
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
    srv.close()
    lat = np.sort(np.array(lat))
    return {"solves_per_s": count / el, "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)), "mean_outer_iterations": float(np.mean(its)),
            "requests": count, "arrival_rate_per_s": lam, "busy_fraction": busy / el, "states": states,
            "elapsed_s": el, "tolerance": tol, "precision": "mixed" if mixed else "fp64",
            "workload": f"Poisson arrivals (seed 5) at {lam}/s, n log-uniform over 2^k+1, k in [{kmin}, {kmax}], "
                        "A ~ U[0, 0.15], Re ~ U[50, 150], Ri log-uniform [0.1, 1], Pr 0.71; each solve to mass "
                        f"and energy residual {tol:g} (cap {max_outer} outer iterations); {slots} slots per size "
                        "class"}


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
        out[name] = {"ms_per_solve": ms, "iterations": its, "relres": rr}
        c.close()
    out["mixed_speedup"] = out["fp64"]["ms_per_solve"] / out["mixed"]["ms_per_solve"]
    return out


def run_ref_ogrid(n0=257, n1=513, steps=10, warmup=2, cpu_steps=2, no_cpu=False):
    """The reference solver's own O-grid time step (solver.cpp:1057-1852), resident on the
    GPU, from the reference's initial state on an n0 x n1 O-grid around an ellipse of
    aspect ratio 2 (outer radius 20), the reference's constants (Re 1000, dt 1e-4).
    Every inner solve stops at 0.75e-2 of its first residual (2-norm; the reference's
    Gauss-Seidel and SIP use the L1 norm); the CPU oracle's reference-faithful step
    (SIP + Gauss-Seidel, serial) is timed beside it on the same grid."""
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
    out = {"grid": [n0, n1], "ms_per_step": ms, "steps": steps, "last_iters": s["iters"], "dmax": s["dmax"],
           "workload": "reference O-grid time step, ellipse ar 2, r_out 20, Re 1000, dt 1e-4, "
                       "inner solves to 0.75e-2 relative"}
    if not no_cpu:
        import ctypes as C
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
        t0 = time.perf_counter()
        for _ in range(cpu_steps):
            o = ref_step(gd, ops, st, pd)
            st = {k: o[k] for k in ("u", "v", "T", "p", "up", "vp")}
        cms = (time.perf_counter() - t0) / cpu_steps * 1e3
        out["cpu_oracle_ms_per_step"] = cms
        out["cpu_sample"] = f"{cpu_steps} steps after {warmup}, SIP + Gauss-Seidel as in the reference, 1 thread"
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
    cfg = {"workload": f"steady mixed-convection cavity {n}x{n}, skew 0.1 + tanh 1.5, Re 400, Ri 1, Pr 0.71, "
                       f"SIMPLE (0.7/0.3), one outer iteration per step, {main_prec} path",
           "grid": [n, n], "precision": main_prec, **info}
    if "fp64" in res and "mixed" in res:
        cfg["fp64_value"] = res["fp64"][0]
        cfg["fp64_ms_per_step"] = res["fp64"][1]
        cfg["fp64_roofline"] = res["fp64"][2]
        cfg["mixed_over_fp64"] = res["mixed"][0] / res["fp64"][0]
    if "batch" in sections:
        cfg["batched"] = run_batched(main_prec == "mixed")
        print(f"[bench] batched: {cfg['batched']['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    if "serve" in sections:
        cfg["served"] = run_served(main_prec == "mixed", a.serve_rate, a.serve_count)
        print(f"[bench] served: {cfg['served']['solves_per_s']:.2f} solves/s", file=sys.stderr, flush=True)
    if "piso" in sections:
        cfg["p_isolate"] = run_p_isolate()
        print("[bench] p' isolate done", file=sys.stderr, flush=True)
    if "ogrid" in sections:
        cfg["ref_ogrid"] = run_ref_ogrid(no_cpu=a.no_cpu)
        print(f"[bench] reference O-grid step: {cfg['ref_ogrid']['ms_per_step']:.2f} ms", file=sys.stderr,
              flush=True)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(grid, a.cpu_iters, dict(Re=400.0, Ri=1.0, Pr=0.71))
    out = {
        "metric": METRIC, "value": value, "unit": "cell-iter/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if main_prec == "mixed" else "f64", "data": "synthetic",
        "config": cfg, "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
