#!/usr/bin/env python3
"""Benchmark: HSD wall time to the 1e-8 relative-gap stop test on one MI355X.

Workload (default): block-angular multicommodity flow on a 150x150 4-connected
grid with K=8 commodities (m=269,392, n=804,600, nnz=2,234,968), plain FP64
direct path: every normal-equation solve of HELP's HSD driver goes through the
MI355X backend (ADA^T assembly, supernodal Cholesky, triangular solves).

A "step" is one complete HSD solve (construct, iterate to the stop test,
destruct) - the span HELP times as HSD_Main (Main.cpp:95-97).  The JSON line
reports the mean seconds per step (lower is better), the dominant kernel's
roofline measured with HIP events on the library's stream in one extra,
untimed solve, and the CPU oracle timed on a bounded sample of the same
workload.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"hbm": 8000.0, "mfma_fp64": 78.6, "mfma_fp32": 157.3}   # GB/s, TFLOP/s (MI355X spec)


def kernel_table(L, table):
    rows = []
    for k, name in enumerate(L.KERNEL_NAMES):
        ms, n, w = C.c_double(), C.c_int64(), C.c_double()
        L.lib().help_gpu_backend_kernel_stats(C.byref(table), k, C.byref(ms), C.byref(n), C.byref(w))
        if n.value:
            rows.append(dict(kernel=name, ms=ms.value, launches=n.value, work=w.value,
                             bound=L.KERNEL_BOUND[k]))
    return rows


def roofline(rows, mode, traffic_file):
    dom = max(rows, key=lambda r: r["ms"])
    avg_ms = dom["ms"] / dom["launches"]
    work_per_launch = dom["work"] / dom["launches"]
    if dom["bound"] == "hbm":
        achieved = work_per_launch / (avg_ms * 1e-3) / 1e9
        peak, unit = PEAK["hbm"], "GB/s"
        bound = "hbm"
    else:
        achieved = work_per_launch / (avg_ms * 1e-3) / 1e12
        peak = PEAK["mfma_fp32"] if mode == "fp32" else PEAK["mfma_fp64"]
        unit, bound = "TFLOP/s", "mfma"
    traffic = None
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            t = json.load(f)
        traffic = t.get(dom["kernel"])
    return dict(kernel=dom["kernel"], bound=bound, achieved=round(achieved, 4), peak=peak, unit=unit,
                frac=round(achieved / peak, 5), traffic=traffic, avg_launch_ms=round(avg_ms, 5),
                work_per_launch=work_per_launch, launches=dom["launches"])


def cpu_baseline(inst, n_renew, n_solve, host_driver_s, sample_renews=1, sample_solves=1):
    """CPU FP64 oracle (single thread) on a bounded sample of the same workload:
    construct + `sample_renews` renews + `sample_solves` solves, scaled to the
    GPU run's renew/solve counts; the HSD host driver time is shared."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as O
    from help_amd import LinearEquation
    from help_amd import _lib as L
    be = LinearEquation(symbolic_only=True)
    be.construct(inst)
    perm, _, _ = be.symbolic_arrays()
    be.destruct()
    ob = O.OracleBackend(1e-5, perm)
    rng = np.random.default_rng(0)
    d = np.exp(rng.uniform(-3, 3, inst.n) * np.log(10.0))
    b1 = rng.standard_normal(inst.n)
    b2 = rng.standard_normal(inst.m)
    x1 = np.empty(inst.n)
    x2 = np.empty(inst.m)
    lp = inst.as_struct()
    o = O.lib()
    t0 = time.time()
    o.or_be_construct(C.c_void_p(ob.ptr), C.byref(lp))
    t_con = time.time() - t0
    t0 = time.time()
    for _ in range(sample_renews):
        o.or_be_renew(C.c_void_p(ob.ptr), L.dptr(d))
    t_ren = (time.time() - t0) / sample_renews
    t0 = time.time()
    for _ in range(sample_solves):
        o.or_be_solve(C.c_void_p(ob.ptr), L.dptr(d), L.dptr(b1), L.dptr(b2), L.dptr(x1), L.dptr(x2))
    t_sol = (time.time() - t0) / sample_solves
    ob.close()
    est = t_con + t_ren * n_renew + t_sol * n_solve + host_driver_s
    sample = (f"oracle construct + {sample_renews} renew (assembly + left-looking FP64 Cholesky, "
              f"{t_ren:.2f}s) + {sample_solves} KKT solve ({t_sol:.3f}s) on the same instance, scaled to "
              f"the GPU run's {n_renew} renews and {n_solve} solves plus its {host_driver_s:.2f}s of host "
              f"HSD driver work")
    return dict(value=round(est, 3), unit="s", cores=1, kind="port", sample=sample)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=1, choices=[0, 1, 2])
    ap.add_argument("--mode", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--driver", default="gpu", choices=["gpu", "host"],
                    help="gpu: device-resident HSD vectors; host: HELP-order host driver")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--inverse-solve", type=int, default=None, choices=[-1, 0, 1],
                    help="explicit L11 inverse for supernodes wider than 64 columns: "
                         "0 or 1 on (default), -1 off (block substitution)")
    ap.add_argument("--refine-max", type=int, default=None,
                    help="FP64 refinement iterations per solve (default: the library's, "
                         "0 in the FP64 mode as in UseCholmod.cpp, 20 in the FP32 mode)")
    args = ap.parse_args()
    if args.gpus != 1:
        msg = ("bench.py measures one MI355X: this backend is single-GPU "
               f"(--gpus {args.gpus} is not supported; run with --gpus 1)")
        print(msg, file=sys.stderr)
        print(json.dumps({"error": msg, "n_gpus": args.gpus}))
        sys.exit(2)

    from help_amd import _lib as L
    from help_amd import instances
    from help_amd.hsd import GpuBackend, solve_lp

    t0 = time.time()
    inst = instances.by_config(args.config)
    gen_s = time.time() - t0
    mode = L.MODE_DIRECT_FP32 if args.mode == "fp32" else L.MODE_DIRECT_FP64
    print(f"[bench] config {args.config} m={inst.m} n={inst.n} nnz={inst.nnz} mode={args.mode} "
          f"(generated in {gen_s:.1f}s)", file=sys.stderr, flush=True)

    extra = {}
    if args.refine_max is not None:
        extra = {"refine_fp64" if args.mode == "fp64" else "refine_max": args.refine_max}
    if args.inverse_solve is not None:
        extra["inverse_solve"] = args.inverse_solve
    for w in range(args.warmup):
        gb = GpuBackend(mode=mode, **extra)
        res, _ = solve_lp(inst, gb.table, device=args.driver == "gpu")
        gb.close()
        print(f"[bench] warmup {w}: status {res.status} iters {res.iterations} {res.seconds:.2f}s",
              file=sys.stderr, flush=True)

    L.lib().help_device_synchronize()
    gb = GpuBackend(mode=mode, profile=0, **extra)
    times, results = [], []
    for k in range(args.steps):
        ts = time.time()
        res, _ = solve_lp(inst, gb.table, device=args.driver == "gpu")
        L.lib().help_device_synchronize()
        times.append(time.time() - ts)
        results.append(res)
        print(f"[bench] step {k}: status {res.status} iters {res.iterations} gap {res.rel_gap:.2e} "
              f"{times[-1]:.3f}s (renew {res.seconds_renew:.2f}s, solve {res.seconds_solve:.2f}s)",
              file=sys.stderr, flush=True)
    n_sol, n_ref, max_rel = gb.solve_stats()
    gb.close()
    res = results[-1]
    value = sum(times) / len(times)
    # kernel-class timings come from one more, untimed solve with HIP events
    # around every launch on the library's stream (the events cost ~20% of
    # the wall time, so the timed steps run without them)
    rows, prof_s = [], None
    if not args.no_profile:
        gp = GpuBackend(mode=mode, profile=1, **extra)
        ts = time.time()
        rp, _ = solve_lp(inst, gp.table, device=args.driver == "gpu")
        L.lib().help_device_synchronize()
        prof_s = time.time() - ts
        rows = kernel_table(L, gp.table)
        gp.close()
        print(f"[bench] profiled solve: status {rp.status} iters {rp.iterations} {prof_s:.3f}s",
              file=sys.stderr, flush=True)
    roof = roofline(rows, args.mode, os.path.join(ROOT, "profiles", "traffic.json")) if rows else None
    kernels = {r["kernel"]: dict(ms_total=round(r["ms"], 3), launches=r["launches"],
                                 rate=round(r["work"] / (r["ms"] * 1e-3) / (1e9 if r["bound"] == "hbm" else 1e12), 3),
                                 unit="GB/s" if r["bound"] == "hbm" else "TFLOP/s") for r in rows}
    # driver work outside the backend (construct, vector updates, reductions)
    host_driver = max(0.0, res.seconds - res.seconds_renew - res.seconds_solve)
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] timing the CPU oracle sample...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(inst, res.n_renew, res.n_solve, host_driver)
    names = {0: "random LP m=2000 n=6000 (A=[B|I], 4 nnz/col)",
             1: "multicommodity flow 150x150 grid, K=8 (m=269392, n=804600, nnz=2234968)",
             2: "staircase T=200, 1000x3000 blocks (m=200000, n=800000)"}
    path = {"fp64": "plain FP64 direct path", "fp32": "FP32 factor + FP64 refinement direct path"}[args.mode]
    line = {
        "metric": f"time to 1e-8 rel gap (HSD_Main span), {path} on 1 MI355X",
        "value": round(value, 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * value, 2), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.mode == "fp64" else "f32 factor + f64 refinement",
        "data": "synthetic (deterministic generator, seed 0)",
        "config": {"workload": names[args.config], "mode": args.mode, "solver": "HELP HSD driver",
                   "status": L.HSD_STATUS.get(res.status, str(res.status)), "iterations": res.iterations,
                   "rel_gap": res.rel_gap, "primal_obj": res.primal_obj,
                   "primal_infeas": res.primal_infeas, "dual_infeas": res.dual_infeas,
                   "ms_per_iteration": round(1e3 * value / max(res.iterations, 1), 3),
                   "renew_s": round(res.seconds_renew, 3), "solve_s": round(res.seconds_solve, 3),
                   "driver": args.driver, "driver_s": round(host_driver, 3),
                   "construct_s": round(res.seconds_construct, 3),
                   "driver_newton_s": round(res.seconds_newton, 3),
                   "driver_direction_s": round(res.seconds_direction, 3),
                   "driver_step_s": round(res.seconds_step, 3),
                   "n_renew": res.n_renew, "n_solve": res.n_solve,
                   "refine_iterations": n_ref, "max_rel_residual": max_rel,
                   "inverse_solve": "off" if args.inverse_solve == -1 else "on",
                   "profiled_solve_s": None if prof_s is None else round(prof_s, 3)},
        "roofline": roof, "cpu_baseline": cpu, "kernels": kernels,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
