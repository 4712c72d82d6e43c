#!/usr/bin/env python3
"""Benchmark: HSD wall time to the 1e-8 relative-gap stop test on one MI355X,
mixed precision against plain FP64.

Default workload: the staircase LP with T=200 periods (1000 x 3000 blocks,
5 N(0,1) nnz per column in-block plus one linking nnz into the next period,
plus slack identity: m=200,000, n=800,000, nnz=3,797,000).  Every
normal-equation solve of HELP's HSD driver goes through the MI355X backend.
Three paths run on the same instance, each as W untimed and K timed steps:

* plain_fp64_amd  FP64 factor, AMD ordering: CHOLMOD's semantics in HELP
                  (UseCholmod.cpp, Setting_CPU.cpp:21-26), one factor solve per
                  system;
* plain_fp64_nd   FP64 factor, nested-dissection ordering, one factor solve;
* mixed_fp32_nd   FP32 factor (MFMA, FP32 accumulation) + FP64 conjugate
                  gradients preconditioned by it, nested dissection.

Other workloads (``--config``): 0 a random LP, 1 the 150x150 K=8
multicommodity-flow grid (plain FP64 vs mixed, AMD), 3 the Zipf LP with
m=10^6 in PCG mode (pcg_fp64: FP64 partial-Cholesky apply; pcg_fp32: FP32
apply, the value), 4 a held-out instance of the warm-start family in PCG
mode (pcg_zero: CG from zero; pcg_learned: CG from the committed network's
guess, help_amd/data/warmstart_net_mcf.npz, the value).

A "step" is one complete HSD solve (construct, iterate to the stop test,
destruct): the span HELP times as HSD_Main (Main.cpp:95-97).  `value` is the
mixed path's mean seconds per step; the plain times and the speedups sit
beside it.  Also reported: the dominant kernel's roofline, measured with HIP
events on the library's stream in one extra untimed solve of the value path;
the dy error of the mixed path against the plain FP64 path on solves taken
from its own trajectory; and MKL PARDISO on all host cores (HELP's
UsePardiso.cpp semantics) timed on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"hbm": 8000.0, "mfma_fp64": 78.6, "mfma_fp32": 157.3}   # GB/s, TFLOP/s (MI355X spec)

# name -> backend mode and options
PATHS = {
    "plain_fp64_amd": dict(mode="fp64", ordering=0),
    "plain_fp64_nd": dict(mode="fp64", ordering=1),
    "mixed_fp32_amd": dict(mode="fp32", ordering=0),
    "mixed_fp32_nd": dict(mode="fp32", ordering=1),
    "pcg_fp64": dict(mode="pcg", pcg_fp32_precond=0),
    "pcg_fp32": dict(mode="pcg", pcg_fp32_precond=1),
    "pcg_zero": dict(mode="pcg", pcg_fp32_precond=1),
    "pcg_learned": dict(mode="pcg_learned", pcg_fp32_precond=1),
}
DEFAULT_PATHS = {0: "plain_fp64_amd,mixed_fp32_amd", 1: "plain_fp64_amd,mixed_fp32_amd",
                 2: "plain_fp64_amd,plain_fp64_nd,mixed_fp32_nd", 3: "pcg_fp64,pcg_fp32",
                 4: "pcg_zero,pcg_learned"}
NAMES = {0: "random LP m=2000 n=6000 (A=[B|I], 4 nnz/col)",
         1: "multicommodity flow 150x150 grid, K=8 (m=269392, n=804600, nnz=2234968)",
         2: "staircase T=200, 1000x3000 blocks (m=200000, n=800000, nnz=3797000)",
         3: "Zipf LP m=1000000, n=4000000 (6 nnz/col, s=1.1), PCG with K=500",
         4: "multicommodity flow 76x76 grid, K=2 (held-out instance of the warm-start family, seed 2002), "
            "PCG with K=50"}
WORKLOAD_KEYS = {0: "random2000", 1: "mcf150x8", 2: "staircase200", 3: "zipf1e6", 4: "mcf76x2"}
NET = os.path.join(ROOT, "help_amd", "data", "warmstart_net_mcf.npz")


def mode_const(L, mode_s):
    return {"fp64": L.MODE_DIRECT_FP64, "fp32": L.MODE_DIRECT_FP32, "pcg": L.MODE_PCG,
            "pcg_learned": L.MODE_PCG_LEARNED}[mode_s]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_table(L, table, fp32_only=False):
    """Per kernel class: total ms (HIP events on the library's stream),
    launches, algorithmic work; fp32_only keeps the launches that ran on the
    FP32 factor."""
    rows = []
    for k, name in enumerate(L.KERNEL_NAMES):
        ms, n, w = C.c_double(), C.c_int64(), C.c_double()
        L.lib().help_gpu_backend_kernel_stats_prec(C.byref(table), k, 1 if fp32_only else 0, C.byref(ms),
                                                   C.byref(n), C.byref(w))
        if n.value:
            rows.append(dict(kernel=name, ms=ms.value, launches=n.value, work=w.value,
                             bound=L.KERNEL_BOUND[k]))
    return rows


def roofline(rows, mode, traffic_file, workload_key):
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
            traffic = json.load(f).get(workload_key, {}).get(dom["kernel"])
    return dict(kernel=dom["kernel"], bound=bound, achieved=round(achieved, 4), peak=peak, unit=unit,
                frac=round(achieved / peak, 5), traffic=traffic, avg_launch_ms=round(avg_ms, 5),
                work_per_launch=work_per_launch, launches=dom["launches"])


def host_cores():
    """Threads the CPU baseline uses: OMP_NUM_THREADS when set (a machine
    shared between GPUs sets it to one GPU's share of the host cores), else
    the affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(inst, gpu_iterations, sample_iters=1):
    """HELP's PARDISO backend (UsePardiso.cpp semantics over MKL PARDISO from
    libmkl_rt, oracle/libpardiso_be.so) under the same HSD driver on all host
    cores, for `sample_iters` iterations (+ the initial point); the full run is
    estimated as construct + per-iteration time x the GPU run's iterations."""
    from help_amd import _lib as L
    from help_amd.hsd import solve_lp
    so = os.path.join(ROOT, "oracle", "libpardiso_be.so")
    if not os.path.exists(so):
        return None
    pb = C.CDLL(so)
    pb.pb_new.restype = C.c_void_p
    pb.pb_new.argtypes = [C.c_int]
    pb.pb_error.restype = C.c_char_p
    pb.pb_stat.restype = C.c_double
    pb.pb_stat.argtypes = [C.c_void_p, C.c_int]
    pb.pb_free.argtypes = [C.c_void_p]
    cores = host_cores()
    ctx = pb.pb_new(cores)
    if not ctx:
        log(f"CPU baseline unavailable: {pb.pb_error().decode()}")
        return None
    table = L.Backend()
    table.ctx = ctx
    for nm in ("construct", "destruct", "renew", "solve"):
        setattr(table, nm, C.cast(getattr(pb, "pb_" + nm), C.c_void_p))
    t0 = time.time()
    res, _ = solve_lp(inst, table, max_iter=sample_iters)
    wall = time.time() - t0
    t_con = pb.pb_stat(ctx, 0)
    n_ren, n_sol = int(pb.pb_stat(ctx, 3)), int(pb.pb_stat(ctx, 4))
    t_ren, t_sol = pb.pb_stat(ctx, 1), pb.pb_stat(ctx, 2)
    pb.pb_free(ctx)
    per_iter = (wall - t_con) / max(n_ren, 1)
    est = t_con + per_iter * (gpu_iterations + 1)
    sample = (f"MKL PARDISO (UsePardiso.cpp semantics: METIS ND ordering, mtype -2, phases 11/22/33) under "
              f"HELP's HSD driver, {cores} threads on {cpu_model()} (nproc {os.cpu_count()}): construct "
              f"{t_con:.2f}s, then {n_ren} renews ({t_ren / max(n_ren, 1):.2f}s each) and {n_sol} solves "
              f"({t_sol / max(n_sol, 1):.3f}s each) in {res.iterations} iteration(s) + the initial point, "
              f"{wall:.1f}s measured; the full run is estimated as construct + {per_iter:.2f}s per iteration "
              f"x ({gpu_iterations} + 1) iterations of the GPU run")
    return dict(value=round(est, 3), unit="s", cores=cores, kind="port", sample=sample,
                measured_s=round(wall, 3))


def path_opts(name, extra, mixed_opts):
    """backend options of a path: its own, the common ones, plus the
    refinement settings for the mixed-precision paths"""
    own = {k: v for k, v in PATHS[name].items() if k != "mode"}
    return {**own, **extra, **(mixed_opts if PATHS[name]["mode"] == "fp32" else {})}


_net = {}


def warmstart_net():
    """The committed warm-start network (help_amd/data), loaded once."""
    if "net" not in _net:
        from help_amd.warmstart import Mpnn
        _net["net"] = Mpnn.load(NET)
    return _net["net"]


def make_backend(name, opts, L, GpuBackend, **more):
    mode_s = PATHS[name]["mode"]
    gb = GpuBackend(mode=mode_const(L, mode_s), **opts, **more)
    if mode_s == "pcg_learned":
        gb.set_warmstart(warmstart_net())
    return gb


def run_path(inst, name, steps, warmup, driver, opts, L, GpuBackend, solve_lp):
    mode_s = PATHS[name]["mode"]
    for w in range(warmup):
        gb = make_backend(name, opts, L, GpuBackend)
        res, _ = solve_lp(inst, gb.table, device=driver == "gpu")
        gb.close()
        log(f"{name} warmup {w}: status {res.status} iters {res.iterations} {res.seconds:.2f}s")
    L.lib().help_device_synchronize()
    gb = make_backend(name, opts, L, GpuBackend, profile=0)
    times, results = [], []
    for k in range(steps):
        ts = time.time()
        res, _ = solve_lp(inst, gb.table, device=driver == "gpu")
        L.lib().help_device_synchronize()
        times.append(time.time() - ts)
        results.append(res)
        log(f"{name} step {k}: status {res.status} iters {res.iterations} gap {res.rel_gap:.2e} "
            f"{times[-1]:.3f}s (construct {res.seconds_construct:.2f}s, renew {res.seconds_renew:.2f}s, "
            f"solve {res.seconds_solve:.2f}s)")
    n_sol, n_ref, max_rel = gb.solve_stats()
    pcg_it = gb.pcg_iterations() if mode_s.startswith("pcg") else None
    gb.close()
    res = results[-1]
    host_driver = max(0.0, res.seconds - res.seconds_renew - res.seconds_solve - res.seconds_construct)
    out = dict(
        seconds=sum(times) / len(times), mode=mode_s,
        status=L.HSD_STATUS.get(res.status, str(res.status)), iterations=res.iterations,
        rel_gap=res.rel_gap, primal_obj=res.primal_obj, primal_infeas=res.primal_infeas,
        dual_infeas=res.dual_infeas, construct_s=round(res.seconds_construct, 3),
        renew_s=round(res.seconds_renew, 3), solve_s=round(res.seconds_solve, 3),
        driver_s=round(host_driver, 3), n_renew=res.n_renew, n_solve=res.n_solve,
        max_rel_residual=max_rel,
        ms_per_iteration=round(1e3 * (sum(times) / len(times)) / max(res.iterations, 1), 3))
    if "ordering" in PATHS[name]:
        out["ordering"] = "nd" if PATHS[name]["ordering"] else "amd"
        out["refine_iterations"] = int(n_ref // max(steps, 1))
    else:
        out["stall_reason"] = res.stall_reason
        out["pcg_iterations"] = int(pcg_it // max(steps, 1))
    return out


def dy_errors(inst, name, capture, extra):
    """dy of the `name` path on the bottom-system solves of chosen iterations
    of its own trajectory (HELP's driver on the host, HSDSolver.cpp:770),
    against the plain FP64 path (AMD ordering, CHOLMOD semantics) on the same
    (d, b1, b2); plus the FP64 backward error of each."""
    import numpy as np
    from help_amd import LinearEquation
    from help_amd import _lib as L
    from help_amd.hsd import RecordingBackend, solve_lp
    mode = mode_const(L, PATHS[name]["mode"])
    recs = {}

    def keep(renews, k, d, b1, b2, x1, x2):
        if k == 0 and renews in capture:
            recs[renews] = (d.copy(), b1.copy(), b2.copy(), x2.copy())

    rb = RecordingBackend(inst, keep, mode=mode, **extra)
    res, _ = solve_lp(inst, rb.table)
    A = inst.to_scipy().tocsr()
    At = A.T.tocsr()
    ref = LinearEquation(mode=L.MODE_DIRECT_FP64, ordering=0)
    ref.construct(inst)
    out = []
    for k in sorted(recs):
        d, b1, b2, x2 = recs[k]
        ref.renew(d)
        _, o2 = ref.solve(d, b1, b2)
        rhs = A @ (b1 / d) - b2
        bwd = lambda x: float(np.linalg.norm(rhs - (A @ ((At @ x) / d) + 1e-5 * x)) / np.linalg.norm(rhs))
        out.append(dict(iteration=k, d_decades=round(float(np.log10(d.max() / d.min())), 2),
                        dy_rel_err_vs_plain_fp64=float(np.linalg.norm(x2 - o2) / np.linalg.norm(o2)),
                        backward_err=bwd(x2), backward_err_plain_fp64=bwd(o2)))
        log(f"dy error {name} iteration {k}: {out[-1]}")
    ref.destruct()
    return out, L.HSD_STATUS.get(res.status, str(res.status)), res.iterations


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, choices=[0, 1, 2, 3, 4])
    ap.add_argument("--paths", default=None,
                    help="comma list of " + ",".join(PATHS) + "; the last one is `value` "
                         "(default per config: " + "; ".join(f"{k}: {v}" for k, v in DEFAULT_PATHS.items()) + ")")
    ap.add_argument("--driver", default="gpu", choices=["gpu", "host"],
                    help="gpu: device-resident HSD vectors; host: HELP-order host driver")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-dy-error", action="store_true")
    ap.add_argument("--refine-tol", type=float, default=None,
                    help="mixed paths: FP64 refinement stops at this relative residual (library default 1e-12)")
    ap.add_argument("--refine-max", type=int, default=None,
                    help="mixed paths: refinement iterations per solve (library default 20)")
    ap.add_argument("--refine-switch", type=float, default=None,
                    help="mixed paths: average refinement steps per solve above which the next "
                         "renews factor in FP64 (library default 2.5)")
    ap.add_argument("--inverse-solve", type=int, default=None, choices=[-1, 0, 1],
                    help="explicit L11 inverse for supernodes wider than 64 columns: "
                         "0 or 1 on (default), -1 off (block substitution)")
    ap.add_argument("--factor-graph", type=int, default=None, choices=[0, 1],
                    help="1: replay the factorization's launches as a captured HIP graph")
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
    log(f"config {args.config} m={inst.m} n={inst.n} nnz={inst.nnz} (generated in {time.time() - t0:.1f}s)")
    paths = (args.paths or DEFAULT_PATHS[args.config]).split(",")
    for p in paths:
        if p not in PATHS:
            raise SystemExit(f"unknown path {p}")
    extra = {}
    if args.inverse_solve is not None:
        extra["inverse_solve"] = args.inverse_solve
    if args.factor_graph is not None:
        extra["factor_graph"] = args.factor_graph

    mixed = {}
    if args.refine_tol is not None:
        mixed["refine_tol"] = args.refine_tol
    if args.refine_max is not None:
        mixed["refine_max"] = args.refine_max
    if args.refine_switch is not None:
        mixed["refine_switch"] = args.refine_switch
    results = {}
    for p in paths:
        results[p] = run_path(inst, p, args.steps, args.warmup, args.driver, path_opts(p, extra, mixed), L,
                              GpuBackend, solve_lp)
    vp = paths[-1]
    vres = results[vp]
    mode_s = PATHS[vp]["mode"]
    direct = mode_s in ("fp64", "fp32")

    # kernel-class timings of the value path: one more, untimed solve with HIP
    # events around every launch on the library's stream (the events cost
    # time, so the timed steps run without them)
    rows, prof_s = [], None
    if not args.no_profile:
        gp = make_backend(vp, path_opts(vp, extra, mixed), L, GpuBackend, profile=1)
        ts = time.time()
        rp, _ = solve_lp(inst, gp.table, device=args.driver == "gpu")
        L.lib().help_device_synchronize()
        prof_s = time.time() - ts
        rows = kernel_table(L, gp.table)
        rows32 = kernel_table(L, gp.table, fp32_only=True) if mode_s == "fp32" else []
        gp.close()
        log(f"profiled solve ({vp}): status {rp.status} iters {rp.iterations} {prof_s:.3f}s")
    # HBM bytes per dispatch from the rocprofv3 counter runs of the same
    # workload (profiles/traffic.json; null when that workload was not counted)
    wkey = WORKLOAD_KEYS[args.config] + f"_{mode_s}" + ("_nd" if PATHS[vp].get("ordering") else "")
    # the mixed path factors in FP32 until it switches to FP64: its roofline
    # object takes the dominant class's FP32 launches against the FP32 peak,
    # and reports the FP64 launches of the same class beside them
    roof = None
    if rows and mode_s == "fp32" and rows32:
        dom = max(rows, key=lambda r: r["ms"])
        d32 = next((r for r in rows32 if r["kernel"] == dom["kernel"]), None)
        if d32:
            roof = roofline([d32], "fp32", os.path.join(ROOT, "profiles", "traffic.json"), wkey)
            rest = dict(kernel=dom["kernel"], ms=dom["ms"] - d32["ms"], launches=dom["launches"] - d32["launches"],
                        work=dom["work"] - d32["work"], bound=dom["bound"])
            if rest["launches"] > 0:
                r64 = roofline([rest], "fp64", os.path.join(ROOT, "profiles", "traffic.json"), wkey + "_fp64part")
                roof["fp64_launches"] = {k: r64[k] for k in ("achieved", "peak", "frac", "avg_launch_ms", "launches")}
            roof["precision"] = "FP32 launches of the class (the FP64 ones after the switch are under fp64_launches)"
    if roof is None and rows:
        roof = roofline(rows, mode_s, os.path.join(ROOT, "profiles", "traffic.json"), wkey)
    kernels = {r["kernel"]: dict(ms_total=round(r["ms"], 3), launches=r["launches"],
                                 rate=round(r["work"] / (r["ms"] * 1e-3) / (1e9 if r["bound"] == "hbm" else 1e12), 3),
                                 unit="GB/s" if r["bound"] == "hbm" else "TFLOP/s") for r in rows}

    dy = None
    if not args.no_dy_error and direct:
        it = max(vres["iterations"], 1)
        cap = sorted({1, max(1, it // 2), it})
        dy_list, st, its = dy_errors(inst, vp, cap, path_opts(vp, extra, mixed))
        dy = dict(path=vp, driver="host (HELP order)", status=st, iterations=its, solves=dy_list)

    cpu = None
    if not args.no_cpu_baseline and args.config != 3:
        # (the Zipf LP's heavy rows make A D^-1 A^T dense: no direct CPU factor)
        log("timing the CPU baseline sample (MKL PARDISO, all host cores)...")
        cpu = cpu_baseline(inst, vres["iterations"])

    value = vres["seconds"]
    speed = {}
    for p in paths[:-1]:
        speed[f"speedup_vs_{p}"] = round(results[p]["seconds"] / value, 3)
    for p in paths:
        results[p]["seconds"] = round(results[p]["seconds"], 4)
    metric = "time to 1e-8 rel gap (HSD_Main span): " + {
        "fp32": "mixed precision (FP32 factor + FP64 refinement) vs plain FP64",
        "fp64": "plain FP64",
        "pcg": "PCG with the FP32 partial-Cholesky apply vs the FP64 apply" if args.config == 3
               else "PCG from zero (the learned warm start's baseline)",
        "pcg_learned": "PCG from the learned warm start vs from zero"}[mode_s] + ", on 1 MI355X"
    dtype = {"fp64": "f64", "fp32": "f32 factor + f64 refinement", "pcg": "f32 preconditioner + f64 CG",
             "pcg_learned": "bf16 network + f32 preconditioner + f64 CG"}[mode_s]
    line = {
        "metric": metric,
        "value": round(value, 4), "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * value, 2), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (deterministic generator, seed 0)",
        "config": {"workload": NAMES[args.config], "value_path": vp, "driver": args.driver,
                   "solver": "HELP HSD driver", **speed,
                   "paths": results,
                   "inverse_solve": "off" if args.inverse_solve == -1 else "on",
                   "refine_tol": args.refine_tol if args.refine_tol is not None else 1e-12,
                   "profiled_solve_s": None if prof_s is None else round(prof_s, 3)},
        "dy_error": dy,
        "roofline": roof, "cpu_baseline": cpu, "kernels": kernels,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
