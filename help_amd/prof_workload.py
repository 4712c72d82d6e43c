"""Small fixed workloads for profiler runs (kernel traces, hardware counters).

    python -m help_amd.prof_workload solve --config 1 --renews 1 --solves 3
        construct the backend on a benchmark instance, then renew and solve
        with a fixed log-uniform d (the factor and solve kernels);
    python -m help_amd.prof_workload solve --config 2 --mode fp32 --ordering 1
        the staircase LP's FP32 factor under nested dissection;
    python -m help_amd.prof_workload solve --config 3 --mode pcg --pcg-max-iter 60
        the same in the matrix-free PCG mode, each solve capped at a fixed
        number of CG iterations;
    python -m help_amd.prof_workload trace --config 3 --mode pcg --pcg-max-iter 3000
        HELP's HSD driver over the backend, one line per solve (refinement
        steps or CG iterations, residual, time);
    python -m help_amd.prof_workload copy --mib 256
        one device-to-device copy of known size: the reference that scales
        the FETCH_SIZE / WRITE_SIZE counters to bytes.
"""
import argparse
import sys
import time

import numpy as np


def _solve(args):
    from . import _lib as L
    from . import instances
    from .backend import LinearEquation
    inst = instances.by_config(args.config)
    mode = {"fp64": L.MODE_DIRECT_FP64, "fp32": L.MODE_DIRECT_FP32, "pcg": L.MODE_PCG}[args.mode]
    extra = ({"pcg_max_iter": args.pcg_max_iter} if args.mode == "pcg"
             else {"ordering": args.ordering, "ea_depth": args.ea_depth})
    be = LinearEquation(mode=mode, **extra)
    be.construct(inst)
    rng = np.random.default_rng(args.seed)
    d = np.exp(rng.uniform(-args.decades, args.decades, inst.n) * np.log(10.0))
    b1 = rng.standard_normal(inst.n)
    b2 = rng.standard_normal(inst.m)
    for k in range(args.renews):
        t = time.time()
        be.renew(d)
        L.lib().help_device_synchronize()
        print(f"renew {k}: {1e3 * (time.time() - t):.2f} ms", flush=True)
    for k in range(args.solves):
        t = time.time()
        be.solve(d, b1, b2)
        extra = f", {be.solve_stats.pcg_iters} CG iterations" if args.mode == "pcg" else ""
        print(f"solve {k}: {1e3 * (time.time() - t):.2f} ms{extra}", flush=True)
    be.destruct()


def _trace(args):
    """HELP's HSD driver (host) over the backend, one line per solve: the
    renew it follows, refinement steps or CG iterations, relative residual
    and solve time; then the run's result."""
    from . import _lib as L
    from . import instances
    from .hsd import RecordingBackend, solve_lp
    inst = instances.by_config(args.config)
    mode = {"fp64": L.MODE_DIRECT_FP64, "fp32": L.MODE_DIRECT_FP32, "pcg": L.MODE_PCG}[args.mode]
    opts = dict(pcg_k=args.pcg_k, pcg_max_iter=args.pcg_max_iter, pcg_fp32_precond=args.pcg_fp32)
    if args.mode != "pcg":
        opts = dict(ordering=args.ordering, refine_switch=args.refine_switch)
    print(f"{args.mode} config {args.config} m={inst.m} n={inst.n} nnz={inst.nnz} {opts}", flush=True)
    holder = {}

    def on_solve(renews, k, d, b1, b2, x1, x2):
        s = holder["rb"].le.solve_stats
        print(f"renew {renews} solve {k}: refine {s.refine_steps} cg {s.pcg_iters} rel {s.rel_residual:.2e} "
              f"{s.ms_solve:.1f} ms", flush=True)

    rb = RecordingBackend(inst, on_solve, mode=mode, **opts)
    holder["rb"] = rb
    t = time.time()
    res, _ = solve_lp(inst, rb.table, max_iter=args.max_iter)
    print(f"result: status {L.HSD_STATUS[res.status]} stall {res.stall_reason} backend_error {res.backend_error} "
          f"iterations {res.iterations} gap {res.rel_gap:.3e} {time.time() - t:.1f}s", flush=True)


def _copy(args):
    import torch
    n = args.mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {args.mib} MiB", flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m help_amd.prof_workload")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("solve")
    s.add_argument("--config", type=int, default=1)
    s.add_argument("--mode", default="fp64", choices=["fp64", "fp32", "pcg"])
    s.add_argument("--pcg-max-iter", type=int, default=60,
                   help="PCG mode: CG iterations per solve (a fixed amount of work)")
    s.add_argument("--ordering", type=int, default=0, choices=[0, 1],
                   help="direct modes: 0 AMD (default), 1 nested dissection")
    s.add_argument("--renews", type=int, default=1)
    s.add_argument("--solves", type=int, default=3)
    s.add_argument("--decades", type=float, default=2.0)
    s.add_argument("--seed", type=int, default=0)
    s.add_argument("--ea-depth", type=int, default=0, choices=[0, 4, 8],
                   help="direct modes: extend-add rows per lane in flight (0: by front size)")
    t = sub.add_parser("trace")
    t.add_argument("--config", type=int, default=2)
    t.add_argument("--mode", default="fp32", choices=["fp64", "fp32", "pcg"])
    t.add_argument("--ordering", type=int, default=1, choices=[0, 1])
    t.add_argument("--refine-switch", type=float, default=2.5)
    t.add_argument("--pcg-k", type=int, default=500)
    t.add_argument("--pcg-max-iter", type=int, default=5000)
    t.add_argument("--pcg-fp32", type=int, default=1, choices=[0, 1])
    t.add_argument("--max-iter", type=int, default=0)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=256)
    args = ap.parse_args(argv)
    if args.cmd == "solve":
        _solve(args)
    elif args.cmd == "trace":
        _trace(args)
    else:
        _copy(args)
    return 0


if __name__ == "__main__":
    sys.exit(main())
