"""Writes mcf150_k8_oracle.json: the CPU FP64 oracle's HSD runs on the
150x150-grid, K=8 multicommodity-flow LP (m=269,392, n=804,600,
nnz=2,234,968; generator seed 0), with the backend's own fill-reducing
ordering, so a GPU test can compare the full-size run's status, iteration
count and objective without re-running the ~35-minute single-core oracle.

Two runs of the same oracle backend, differing only in how the HSD driver
sums its reductions: "exact" (sequentially, in HELP's order, so the iterates
are those of HELP's own HSDSolver.cpp with this backend) and "chunked"
(64 fixed chunks).  Each records the final status, iterations, objectives,
gap and infeasibilities and the per-iteration log (iteration, primal and
dual objective, gap, primal and dual infeasibility).  Run from the
repository root (a run is skipped when the file already holds it):
    python tests/golden/make_mcf150_fixture.py [exact|chunked ...]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from help_amd import LinearEquation, instances  # noqa: E402
from help_amd import _lib as L  # noqa: E402
from help_amd.hsd import solve_lp  # noqa: E402
import oracle_bind as O  # noqa: E402


def main(modes):
    path = os.path.join(HERE, "mcf150_k8_oracle.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    inst = instances.mcf(150, 8, 20, 0)
    out["instance"] = dict(generator="mcf", params=dict(side=150, commodities=8, pairs=20, seed=0),
                           m=inst.m, n=inst.n, nnz=inst.nnz)
    out["backend"] = "oracle (CHOLMOD semantics, beta 1e-5) with the MI355X backend's AMD ordering"
    runs = out.setdefault("runs", {})
    be = LinearEquation(symbolic_only=True)
    be.construct(inst)
    perm, _, _ = be.symbolic_arrays()
    be.destruct()
    for mode in modes:
        if mode in runs:
            continue
        ob = O.OracleBackend(1e-5, perm)
        t0 = time.time()
        res, rows = solve_lp(inst, ob.table, exact=mode == "exact", log=lambda *r: print(*r, flush=True))
        ob.close()
        runs[mode] = dict(
            driver=f"help_hsd_solve ({'sequential reductions in HELP order' if mode == 'exact' else '64-chunk reductions'})",
            status=L.HSD_STATUS[res.status], stall_reason=res.stall_reason, iterations=res.iterations,
            primal_obj=res.primal_obj, dual_obj=res.dual_obj, rel_gap=res.rel_gap,
            primal_infeas=res.primal_infeas, dual_infeas=res.dual_infeas,
            seconds=round(time.time() - t0, 1), log=[list(r) for r in rows])
        with open(path, "w") as f:
            json.dump(out, f, indent=0)
        print(mode, runs[mode]["status"], runs[mode]["iterations"], runs[mode]["primal_obj"])


if __name__ == "__main__":
    main(sys.argv[1:] or ["exact", "chunked"])
