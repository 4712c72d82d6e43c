"""Writes hsd_objectives.json: optimal objectives of small seeded instances
from the build's own generators, computed by scipy's HiGHS (an independent
LP solver), for the end-to-end HSD tests.  Run from the repository root:
    python tests/golden/make_hsd_fixtures.py
"""
import json
import os
import sys

import numpy as np
import scipy.sparse as sp
from scipy.optimize import linprog

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from help_amd import instances  # noqa: E402

CASES = {
    "random_300x600_s1000": ("random_lp", dict(m=300, nb=600, k=4, seed=1000)),
    "random_500x1000_s1001": ("random_lp", dict(m=500, nb=1000, k=4, seed=1001)),
    "mcf_8x8_k2_s1002": ("mcf", dict(side=8, commodities=2, pairs=4, seed=1002)),
    "mcf_12x12_k3_s1003": ("mcf", dict(side=12, commodities=3, pairs=6, seed=1003)),
    "staircase_5x40x120_s1004": ("staircase", dict(periods=5, block_rows=40, block_cols=120, block_k=5,
                                                   seed=1004)),
    # bounded and free columns (HELP's n_UB and n_FR blocks)
    "random_300x600_ub150_s1010": ("random_lp", dict(m=300, nb=600, k=4, seed=1010, ub_cols=150)),
    "random_300x600_fr20_s1010": ("random_lp", dict(m=300, nb=600, k=4, seed=1010, free_cols=20)),
    "random_300x600_ub100_fr2_s1013": ("random_lp", dict(m=300, nb=600, k=4, seed=1013, ub_cols=100,
                                                         free_cols=2)),
}


def bounds_of(inst):
    """linprog bounds of HELP's column layout [UB | LB-only | free]."""
    out = []
    for j in range(inst.n):
        if j < inst.n_ub:
            out.append((0, float(inst.ub[j])))
        elif j >= inst.n - inst.n_fr:
            out.append((None, None))
        else:
            out.append((0, None))
    return out


def main():
    out = {}
    for name, (gen, kw) in CASES.items():
        inst = getattr(instances, gen)(**kw)
        A = sp.csc_matrix((inst.val, inst.rowind, inst.colptr), shape=(inst.m, inst.n))
        r = linprog(inst.c, A_eq=A, b_eq=inst.b, bounds=bounds_of(inst), method="highs")
        assert r.status == 0, (name, r.message)
        out[name] = dict(generator=gen, params=kw, m=inst.m, n=inst.n, nnz=inst.nnz,
                         objective=float(r.fun))
        print(name, inst.m, inst.n, r.fun)
    with open(os.path.join(HERE, "hsd_objectives.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
