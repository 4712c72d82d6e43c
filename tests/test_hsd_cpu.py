"""End-to-end HSD on the host with the CPU oracle backend.

* The optimal objective is pinned by scipy's HiGHS (fixtures in
  tests/golden/hsd_objectives.json, written by make_hsd_fixtures.py).
* The C++ driver restatement in exact mode is bit-identical to HELP's own
  driver (HSDSolver.cpp, built from the reference sources by oracle/Makefile
  into oracle/_ref/help_ref_cpu) run on the same instance written as MPS,
  both with the oracle backend in natural order.
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_bind as O
from help_amd import _lib as L
from help_amd import instances
from help_amd.hsd import solve_lp

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "hsd_objectives.json")))
REF_BIN = os.path.join(O.REF_DIR, "help_ref_cpu")


def _instance(spec):
    return getattr(instances, spec["generator"])(**spec["params"])


@pytest.mark.parametrize("name", sorted(FIX))
def test_hsd_objective_matches_highs(name):
    spec = FIX[name]
    inst = _instance(spec)
    assert (inst.m, inst.n, inst.nnz) == (spec["m"], spec["n"], spec["nnz"])
    ob = O.OracleBackend(1e-5)
    try:
        res, rows = solve_lp(inst, ob.table, exact=True)
    finally:
        ob.close()
    assert L.HSD_STATUS[res.status] == "OK"
    assert res.rel_gap <= 1e-8
    obj = spec["objective"]
    assert abs(res.primal_obj - obj) <= 1e-6 * (1 + abs(obj))
    # the iteration log is monotone in the gap's final order of magnitude
    assert rows[-1][3] <= 1e-8


def _ref_result(mps):
    out = subprocess.run([REF_BIN, mps], capture_output=True, text=True, timeout=300).stdout
    line = [l for l in out.splitlines() if l.startswith("RESULT")][-1]
    kv = dict(re.findall(r"(\w+)=(\S+)", line))
    return kv


def _compare_with_reference_driver(inst, tmp_path):
    mps = str(tmp_path / "inst.mps")
    instances.write_mps(inst, mps)
    ref = _ref_result(mps)
    ob = O.OracleBackend(1e-5)
    try:
        res, rows = solve_lp(inst, ob.table, exact=True)
    finally:
        ob.close()
    assert int(ref["status"]) == res.status
    assert float(ref["pobj"]) == res.primal_obj
    assert float(ref["dobj"]) == res.dual_obj
    assert int(ref["m"]) == inst.m and int(ref["n"]) == inst.n
    return res


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference driver not built (no reference sources)")
@pytest.mark.parametrize("name", ["random_300x600_s1000", "mcf_8x8_k2_s1002", "random_300x600_ub150_s1010",
                                  "random_300x600_fr20_s1010", "random_300x600_ub100_fr2_s1013"])
def test_driver_bit_identical_to_reference_driver(name, tmp_path):
    """The bounded/free cases go through MPSRead's BOUNDS section, Crushing.cpp
    and Presolve_FinalizeModel on the reference side and exercise the n_UB
    block (HSDSolver.cpp:70-82,99-100,160-165,203-207) and the free-column
    terms (HSDSolver.cpp:84-90,101-102) of both drivers."""
    res = _compare_with_reference_driver(_instance(FIX[name]), tmp_path)
    assert L.HSD_STATUS[res.status] == "OK"


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference driver not built (no reference sources)")
def test_driver_matches_reference_driver_stall(tmp_path):
    """Bounded and free columns together on this seed make HELP's own driver
    stop on its inaccurate-Newton-direction test (HSDSolver.cpp:814-818);
    the restatement stops at the same iterate with the same status."""
    inst = instances.random_lp(300, 600, 4, seed=1011, ub_cols=150, free_cols=5)
    res = _compare_with_reference_driver(inst, tmp_path)
    assert L.HSD_STATUS[res.status] == "STALLED" and res.stall_reason == 2
