"""The C-ABI library on the host: it loads, exports every symbol its header
declares, reports a missing GPU as an error (no CPU fallback), and its
deterministic instance generators and MPS writer behave.  CPU only."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from help_amd import LinearEquation, instances
from help_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "help_mi355x.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set(re.findall(r"\b(help_\w+)\s*\(", text))
    # drop typedef'd function-pointer names
    names -= set(re.findall(r"\(\s*\*\s*(help_\w+)\s*\)", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(L.LIB_PATH)
    names = _declared_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_options_struct_matches_the_header():
    """help_opts_default through the Python mirror of help_opts: every field
    lands where the header puts it (a field added on one side only shifts
    the later ones)."""
    o = L.Opts()
    L.lib().help_opts_default(C.byref(o))
    assert o.mode == L.MODE_DIRECT_FP64 and o.beta == 1e-5
    assert o.refine_max == 20 and o.refine_tol == 1e-12
    assert o.pcg_k == 500 and o.pcg_max_iter == 5000 and o.pcg_tol == 1e-10 and o.pcg_fp32_precond == 1
    assert o.ordering == L.ORDER_AMD and o.factor_graph == 0 and o.inverse_solve == 0
    assert o.refine_switch == 2.5 and o.ea_depth == 0
    fields = re.search(r"typedef struct help_opts \{(.*?)\} help_opts;", open(HEADER).read(), re.S).group(1)
    fields = re.sub(r"/\*.*?\*/", "", fields, flags=re.S)
    declared = re.findall(r"\b(?:int32_t|double)\s+(\w+)\s*;", fields)
    assert declared == [f[0] for f in L.Opts._fields_]


def test_no_gpu_is_an_error_not_a_fallback():
    if L.lib().help_device_count() > 0:
        pytest.skip("a GPU is visible")
    inst = instances.random_lp(50, 100, 3, seed=1)
    be = LinearEquation()
    with pytest.raises(Exception) as e:
        be.construct(inst)
    assert "NOGPU" in str(e.value) or str(L.HELP_ERR_NOGPU) in str(e.value) or "device" in str(e.value)


def test_generators_are_deterministic():
    a = instances.mcf(side=9, commodities=2, pairs=4, seed=5)
    b = instances.mcf(side=9, commodities=2, pairs=4, seed=5)
    c = instances.mcf(side=9, commodities=2, pairs=4, seed=6)
    for f in ("colptr", "rowind", "val", "b", "c"):
        assert np.array_equal(getattr(a, f), getattr(b, f))
    assert not np.array_equal(a.c, c.c)


@pytest.mark.parametrize("cfg,m,n,nnz", [(0, 2000, 6000, 18000), (1, 269392, 804600, 2234968),
                                         (2, 200000, 800000, 3797000)])
def test_workload_sizes(cfg, m, n, nnz):
    inst = instances.by_config(cfg)
    assert (inst.m, inst.n, inst.nnz) == (m, n, nnz)
    # rows ascending within each column (the order HELP's presolve leaves)
    cp, ri = inst.colptr, inst.rowind
    starts = cp[:-1]
    d = np.diff(ri)
    inside = np.ones(len(d), bool)
    inside[cp[1:-1] - 1] = False
    assert np.all(d[inside[:len(d)]] > 0)


def test_feasible_construction():
    """b = A x0 with x0 > 0 and c = A^T y0 + s0 with s0 > 0: primal and dual
    strictly feasible points exist (the values are rounded to 6 decimals,
    so the identities hold to that rounding)."""
    inst = instances.random_lp(300, 600, 4, seed=2)
    A = np.zeros((inst.m, inst.n))
    for j in range(inst.n):
        A[inst.rowind[inst.colptr[j]:inst.colptr[j + 1]], j] = inst.val[inst.colptr[j]:inst.colptr[j + 1]]
    # the identity block A=[B | I]: the last m columns are unit vectors
    assert np.array_equal(A[:, -inst.m:], np.eye(inst.m))
    assert np.all(np.abs(inst.val * 1e6 - np.round(inst.val * 1e6)) < 1e-6)


def test_mps_writer_round_trip(tmp_path):
    """Fixed-field MPS (MPSRead.cpp semantics): names R%07d / C%07d, values
    printed with 6 decimals, which the generator's rounding makes exact."""
    inst = instances.mcf(side=6, commodities=2, pairs=3, seed=7)
    path = str(tmp_path / "t.mps")
    instances.write_mps(inst, path)
    rows, cols = {}, {}
    vals, rhs, cost = {}, {}, {}
    section = None
    for line in open(path):
        if not line.startswith(" "):
            section = line.split()[0]
            continue
        f = line.split()
        if section == "COLUMNS":
            j = int(f[0][1:])
            if f[1] == "COST":
                cost[j] = float(f[2])
            else:
                vals[(int(f[1][1:]), j)] = float(f[2])
        elif section == "RHS":
            rhs[int(f[1][1:])] = float(f[2])
    for j in range(inst.n):
        for p in range(inst.colptr[j], inst.colptr[j + 1]):
            assert vals[(inst.rowind[p], j)] == inst.val[p]
        assert cost.get(j, 0.0) == inst.c[j]
    for i in range(inst.m):
        assert rhs.get(i, 0.0) == inst.b[i]
    assert len(vals) == inst.nnz


def test_csc_binary_roundtrip(tmp_path):
    """The binary CSC instance file (for LPs beyond MPS's limits, e.g. the
    Zipf LP with m = 1e6, n = 4e6) reads back bit for bit, bounds and free
    columns included; corrupted, truncated or foreign files are rejected."""
    import numpy as np
    from help_amd import instances
    from help_amd.backend import HelpError
    cases = [instances.random_lp(300, 600, 4, seed=61, ub_cols=120, free_cols=7),
             instances.zipf(2000, 6000, 6, 1.1, seed=62)]
    for k, inst in enumerate(cases):
        path = str(tmp_path / f"inst{k}.bin")
        instances.write_csc(inst, path)
        back = instances.read_csc(path)
        assert (back.m, back.n, back.n_ub, back.n_fr) == (inst.m, inst.n, inst.n_ub, inst.n_fr)
        for a in ("colptr", "rowind", "val", "b", "c"):
            assert np.array_equal(getattr(back, a), getattr(inst, a)), a
        if inst.n_ub:
            assert np.array_equal(back.ub, inst.ub)
    raw = open(path, "rb").read()
    bad = bytearray(raw)
    bad[len(bad) // 2] ^= 0x40                      # one flipped bit in the payload
    for name, data in (("flip", bytes(bad)), ("short", raw[:-9]), ("magic", b"NOTCSC01" + raw[8:])):
        p = str(tmp_path / f"{name}.bin")
        open(p, "wb").write(data)
        with pytest.raises(HelpError):
            instances.read_csc(p)
    with pytest.raises(HelpError):
        instances.read_csc(str(tmp_path / "missing.bin"))
