"""Parity of the MI355X backend with the CPU FP64 oracle (GPU tests).

Bars:
* ADA^T assembly and the SpMV gathers: bit-exact (same order, no FMA).
* Factor: every column of L within 1e-12 of the oracle's, relative to the
  column's largest entry (different but backward-stable summation orders).
* Solve: x2 (= dy) within 1e-10 relative of the oracle's on well-conditioned
  d; x1 likewise.
* HSD end to end: same status, iterations within +-1, objective within 1e-8.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_bind as O
from conftest import log_uniform
from help_amd import LinearEquation, instances
from help_amd import _lib as L

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_factor(inst, perm, dvec, beta=1e-5):
    arp, aci = O.adat_pattern(inst)
    aval = O.adat_values(inst, dvec, arp, aci)
    parent, cc, Lp, Li = O.symbolic(inst.m, arp, aci, perm)
    st, Lx = O.cholesky(inst.m, arp, aci, aval, beta, perm, Lp, Li)
    assert st == 0
    return Lp, Li, Lx


@pytest.fixture(scope="module")
def be0(cfg0):
    _flush("construct cfg0 m=2000 n=6000")
    be = LinearEquation()
    be.construct(cfg0)
    yield be
    be.destruct()


def test_adat_bit_exact(be0, cfg0):
    rng = np.random.default_rng(1)
    d = log_uniform(rng, cfg0.n, -8, 8)
    _flush("adat cfg0")
    g = be0.adat_values(d)
    arp, aci = O.adat_pattern(cfg0)
    o = O.adat_values(cfg0, d, arp, aci)
    assert np.array_equal(g.view(np.uint64), o.view(np.uint64))


def test_adat_bit_exact_mcf(small_mcf):
    be = LinearEquation()
    be.construct(small_mcf)
    rng = np.random.default_rng(2)
    d = log_uniform(rng, small_mcf.n, -4, 4)
    _flush("adat mcf")
    g = be.adat_values(d)
    arp, aci = O.adat_pattern(small_mcf)
    o = O.adat_values(small_mcf, d, arp, aci)
    assert np.array_equal(g.view(np.uint64), o.view(np.uint64))
    be.destruct()


def test_spmv_bit_exact(be0, cfg0):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(cfg0.n)
    y = rng.standard_normal(cfg0.m)
    _flush("spmv cfg0")
    g1 = be0.spmv(x, trans=False)
    g2 = be0.spmv(y, trans=True)
    rp, ci, pos = O.csr(cfg0)
    csrval = np.ascontiguousarray(cfg0.val[pos])
    o1 = np.zeros(cfg0.m)
    O.lib().or_csrmv_n(cfg0.m, 1.0, O.d(csrval), O.i32(ci), O.i64(rp), O.d(x), 0.0, O.d(o1))
    o2 = np.zeros(cfg0.n)
    O.lib().or_at_times(cfg0.m, cfg0.n, O.i64(cfg0.colptr), O.i32(cfg0.rowind), O.d(cfg0.val), 1,
                        O.d(y), O.d(o2))
    assert np.array_equal(g1.view(np.uint64), o1.view(np.uint64))
    assert np.array_equal(g2.view(np.uint64), o2.view(np.uint64))


def test_factor_matches_oracle(be0, cfg0):
    rng = np.random.default_rng(4)
    d = log_uniform(rng, cfg0.n, -2, 2)
    _flush("renew cfg0 fp64")
    be0.renew(d)
    assert be0.renew_stats.n_tiny_pivots == 0
    perm, parent, cc = be0.symbolic_arrays()
    Lp, Li, Lx = _oracle_factor(cfg0, perm, d)
    worst = 0.0
    for j in list(range(0, cfg0.m, 7)) + [cfg0.m - 1]:
        rows, vals = be0.factor_column(j)
        ref = dict(zip(Li[Lp[j]:Lp[j + 1]], Lx[Lp[j]:Lp[j + 1]]))
        scale = max(abs(v) for v in ref.values())
        got = dict(zip(rows, vals))
        for r, v in ref.items():
            worst = max(worst, abs(got.get(r, 0.0) - v) / scale)
        for r, v in got.items():   # relaxed-supernode zeros must stay zero
            if r not in ref:
                worst = max(worst, abs(v) / scale)
    assert worst <= 1e-12, worst


@pytest.mark.parametrize("mode", [L.MODE_DIRECT_FP64, L.MODE_DIRECT_FP32])
def test_solve_matches_oracle(cfg0, mode):
    be = LinearEquation(mode=mode)
    be.construct(cfg0)
    rng = np.random.default_rng(5)
    d = log_uniform(rng, cfg0.n, -2, 2)
    b1 = rng.standard_normal(cfg0.n)
    b2 = rng.standard_normal(cfg0.m)
    _flush(f"solve cfg0 mode={mode}")
    be.renew(d)
    x1, x2 = be.solve(d, b1, b2)
    perm, _, _ = be.symbolic_arrays()
    Lp, Li, Lx = _oracle_factor(cfg0, perm, d)
    o1, o2 = O.solve_kkt(cfg0, perm, Lp, Li, Lx, d, b1, b2)
    e2 = np.linalg.norm(x2 - o2) / np.linalg.norm(o2)
    e1 = np.linalg.norm(x1 - o1) / np.linalg.norm(o1)
    _flush(f"  dy err {e2:.3e}  x1 err {e1:.3e}  refine {be.solve_stats.refine_steps}")
    assert e2 <= 1e-10 and e1 <= 1e-10
    be.destruct()


@pytest.mark.parametrize("flag", [L.FAULT_REFINE_NOCONV, L.FAULT_FP32_BREAKDOWN])
def test_fp32_fallback_to_fp64(cfg0, flag):
    """A refinement failure of the FP32 factor (forced from the host) makes
    the solve refactor in FP64 with the same d and keep FP64 afterwards."""
    be = LinearEquation(mode=L.MODE_DIRECT_FP32, fault_flags=flag)
    be.construct(cfg0)
    rng = np.random.default_rng(8)
    d = log_uniform(rng, cfg0.n, -2, 2)
    b1 = rng.standard_normal(cfg0.n)
    b2 = rng.standard_normal(cfg0.m)
    _flush(f"fp32 fallback cfg0 flag={flag}")
    be.renew(d)
    x1, x2 = be.solve(d, b1, b2)
    assert be.solve_stats.fallback == 1
    perm, _, _ = be.symbolic_arrays()
    Lp, Li, Lx = _oracle_factor(cfg0, perm, d)
    o1, o2 = O.solve_kkt(cfg0, perm, Lp, Li, Lx, d, b1, b2)
    assert np.linalg.norm(x2 - o2) <= 1e-10 * np.linalg.norm(o2)
    assert np.linalg.norm(x1 - o1) <= 1e-10 * np.linalg.norm(o1)
    # the next renew factors in FP64: columns equal the FP64 oracle's
    be.renew(d)
    worst = 0.0
    for j in range(0, cfg0.m, 97):
        rows, vals = be.factor_column(j)
        ref = dict(zip(Li[Lp[j]:Lp[j + 1]], Lx[Lp[j]:Lp[j + 1]]))
        scale = max(abs(v) for v in ref.values())
        got = dict(zip(rows, vals))
        worst = max([worst] + [abs(got.get(r, 0.0) - v) / scale for r, v in ref.items()])
    assert worst <= 1e-12, worst
    x1b, x2b = be.solve(d, b1, b2)
    assert be.solve_stats.fallback == 0
    assert np.linalg.norm(x2b - o2) <= 1e-10 * np.linalg.norm(o2)
    be.destruct()


def test_solve_mcf_large_fronts():
    inst = instances.mcf(side=60, commodities=4, pairs=10, seed=7)
    be = LinearEquation()
    be.construct(inst)
    info = be.symbolic_info()
    _flush(f"mcf60 m={inst.m} levels={info.n_levels} max_front={info.max_front}")
    assert info.max_front > 256   # exercises the tiled large-front path
    rng = np.random.default_rng(6)
    b1 = rng.standard_normal(inst.n)
    b2 = rng.standard_normal(inst.m)
    A = inst.to_scipy()
    perm, _, _ = be.symbolic_arrays()
    for lo, hi in [(-2, 2), (-5, 5)]:
        d = log_uniform(rng, inst.n, lo, hi)
        be.renew(d)
        x1, x2 = be.solve(d, b1, b2)
        # backward error of the normal-equation solve (scale-free)
        M = (A @ (A.T.multiply(1.0 / d[:, None]))).tocsr()
        r = A @ (b1 / d) - b2
        res = r - (M @ x2 + 1e-5 * x2)
        bwd = np.linalg.norm(res) / (abs(M).sum(axis=1).max() * np.linalg.norm(x2) + np.linalg.norm(r))
        _flush(f"  d in 1e{lo}..1e{hi}: backward error {bwd:.3e}")
        assert bwd <= 1e-13
        if hi <= 2:   # well conditioned: forward agreement with the oracle
            Lp, Li, Lx = _oracle_factor(inst, perm, d)
            o1, o2 = O.solve_kkt(inst, perm, Lp, Li, Lx, d, b1, b2)
            e2 = np.linalg.norm(x2 - o2) / np.linalg.norm(o2)
            _flush(f"  dy err {e2:.3e}")
            assert e2 <= 1e-10
    be.destruct()


@pytest.mark.parametrize("mode,inverse", [(L.MODE_DIRECT_FP64, 1), (L.MODE_DIRECT_FP32, 1),
                                          (L.MODE_DIRECT_FP64, -1), (L.MODE_DIRECT_FP32, -1)])
def test_solve_with_explicit_inverse(mode, inverse):
    """Supernodes wider than one 64-column block solved through the explicit
    inverse of L11 (inverse=1, the default) or by block substitution with the
    diagonal-block inverses (inverse=-1): solutions within 1e-10 of the
    oracle on well-conditioned d."""
    inst = instances.mcf(side=60, commodities=4, pairs=10, seed=7)
    be = LinearEquation(mode=mode, inverse_solve=inverse)
    be.construct(inst)
    sf, _, _, _ = be.supernodes()
    wide = int((np.diff(sf) > 64).sum())
    _flush(f"explicit inverse mode={mode} inverse={inverse}: {wide} supernodes wider than 64 columns")
    assert wide > 0
    rng = np.random.default_rng(12)
    d = log_uniform(rng, inst.n, -2, 2)
    b1 = rng.standard_normal(inst.n)
    b2 = rng.standard_normal(inst.m)
    be.renew(d)
    x1, x2 = be.solve(d, b1, b2)
    perm, _, _ = be.symbolic_arrays()
    Lp, Li, Lx = _oracle_factor(inst, perm, d)
    o1, o2 = O.solve_kkt(inst, perm, Lp, Li, Lx, d, b1, b2)
    e2 = np.linalg.norm(x2 - o2) / np.linalg.norm(o2)
    e1 = np.linalg.norm(x1 - o1) / np.linalg.norm(o1)
    _flush(f"  dy err {e2:.3e}  x1 err {e1:.3e}  refine {be.solve_stats.refine_steps}")
    assert e2 <= 1e-10 and e1 <= 1e-10
    be.destruct()


@pytest.mark.parametrize("mode,inverse,dec", [(L.MODE_DIRECT_FP64, 1, 2), (L.MODE_DIRECT_FP64, -1, 2),
                                              (L.MODE_DIRECT_FP32, 1, 2), (L.MODE_DIRECT_FP32, 1, 6)])
def test_solve2_matches_two_solves(mode, inverse, dec):
    """help_gpu_solve2 (two right-hand sides through each triangular-solve
    launch: in the FP64 mode one factor solve each, in the FP32 mode the two
    refinements in lockstep, with d over 6 decades for many steps) returns
    exactly the two single solves' results and refinement counts, on an instance with wide supernodes (explicit inverse
    and block substitution) and many tiny ones; the first solution also
    within 1e-10 of the oracle."""
    inst = instances.mcf(side=60, commodities=4, pairs=10, seed=7)
    be = LinearEquation(mode=mode, inverse_solve=inverse)
    be.construct(inst)
    rng = np.random.default_rng(21)
    d = log_uniform(rng, inst.n, -dec, dec)
    b1a, b1b = rng.standard_normal(inst.n), rng.standard_normal(inst.n)
    b2a, b2b = rng.standard_normal(inst.m), rng.standard_normal(inst.m)
    be.renew(d)
    for rep in range(2):   # the second pass reuses the update buffers left by the first
        (x1a, x2a), (x1b, x2b) = be.solve2(d, b1a, b2a, b1b, b2b)
        steps2 = be.solve_stats.refine_steps
        y1a, y2a = be.solve(d, b1a, b2a)
        steps1 = be.solve_stats.refine_steps
        y1b, y2b = be.solve(d, b1b, b2b)
        _flush(f"solve2 mode={mode} inverse={inverse} decades={dec} pass {rep}: refinement steps "
               f"{steps2} / {steps1} (then {be.solve_stats.refine_steps}), max |diff| "
               f"{max(np.abs(x2a - y2a).max(), np.abs(x2b - y2b).max()):.3e}")
        assert steps2 == steps1
        for u, v in ((x1a, y1a), (x2a, y2a), (x1b, y1b), (x2b, y2b)):
            assert np.array_equal(u, v)
        b1a, b1b = b1b, b1a   # swap the roles for the second pass
        b2a, b2b = b2b, b2a
    if dec == 2:   # the forward error bound holds for the moderately conditioned systems
        perm, _, _ = be.symbolic_arrays()
        Lp, Li, Lx = _oracle_factor(inst, perm, d)
        o1, o2 = O.solve_kkt(inst, perm, Lp, Li, Lx, d, b1a, b2a)
        e2 = np.linalg.norm(x2b - o2) / np.linalg.norm(o2)
        _flush(f"  dy err {e2:.3e}")
        assert e2 <= 1e-10
    be.destruct()


@pytest.mark.parametrize("mode,device", [(L.MODE_DIRECT_FP64, False), (L.MODE_DIRECT_FP32, False),
                                         (L.MODE_DIRECT_FP64, True), (L.MODE_DIRECT_FP32, True)])
def test_hsd_end_to_end(cfg0, mode, device):
    from help_amd.hsd import GpuBackend, solve_lp
    _flush(f"hsd cfg0 mode={mode} device_driver={device}")
    gb = GpuBackend(mode=mode)
    res, rows = solve_lp(cfg0, gb.table, device=device)
    gb.close()
    be = LinearEquation(symbolic_only=True)
    be.construct(cfg0)
    perm, _, _ = be.symbolic_arrays()
    ob = O.OracleBackend(1e-5, perm)
    ref, _ = solve_lp(cfg0, ob.table)
    ob.close()
    _flush(f"  gpu {res.status} {res.iterations} {res.primal_obj!r} | oracle {ref.status} "
           f"{ref.iterations} {ref.primal_obj!r}")
    assert res.backend_error == 0
    assert res.status == ref.status == 0
    assert abs(res.iterations - ref.iterations) <= 1
    assert abs(res.primal_obj - ref.primal_obj) / (1 + abs(ref.primal_obj)) <= 1e-8
    assert res.rel_gap <= 1e-8


@pytest.mark.parametrize("side,k", [(30, 8), (40, 6)])
def test_hsd_end_to_end_mcf(side, k):
    """Multicommodity-flow instances of the benchmark's family: the plain FP64
    path (one factor solve per system, as UseCholmod.cpp) reaches the
    oracle's status, iteration count within +-1 and objective."""
    from help_amd.hsd import GpuBackend, solve_lp
    inst = instances.mcf(side=side, commodities=k, pairs=20, seed=50 + side)
    _flush(f"hsd mcf side={side} K={k} m={inst.m} n={inst.n}")
    gb = GpuBackend(mode=L.MODE_DIRECT_FP64)
    res, _ = solve_lp(inst, gb.table, device=True)
    gb.close()
    be = LinearEquation(symbolic_only=True)
    be.construct(inst)
    perm, _, _ = be.symbolic_arrays()
    ob = O.OracleBackend(1e-5, perm)
    ref, _ = solve_lp(inst, ob.table)
    ob.close()
    _flush(f"  gpu {res.status} {res.iterations} {res.primal_obj!r} | oracle {ref.status} "
           f"{ref.iterations} {ref.primal_obj!r}")
    assert res.status == ref.status == 0
    assert abs(res.iterations - ref.iterations) <= 1
    assert abs(res.primal_obj - ref.primal_obj) / (1 + abs(ref.primal_obj)) <= 1e-8


def test_reference_driver_with_mi355x_backend(cfg0, tmp_path):
    exe = os.path.join(O.REF_DIR, "help_ref_mi355x")
    cpu = os.path.join(O.REF_DIR, "help_ref_cpu")
    if not (os.path.exists(exe) and os.path.exists(cpu)):
        pytest.skip("reference driver build not present")
    mps = str(tmp_path / "cfg0.mps")
    instances.write_mps(cfg0, mps)
    be = LinearEquation(symbolic_only=True)
    be.construct(cfg0)
    perm, _, _ = be.symbolic_arrays()
    pf = str(tmp_path / "perm.bin")
    perm.tofile(pf)
    _flush("reference HELP driver + UseMI355X.cpp on cfg0")
    out_g = subprocess.run([exe, mps], capture_output=True, text=True, timeout=120).stdout
    out_c = subprocess.run([cpu, mps, pf], capture_output=True, text=True, timeout=120).stdout

    def parse(out):
        rows = [l for l in out.splitlines() if l.startswith("|") and "Iter" not in l]
        res = [l for l in out.splitlines() if l.startswith("RESULT")][0]
        kv = dict(t.split("=") for t in res.split()[1:])
        return len(rows), kv

    ng, kg = parse(out_g)
    nc, kc = parse(out_c)
    _flush(f"  mi355x iters {ng} {kg['pobj']} | cpu iters {nc} {kc['pobj']}")
    assert kg["status"] == kc["status"] == "0"
    assert abs(ng - nc) <= 1
    po_g, po_c = float(kg["pobj"]), float(kc["pobj"])
    assert abs(po_g - po_c) / (1 + abs(po_c)) <= 1e-8


@pytest.mark.parametrize("fp64", [1, 0])
def test_mfma_lane_layout(fp64):
    """In-bounds probe of V_MFMA_{F64,F32}_16X16X4 lane layout: A[i][k] in lane
    i+16k, B[k][j] in lane j+16k; D[i][j] in lane j+16(i%4) element i/4 (F64)
    or lane j+16(i/4) element i%4 (F32)."""
    rng = np.random.default_rng(9)
    A = rng.integers(-4, 5, (16, 4)).astype(np.float64)
    B = rng.integers(-4, 5, (4, 16)).astype(np.float64)
    a = np.array([A[l % 16, l // 16] for l in range(64)])
    b = np.array([B[l // 16, l % 16] for l in range(64)])
    out = np.zeros(256)
    st = L.lib().help_selftest_mfma(fp64, L.dptr(a), L.dptr(b), L.dptr(out))
    assert st == 0
    Dm = A @ B
    if fp64:
        exp = np.array([Dm[4 * r + l // 16, l % 16] for l in range(64) for r in range(4)])
    else:
        exp = np.array([Dm[4 * (l // 16) + r, l % 16] for l in range(64) for r in range(4)])
    _flush(f"mfma fp64={fp64} max|diff|={np.abs(out - exp).max()}")
    assert np.array_equal(out, exp)


def _first_bad_supernode(be, inst, d):
    perm, parent, cc = be.symbolic_arrays()
    Lp, Li, Lx = _oracle_factor(inst, perm, d)
    sf, nr, lv, tl = be.supernodes()
    order = np.lexsort((np.arange(len(nr)), lv))
    for s in order:
        worst = 0.0
        for j in range(sf[s], sf[s + 1]):
            rows, vals = be.factor_column(j)
            ref = dict(zip(Li[Lp[j]:Lp[j + 1]], Lx[Lp[j]:Lp[j + 1]]))
            scale = max(abs(v) for v in ref.values())
            got = dict(zip(rows, vals))
            for r, v in ref.items():
                worst = max(worst, abs(got.get(r, 0.0) - v) / scale)
        if worst > 1e-10:
            kids = [c for c in range(len(nr)) if sf[c + 1] - 1 >= 0 and parent[sf[c + 1] - 1] >= sf[s]
                    and parent[sf[c + 1] - 1] < sf[s + 1]]
            return (int(s), int(lv[s]), int(sf[s + 1] - sf[s]), int(nr[s]), int(tl[s]), worst,
                    [(int(c), int(tl[c]), int(nr[c])) for c in kids])
    return None


@pytest.mark.parametrize("path,depth", [(0, 0), (1, 0), (2, 0), (1, 4), (1, 8)])
def test_factor_paths_diagnostic(cfg0, path, depth):
    """Every supernode's columns against the oracle with the automatic front
    schedule (0), every front through the tiled kernels (1), and every small
    front that fits one wave through the 64-lane one-wave kernel (2); with
    every front tiled, the extend-add also with 4 and with 8 child rows per
    lane in flight in every launch (depth 0: by front size)."""
    be = LinearEquation(front_path=path, ea_depth=depth)
    be.construct(cfg0)
    rng = np.random.default_rng(4)
    d = log_uniform(rng, cfg0.n, -2, 2)
    _flush(f"renew cfg0 front_path={path} ea_depth={depth}")
    be.renew(d)
    bad = _first_bad_supernode(be, cfg0, d)
    _flush(f"  tiny pivots {be.renew_stats.n_tiny_pivots}; first bad supernode "
           f"(s, level, ncol, nrow, tiled, err, children) = {bad}")
    be.destruct()
    assert bad is None


@pytest.mark.parametrize("mode", [L.MODE_DIRECT_FP64, L.MODE_DIRECT_FP32])
def test_factor_graph_replays_match_eager(small_mcf, mode):
    """The factorization captured once as a HIP graph and replayed on later
    renews gives the eager launches' factor bit for bit, replay after replay,
    with other eager work and synchronizations between the replays."""
    inst = small_mcf
    eager = LinearEquation(mode=mode)
    graph = LinearEquation(mode=mode, factor_graph=1)
    eager.construct(inst)
    graph.construct(inst)
    rng = np.random.default_rng(21)
    cols = rng.choice(inst.m, 40, replace=False)
    for rep in range(4):
        d = log_uniform(rng, inst.n, -2, 2)
        _flush(f"factor graph mode={mode} renew {rep}")
        eager.renew(d)
        graph.renew(d)
        for j in cols:
            re_, ve = eager.factor_column(int(j))
            rg, vg = graph.factor_column(int(j))
            assert np.array_equal(re_, rg) and np.array_equal(ve, vg)
        b1, b2 = rng.standard_normal(inst.n), rng.standard_normal(inst.m)
        xe = eager.solve(d, b1, b2)
        xg = graph.solve(d, b1, b2)
        assert np.array_equal(xe[0], xg[0]) and np.array_equal(xe[1], xg[1])
    eager.destruct()
    graph.destruct()
