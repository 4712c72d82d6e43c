"""GPU parity of the matrix-free PCG mode (HELP PCG.cpp / UsePCG_CPU.cpp):
the partial-Cholesky preconditioner against the CPU oracle, the solve against
dense and sparse direct solves, and the HSD driver end to end."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle_bind as O
from help_amd import LinearEquation, instances
from help_amd import _lib as L

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _flush(msg):
    print(msg, flush=True)


def _csc(inst):
    return sp.csc_matrix((inst.val, inst.rowind, inst.colptr), shape=(inst.m, inst.n))


def _oracle_pchol(inst, d, K):
    rp, ci, pos = O.csr(inst)
    cv = np.ascontiguousarray(inst.val[pos])
    pc = O.PChol()
    O.lib().or_pchol_build(inst.m, inst.n, O.i64(rp), O.i32(ci), O.d(cv), O.d(d), 0.0, 0.0, K, pc)
    m = inst.m
    out = (np.ctypeslib.as_array(pc.perm, shape=(m,)).copy(),
           np.ctypeslib.as_array(pc.D, shape=(K,)).copy(),
           np.ctypeslib.as_array(pc.diag_s, shape=(m,)).copy(),
           np.ctypeslib.as_array(pc.L, shape=(m * K,)).reshape(m, K).copy())
    O.lib().or_pchol_free(pc)
    return out


@pytest.mark.parametrize("name,maker,K", [
    ("random", lambda: instances.random_lp(300, 600, 4, seed=21), 20),
    ("zipf-heavy-rows", lambda: instances.zipf(3000, 9000, 6, 1.1, seed=22), 40)])
def test_partial_cholesky_matches_oracle(name, maker, K):
    inst = maker()
    rp = O.csr(inst)[0]
    _flush(f"pcg factor {name} m={inst.m} n={inst.n} K={K} max row nnz={int(np.diff(rp).max())}")
    rng = np.random.default_rng(1)
    d = np.exp(rng.uniform(-2, 2, inst.n) * np.log(10.0))
    be = LinearEquation(mode=L.MODE_PCG, pcg_k=K)
    be.construct(inst)
    be.renew(d)
    perm, D, S, Lg = be.pcg_factor()
    be.destruct()
    operm, oD, oS, oL = _oracle_pchol(inst, d, K)
    assert np.array_equal(perm, operm)
    assert np.allclose(D, oD, rtol=1e-11, atol=0)
    assert np.allclose(S, oS, rtol=1e-9, atol=1e-12 * np.abs(oS).max())
    assert np.allclose(Lg, oL, rtol=1e-9, atol=1e-11)


def test_pcg_solve_matches_dense():
    inst = instances.random_lp(400, 800, 4, seed=23)
    rng = np.random.default_rng(2)
    d = np.exp(rng.uniform(-2, 2, inst.n) * np.log(10.0))
    b1 = rng.standard_normal(inst.n)
    b2 = rng.standard_normal(inst.m)
    _flush("pcg solve vs dense, random 400x800")
    be = LinearEquation(mode=L.MODE_PCG, pcg_k=30, pcg_tol=1e-12)
    be.construct(inst)
    be.renew(d)
    x1, x2 = be.solve(d, b1, b2)
    st = be.solve_stats
    be.destruct()
    A = _csc(inst).toarray()
    M = A @ np.diag(1.0 / d) @ A.T          # gamma = delta = 0 (UsePCG_CPU.cpp:154-155)
    x2r = np.linalg.solve(M, A @ (b1 / d) - b2)
    x1r = (b1 - A.T @ x2r) / d
    _flush(f"  iterations {st.pcg_iters} rel residual {st.rel_residual:.2e}")
    assert 0 < st.pcg_iters < 400 and st.rel_residual <= 1e-12
    assert np.linalg.norm(x2 - x2r) <= 1e-8 * np.linalg.norm(x2r)
    assert np.linalg.norm(x1 - x1r) <= 1e-8 * np.linalg.norm(x1r)


def test_pcg_solve_heavy_rows_matches_sparse_direct():
    """Rows with thousands of nonzeros go through the chunked product."""
    inst = instances.zipf(3000, 9000, 6, 1.1, seed=24)
    rng = np.random.default_rng(3)
    d = np.exp(rng.uniform(-1, 1, inst.n) * np.log(10.0))
    b1 = rng.standard_normal(inst.n)
    b2 = rng.standard_normal(inst.m)
    _flush("pcg solve vs sparse direct, zipf 3000x9000")
    be = LinearEquation(mode=L.MODE_PCG, pcg_k=50, pcg_tol=1e-12, pcg_max_iter=3000)
    be.construct(inst)
    be.renew(d)
    x1, x2 = be.solve(d, b1, b2)
    st = be.solve_stats
    be.destruct()
    A = _csc(inst)
    M = (A @ sp.diags(1.0 / d) @ A.T).tocsc()
    x2r = spla.spsolve(M, A @ (b1 / d) - b2)
    _flush(f"  iterations {st.pcg_iters} rel residual {st.rel_residual:.2e}")
    assert st.rel_residual <= 1e-12
    assert np.linalg.norm(x2 - x2r) <= 1e-7 * np.linalg.norm(x2r)


@pytest.mark.parametrize("device", [False, True])
def test_hsd_pcg_mode_reaches_highs_objective(device):
    from help_amd.hsd import GpuBackend, solve_lp
    spec = json.load(open(os.path.join(HERE, "golden", "hsd_objectives.json")))["random_300x600_s1000"]
    inst = getattr(instances, spec["generator"])(**spec["params"])
    _flush(f"hsd pcg mode device_driver={device}")
    gb = GpuBackend(mode=L.MODE_PCG, pcg_k=50)
    res, rows = solve_lp(inst, gb.table, device=device)
    gb.close()
    _flush(f"  status {res.status} iters {res.iterations} pobj {res.primal_obj!r} (HiGHS {spec['objective']!r})")
    assert res.backend_error == 0
    assert L.HSD_STATUS[res.status] == "OK"
    assert abs(res.primal_obj - spec["objective"]) <= 1e-6 * (1 + abs(spec["objective"]))


@pytest.mark.parametrize("fp32", [1, 0])
@pytest.mark.parametrize("name,maker,K", [
    ("random", lambda: instances.random_lp(300, 600, 4, seed=31), 20),
    ("zipf-heavy-rows", lambda: instances.zipf(3000, 9000, 6, 1.1, seed=32), 40)])
def test_preconditioner_apply_matches_oracle(name, maker, K, fp32):
    """The apply z = M^-1 r of PCG.cpp:183-196 (L^-T D_S^-1 L^-1 with
    L = [L11 0; L21 I]) against the oracle's or_pchol_apply, which the CPU
    tests pin to the explicit dense inverse.  FP64 factor: within 1e-12
    relative.  FP32 factor (every L21 entry rounded to FP32, relative 6e-8,
    vectors and sums in FP64): within 1e-5 relative."""
    inst = maker()
    rng = np.random.default_rng(7)
    d = np.exp(rng.uniform(-2, 2, inst.n) * np.log(10.0))
    r = rng.standard_normal(inst.m)
    _flush(f"pcg apply {name} m={inst.m} K={K} fp32={fp32}")
    be = LinearEquation(mode=L.MODE_PCG, pcg_k=K, pcg_fp32_precond=fp32)
    be.construct(inst)
    be.renew(d)
    z = be.pcg_apply(r)
    z2 = be.pcg_apply(r)
    be.destruct()
    rp, ci, pos = O.csr(inst)
    cv = np.ascontiguousarray(inst.val[pos])
    pc = O.PChol()
    O.lib().or_pchol_build(inst.m, inst.n, O.i64(rp), O.i32(ci), O.d(cv), O.d(d), 0.0, 0.0, K, pc)
    perm = np.ctypeslib.as_array(pc.perm, shape=(inst.m,)).copy()
    rperm = np.ascontiguousarray(r[perm])
    zp = np.empty(inst.m)
    work = np.empty(2 * inst.m)
    O.lib().or_pchol_apply(pc, O.d(rperm), O.d(zp), O.d(work))
    O.lib().or_pchol_free(pc)
    zo = np.empty(inst.m)
    zo[perm] = zp
    err = np.linalg.norm(z - zo) / np.linalg.norm(zo)
    _flush(f"  rel err vs oracle {err:.3e}")
    assert np.array_equal(z, z2)
    assert err <= (1e-5 if fp32 else 1e-12)


def test_pcg_cap_reports_noconv():
    """A PCG solve stopped by pcg_max_iter short of pcg_tol returns
    HELP_ERR_NOCONV (with its last iterate and the residual in the stats),
    and the HSD drivers stop with STALLED / HELP_STALL_BACKEND instead of
    iterating on an unconverged direction."""
    from help_amd.backend import HelpError
    from help_amd.hsd import GpuBackend, solve_lp
    inst = instances.random_lp(300, 600, 4, seed=41)
    rng = np.random.default_rng(3)
    d = np.exp(rng.uniform(-3, 3, inst.n) * np.log(10.0))
    b1, b2 = rng.standard_normal(inst.n), rng.standard_normal(inst.m)
    be = LinearEquation(mode=L.MODE_PCG, pcg_k=10, pcg_max_iter=2)
    be.construct(inst)
    be.renew(d)
    with pytest.raises(HelpError) as ei:
        be.solve(d, b1, b2)
    assert ei.value.code == L.HELP_ERR_NOCONV
    assert be.solve_stats.pcg_iters == 2 and be.solve_stats.rel_residual > 1e-10
    be.destruct()
    for device in (True, False):
        gb = GpuBackend(mode=L.MODE_PCG, pcg_k=10, pcg_max_iter=2)
        try:
            res, _ = solve_lp(inst, gb.table, device=device)
        finally:
            gb.close()
        _flush(f"  hsd device={device}: status {res.status} stall {res.stall_reason} "
               f"backend_error {res.backend_error} iterations {res.iterations}")
        assert res.status == 1 and res.stall_reason == 3 and res.backend_error == L.HELP_ERR_NOCONV


def test_pcg_full_size_zipf_residuals():
    """The Zipf LP at its full size (m=10^6, n=4*10^6, nnz=1.9*10^7) in PCG
    mode with K=500: the backend's solution of A D^-1 A^T x2 = r (the PCG
    mode's operator has no diagonal shift, as PCG.cpp's), r = A D^-1 b1 - b2,
    checked by an independent CPU product (scipy) against the stop tolerance
    (the recursive CG residual and the true one differ by rounding), at d = 1
    and at a log-uniform d over +-1 decade, plus x1 = D^-1 (b1 - A^T x2).  The
    oracle's partial Cholesky takes too long at this size; its parity is
    pinned on the smaller instances above."""
    inst = instances.by_config(3)
    A = inst.to_scipy().tocsr()
    At = A.T.tocsr()
    rng = np.random.default_rng(5)
    be = LinearEquation(mode=L.MODE_PCG, pcg_k=500, pcg_tol=1e-10, pcg_max_iter=5000)
    be.construct(inst)
    for name, d in (("d=1", np.ones(inst.n)), ("d log-uniform +-1", np.exp(rng.uniform(-1, 1, inst.n) * np.log(10.0)))):
        b1, b2 = rng.standard_normal(inst.n), rng.standard_normal(inst.m)
        be.renew(d)
        x1, x2 = be.solve(d, b1, b2)
        r = A @ (b1 / d) - b2
        res = r - A @ ((At @ x2) / d)
        rel = np.linalg.norm(res) / np.linalg.norm(r)
        bs = np.linalg.norm(x1 - (b1 - At @ x2) / d) / np.linalg.norm(x1)
        _flush(f"zipf m={inst.m} {name}: {be.solve_stats.pcg_iters} CG iterations, relative residual {rel:.2e} "
               f"(library {be.solve_stats.rel_residual:.2e}), x1 back-substitution {bs:.1e}")
        assert rel <= 1e-9
        assert bs <= 1e-12
    be.destruct()
