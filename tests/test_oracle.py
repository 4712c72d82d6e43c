"""The CPU oracle pinned against the reference's own golden vectors and
against what the mathematics fixes (closed forms, exact arithmetic, dense
textbook factorizations).  CPU only."""
import numpy as np
import pytest

import oracle_bind as O
from help_amd import instances


def _csr(rows, ncol):
    """CSR arrays from a list of {col: val} dicts."""
    rp, ci, cv = [0], [], []
    for r in rows:
        for c in sorted(r):
            ci.append(c)
            cv.append(r[c])
        rp.append(len(ci))
    return (np.array(rp, np.int64), np.array(ci, np.int32), np.array(cv, np.float64))


def _dense(inst):
    A = np.zeros((inst.m, inst.n))
    for j in range(inst.n):
        for p in range(inst.colptr[j], inst.colptr[j + 1]):
            A[inst.rowind[p], j] = inst.val[p]
    return A


# ------------------------------------------------------------- golden vectors
def test_mkltest_csrmv_known_answer():
    """Main.cpp:19-35 (MKLTest): [1 0 2; 0 3 0; 4 0 5] * [3 5 7] = [17 15 47]."""
    rp, ci, cv = _csr([{0: 1, 2: 2}, {1: 3}, {0: 4, 2: 5}], 3)
    x = np.array([3.0, 5.0, 7.0])
    y = np.zeros(3)
    O.lib().or_csrmv_n(3, 1.0, O.d(cv), O.i32(ci), O.i64(rp), O.d(x), 0.0, O.d(y))
    assert y.tolist() == [17.0, 15.0, 47.0]
    # transpose: A^T x = [1*3+4*7, 3*5, 2*3+5*7]
    yt = np.zeros(3)
    O.lib().or_csrmv_t(3, 3, 1.0, O.d(cv), O.i32(ci), O.i64(rp), O.d(x), 0.0, O.d(yt))
    assert yt.tolist() == [31.0, 15.0, 41.0]


def _unittest_matrix():
    """PCG.cpp:404-413 (UnitTest) with its 1-based column indices made
    0-based: A = M (5x5 lower), d = 1/diag([4 1 5 3 2])^2, so
    G = A D^-1 A^T = M diag([4 1 5 3 2])^2 M^T."""
    M = [[1], [2, 1], [3, 5, 1], [2, 4, 3, 1], [3, 3, 0, 2, 1]]
    rows = [{c: float(v) for c, v in enumerate(r)} for r in M]
    rp, ci, cv = _csr(rows, 5)
    d = np.array([0.25 * 0.25, 1, 0.2 * 0.2, 1.0 / 9.0, 0.25])
    Md = np.zeros((5, 5))
    for i, r in enumerate(M):
        Md[i, :len(r)] = r
    G = Md @ np.diag(1.0 / d) @ Md.T
    return rp, ci, cv, d, G


def _pchol(rp, ci, cv, d, m, n, k):
    pc = O.PChol()
    O.lib().or_pchol_build(m, n, O.i64(rp), O.i32(ci), O.d(cv), O.d(d), 0.0, 0.0, k, pc)
    perm = np.ctypeslib.as_array(pc.perm, shape=(m,)).copy()
    L = np.ctypeslib.as_array(pc.L, shape=(m * k,)).reshape(m, k).copy()
    D = np.ctypeslib.as_array(pc.D, shape=(k,)).copy()
    S = np.ctypeslib.as_array(pc.diag_s, shape=(m,)).copy()
    return pc, perm, L, D, S


def test_pcg_unittest_golden_k1():
    """PCG.cpp:415-437: the MATLAB values for one pivot.  The first pivot is
    the largest diagonal, 314, at row 3 (MATLAB's G(1) = 4), and
    inv(L) G(p,p) inv(L)' has Schur diagonal 50.2739, 77.8185, 12.7389,
    142.4395 on the remaining rows of p = [3 1 2 0 4]."""
    rp, ci, cv, d, G = _unittest_matrix()
    pc, perm, L, D, S = _pchol(rp, ci, cv, d, 5, 5, 1)
    try:
        assert perm.tolist() == [3, 1, 2, 0, 4]
        assert D[0] == pytest.approx(314.0, rel=1e-15)
        col = L[perm, 0]
        matlab = [1.0000, 0.2166, 0.6083, 0.1019, 0.4013]
        assert np.allclose(col, matlab, atol=5e-5)
        # exact fractions: G(i,3)/314
        assert np.allclose(col, G[perm, 3] / 314.0, rtol=1e-15, atol=0)
        schur = [50.2739, 77.8185, 12.7389, 142.4395]
        assert np.allclose(S[perm[1:]], schur, atol=5e-5)
    finally:
        O.lib().or_pchol_free(pc)


def test_pcg_unittest_golden_k2_matches_dense_cholesky():
    """PCG.cpp:439-456: two pivots, p = [3 4 2 0 1] (MATLAB G = [4 5 3 1 2]);
    the two columns equal the unit-scaled columns of chol(G(p,p))'."""
    rp, ci, cv, d, G = _unittest_matrix()
    pc, perm, L, D, S = _pchol(rp, ci, cv, d, 5, 5, 2)
    try:
        assert perm.tolist() == [3, 4, 2, 0, 1]
        l = np.linalg.cholesky(G[np.ix_(perm, perm)])
        for k in range(2):
            assert np.allclose(L[perm, k][k:], (l[:, k] / l[k, k])[k:], rtol=1e-13, atol=1e-15)
            assert D[k] == pytest.approx(l[k, k] ** 2, rel=1e-13)
        # Schur diagonal after two pivots: diag(G(p,p) - L D L^T) on rows >= 2
        Lk = L[perm, :]
        Sch = G[np.ix_(perm, perm)] - Lk @ np.diag(D) @ Lk.T
        assert np.allclose(S[perm[2:]], np.diag(Sch)[2:], rtol=1e-12)
    finally:
        O.lib().or_pchol_free(pc)


# -------------------------------------------------------- ADA^T assembly
def test_adat_values_closed_form(small_mcf):
    inst = small_mcf
    rng = np.random.default_rng(5)
    d = np.exp(rng.uniform(-3, 3, inst.n))
    arp, aci = O.adat_pattern(inst)
    vals = O.adat_values(inst, d, arp, aci)
    A = _dense(inst)
    M = A @ np.diag(1.0 / d) @ A.T
    for i in range(inst.m):
        cols = aci[arp[i]:arp[i + 1]]
        assert cols[0] == i                      # diagonal first (Helper.cpp:273-329)
        assert np.all(cols >= i)
        assert len(set(cols.tolist())) == len(cols)
        assert set(cols.tolist()) == set(np.nonzero(np.abs(M[i, i:]) > 0)[0] + i) | {i}
        assert np.allclose(vals[arp[i]:arp[i + 1]], M[i, cols], rtol=1e-13, atol=1e-12)


def test_adat_values_exact_arithmetic():
    """Integer A and power-of-two d: every product and partial sum is exact,
    so the assembly equals the dense product bitwise."""
    rng = np.random.default_rng(11)
    inst = instances.random_lp(60, 120, 4, seed=7)
    inst.val[:] = rng.integers(-3, 4, inst.nnz).astype(np.float64)
    inst.val[inst.val == 0] = 1.0
    d = 2.0 ** rng.integers(-4, 5, inst.n)
    arp, aci = O.adat_pattern(inst)
    vals = O.adat_values(inst, d, arp, aci)
    A = _dense(inst)
    M = A @ np.diag(1.0 / d) @ A.T
    for i in range(inst.m):
        cols = aci[arp[i]:arp[i + 1]]
        assert np.array_equal(vals[arp[i]:arp[i + 1]], M[i, cols])


# -------------------------------------------------- factor and solves
def _factor(inst, d, beta, perm):
    arp, aci = O.adat_pattern(inst)
    aval = O.adat_values(inst, d, arp, aci)
    parent, cc, Lp, Li = O.symbolic(inst.m, arp, aci, perm)
    st, Lx = O.cholesky(inst.m, arp, aci, aval, beta, perm, Lp, Li)
    return st, Lp, Li, Lx


def test_cholesky_equals_dense_textbook():
    inst = instances.random_lp(80, 160, 4, seed=2)
    rng = np.random.default_rng(3)
    d = np.exp(rng.uniform(-2, 2, inst.n))
    perm = rng.permutation(inst.m).astype(np.int32)
    beta = 1e-5
    st, Lp, Li, Lx = _factor(inst, d, beta, perm)
    assert st == 0
    A = _dense(inst)
    M = A @ np.diag(1.0 / d) @ A.T + beta * np.eye(inst.m)
    Lref = np.linalg.cholesky(M[np.ix_(perm, perm)])
    L = np.zeros_like(Lref)
    for j in range(inst.m):
        L[Li[Lp[j]:Lp[j + 1]], j] = Lx[Lp[j]:Lp[j + 1]]
    assert np.allclose(L, Lref, rtol=1e-12, atol=1e-12 * np.abs(Lref).max())
    # structural zeros of the dense factor lie outside the symbolic pattern
    mask = np.zeros_like(L, bool)
    for j in range(inst.m):
        mask[Li[Lp[j]:Lp[j + 1]], j] = True
    assert np.all(np.abs(Lref[~mask]) < 1e-12 * np.abs(Lref).max())


def test_solve_kkt_against_dense(cfg0):
    """SolveLinearEquation semantics (UseCholmod.cpp:157-238): beta on the
    factor only, RHS = A D^-1 b1 - b2, x1 back-substituted."""
    inst = instances.random_lp(150, 300, 4, seed=4)
    rng = np.random.default_rng(9)
    d = np.exp(rng.uniform(-3, 3, inst.n))
    perm = np.arange(inst.m, dtype=np.int32)
    beta = 1e-5
    st, Lp, Li, Lx = _factor(inst, d, beta, perm)
    b1 = rng.standard_normal(inst.n)
    b2 = rng.standard_normal(inst.m)
    x1, x2 = O.solve_kkt(inst, perm, Lp, Li, Lx, d, b1, b2)
    A = _dense(inst)
    M = A @ np.diag(1.0 / d) @ A.T + beta * np.eye(inst.m)
    x2r = np.linalg.solve(M, A @ (b1 / d) - b2)
    x1r = (b1 - A.T @ x2r) / d
    assert np.allclose(x2, x2r, rtol=1e-9, atol=1e-10 * np.abs(x2r).max())
    assert np.allclose(x1, x1r, rtol=1e-9, atol=1e-10 * np.abs(x1r).max())


# ------------------------------------------------------------------- PCG
def test_preconditioner_equals_dense_inverse():
    """PCG.cpp:183-196 against the explicit L^-T D^-1 L^-1 (the idea of
    PCGDebugHelper, PCG.cpp:209-262), with L = [L11 0; L21 I]."""
    inst = instances.random_lp(40, 80, 4, seed=5)
    rng = np.random.default_rng(1)
    d = np.exp(rng.uniform(-1, 1, inst.n))
    rp, ci, pos = O.csr(inst)
    cv = inst.val[pos]
    m, K = inst.m, 6
    pc, perm, L, D, S = _pchol(rp, ci, cv, d, m, inst.n, K)
    try:
        Lp = np.eye(m)
        Lp[:, :K] = L[perm, :]
        Dfull = np.concatenate([D, S[perm[K:]]])
        Minv = np.linalg.inv(Lp).T @ np.diag(1.0 / Dfull) @ np.linalg.inv(Lp)
        r = rng.standard_normal(m)
        out = np.empty(m)
        work = np.empty(2 * m)
        O.lib().or_pchol_apply(pc, O.d(r), O.d(out), O.d(work))
        assert np.allclose(out, Minv @ r, rtol=1e-10, atol=1e-12)
    finally:
        O.lib().or_pchol_free(pc)


def _pcg(inst, d, K, b, x0=None, tol=1e-12, max_iter=1000):
    rp, ci, pos = O.csr(inst)
    cv = np.ascontiguousarray(inst.val[pos])
    pc, perm, L, D, S = _pchol(rp, ci, cv, d, inst.m, inst.n, K)
    x = np.empty(inst.m)
    x0 = np.zeros(inst.m) if x0 is None else x0
    rel = np.zeros(1)
    it = O.lib().or_pcg(inst.m, inst.n, O.i64(rp), O.i32(ci), O.d(cv), O.d(d), 0.0, 0.0, pc, O.d(b),
                        O.d(x0), O.d(x), tol, max_iter, O.d(rel))
    O.lib().or_pchol_free(pc)
    return it, x, rel[0]


def test_pcg_matches_direct_solve():
    inst = instances.random_lp(120, 240, 4, seed=6)
    rng = np.random.default_rng(2)
    d = np.exp(rng.uniform(-2, 2, inst.n))
    b = rng.standard_normal(inst.m)
    A = _dense(inst)
    G = A @ np.diag(1.0 / d) @ A.T
    xr = np.linalg.solve(G, b)
    for K in (0, 10, 40):
        it, x, rel = _pcg(inst, d, K, b)
        assert rel < 1e-12
        assert np.allclose(x, xr, rtol=1e-8, atol=1e-9 * np.abs(xr).max()), K


def test_pcg_finite_termination_and_exact_preconditioner():
    """CG terminates in at most m steps on an SPD system; with K = m the
    partial Cholesky is the full factor and one step suffices."""
    inst = instances.random_lp(12, 24, 3, seed=8)
    rng = np.random.default_rng(4)
    d = np.exp(rng.uniform(-1, 1, inst.n))
    b = rng.standard_normal(inst.m)
    it, x, rel = _pcg(inst, d, 0, b, tol=1e-13, max_iter=200)
    assert it <= inst.m + 2 and rel < 1e-12
    it, x, rel = _pcg(inst, d, inst.m, b, tol=1e-12, max_iter=200)
    assert it <= 2 and rel < 1e-12


def _ref_helper():
    import ctypes as C
    import os
    path = os.path.join(O.REF_DIR, "libref_helper.so")
    if not os.path.exists(path):
        pytest.skip("reference Helper harness not built (no reference sources)")
    lib = C.CDLL(path)
    D, I32, I64 = C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    lib.ref_load.restype = C.c_int
    lib.ref_load.argtypes = [C.c_int32, C.c_int32, I64, I32, D]
    lib.ref_adat_pattern.restype = C.c_int64
    lib.ref_adat_pattern.argtypes = [I64, I32]
    lib.ref_adat_values.restype = None
    lib.ref_adat_values.argtypes = [D, D]
    lib.ref_a_times.restype = None
    lib.ref_a_times.argtypes = [C.c_int32, C.c_int32, D, D]
    return lib


@pytest.mark.parametrize("maker", [
    lambda: instances.random_lp(300, 600, 4, seed=71),
    lambda: instances.mcf(side=8, commodities=2, pairs=4, seed=72),
    lambda: instances.staircase(5, 40, 120, 5, seed=73),
    lambda: instances.zipf(2000, 6000, 6, 1.1, seed=74),
    lambda: instances.random_lp(300, 600, 4, seed=75, ub_cols=100, free_cols=5)])
def test_oracle_matches_reference_helper_bitwise(maker):
    """The oracle's ADA^T pattern (or_adat_count/fill) and values
    (or_adat_values) and its products with A and A^T (or_a_times,
    or_at_times) against HELP's own Helper.cpp compiled from the reference
    sources (ADAt_Allocate / ADAt_Calc in the linker form UsePardiso.cpp
    calls, Helper.cpp:273-377; SetATimesVector, Helper.cpp:110-156), bit for
    bit, on d spread over 16 decades and both signs."""
    import ctypes as C
    lib = _ref_helper()
    inst = maker()
    assert lib.ref_load(inst.m, inst.n, O.i64(inst.colptr), O.i32(inst.rowind), O.d(inst.val)) == 0
    rp = np.empty(inst.m + 1, np.int64)
    nnz = lib.ref_adat_pattern(O.i64(rp), C.cast(None, C.POINTER(C.c_int32)))
    ci = np.empty(nnz, np.int32)
    lib.ref_adat_pattern(O.i64(rp), O.i32(ci))
    arp, aci = O.adat_pattern(inst)
    assert np.array_equal(arp, rp) and np.array_equal(aci, ci)
    rng = np.random.default_rng(76)
    d = np.exp(rng.uniform(-8, 8, inst.n) * np.log(10.0))
    ref_vals = np.empty(nnz)
    lib.ref_adat_values(O.d(d), O.d(ref_vals))
    vals = O.adat_values(inst, d, arp, aci)
    assert np.array_equal(vals.view(np.uint64), ref_vals.view(np.uint64))
    for trans in (0, 1):
        nin, nout = (inst.n, inst.m) if trans == 0 else (inst.m, inst.n)
        v = rng.standard_normal(nin)
        y0 = rng.standard_normal(nout)
        for sign in (1, -1):
            yr = y0.copy()
            lib.ref_a_times(trans, sign, O.d(v), O.d(yr))
            yo = y0.copy()
            f = O.lib().or_a_times if trans == 0 else O.lib().or_at_times
            f(inst.m, inst.n, O.i64(inst.colptr), O.i32(inst.rowind), O.d(inst.val), sign, O.d(v), O.d(yo))
            assert np.array_equal(yo.view(np.uint64), yr.view(np.uint64)), (trans, sign)
