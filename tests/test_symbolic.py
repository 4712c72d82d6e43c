"""Symbolic analysis of the backend (host C++, no GPU): ADA^T pattern in
HELP order, AMD ordering, elimination tree, column counts, supernodes.
Pinned by the oracle's independent structure-merging symbolic factorization,
brute-force dense elimination, and scipy's MMD ordering for fill quality."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle_bind as O
from help_amd import LinearEquation, instances


def _analyze(inst):
    be = LinearEquation(symbolic_only=True)
    be.construct(inst)
    return be


def _pattern_dense(inst):
    A = sp.csc_matrix((np.ones(inst.nnz), inst.rowind, inst.colptr), shape=(inst.m, inst.n))
    M = (A @ A.T).toarray() != 0
    return M


@pytest.mark.parametrize("maker", [lambda: instances.random_lp(200, 400, 4, seed=1),
                                   lambda: instances.mcf(side=8, commodities=2, pairs=4, seed=2)])
def test_adat_pattern_is_the_reference_order(maker):
    """The product's upper ADA^T pattern equals the oracle's linker-form
    pattern entry for entry (Helper.cpp:273-329: diagonal first, then
    discovery order)."""
    inst = maker()
    be = _analyze(inst)
    rp, ci = be.adat_pattern()
    arp, aci = O.adat_pattern(inst)
    assert np.array_equal(rp, arp)
    assert np.array_equal(ci, aci)
    be.destruct()


def test_ordering_is_valid_and_deterministic(cfg0):
    be1, be2 = _analyze(cfg0), _analyze(cfg0)
    p1, _, _ = be1.symbolic_arrays()
    p2, _, _ = be2.symbolic_arrays()
    assert np.array_equal(np.sort(p1), np.arange(cfg0.m))
    assert np.array_equal(p1, p2)


@pytest.mark.parametrize("maker", [lambda: instances.random_lp(400, 800, 4, seed=3),
                                   lambda: instances.mcf(side=10, commodities=3, pairs=5, seed=4),
                                   lambda: instances.staircase(6, 40, 120, 5, seed=5)])
def test_etree_and_colcounts_match_structure_merging(maker):
    inst = maker()
    be = _analyze(inst)
    perm, parent, cc = be.symbolic_arrays()
    arp, aci = O.adat_pattern(inst)
    par_o, cc_o, Lp, Li = O.symbolic(inst.m, arp, aci, perm)
    assert np.array_equal(parent, par_o)
    assert np.array_equal(cc, cc_o)
    assert be.symbolic_info().nnz_l_exact == int(cc.sum())


def test_etree_and_colcounts_brute_force():
    """Dense boolean elimination on a tiny instance: parent(j) is the first
    off-diagonal nonzero of column j of L, colcount(j) its length."""
    inst = instances.random_lp(40, 80, 3, seed=9)
    be = _analyze(inst)
    perm, parent, cc = be.symbolic_arrays()
    M = _pattern_dense(inst)[np.ix_(perm, perm)]
    m = inst.m
    F = M.copy()
    for j in range(m):
        rows = np.nonzero(F[j + 1:, j])[0] + j + 1
        for a in rows:
            F[rows, a] = True
        exp_parent = rows[0] if len(rows) else -1
        assert parent[j] == exp_parent, j
        assert cc[j] == 1 + len(rows), j


def test_supernode_partition_invariants(small_mcf):
    be = _analyze(small_mcf)
    perm, parent, cc = be.symbolic_arrays()
    sf, nr, lv, _ = be.supernodes()
    info = be.symbolic_info()
    m = small_mcf.m
    assert sf[0] == 0 and sf[-1] == m and np.all(np.diff(sf) > 0)
    for s in range(len(nr)):
        f, l = sf[s], sf[s + 1]
        assert nr[s] >= l - f
        # the relaxed structure holds every column's exact structure
        for j in range(f, l):
            assert cc[j] <= nr[s] - (j - f)
        # columns of a supernode are a chain of the elimination tree
        for j in range(f, l - 1):
            assert parent[j] == j + 1
        # levels: a parent supernode sits above its children
        if parent[l - 1] >= 0:
            ps = np.searchsorted(sf, parent[l - 1], side="right") - 1
            assert lv[ps] > lv[s]
    assert info.nnz_l >= info.nnz_l_exact


def test_fill_is_close_to_mmd():
    """Fill of the AMD ordering within 30% of SuperLU's MMD on A A^T."""
    inst = instances.mcf(side=20, commodities=3, pairs=8, seed=6)
    be = _analyze(inst)
    nnz_l = be.symbolic_info().nnz_l_exact
    A = sp.csc_matrix((inst.val, inst.rowind, inst.colptr), shape=(inst.m, inst.n))
    M = (A @ A.T + 1e-3 * sp.eye(inst.m)).tocsc()
    lu = spla.splu(M, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                   options=dict(SymmetricMode=True))
    assert nnz_l <= 1.3 * lu.L.nnz


def _analyze_nd(inst):
    be = LinearEquation(symbolic_only=True, ordering=1)
    be.construct(inst)
    return be


def test_nd_ordering_is_valid_and_deterministic():
    inst = instances.staircase(8, 400, 1200, 5, seed=11)
    p1, _, _ = _analyze_nd(inst).symbolic_arrays()
    p2, _, _ = _analyze_nd(inst).symbolic_arrays()
    assert np.array_equal(np.sort(p1), np.arange(inst.m))
    assert np.array_equal(p1, p2)


def test_nd_etree_and_colcounts_match_structure_merging():
    inst = instances.staircase(8, 400, 1200, 5, seed=12)
    be = _analyze_nd(inst)
    perm, parent, cc = be.symbolic_arrays()
    arp, aci = O.adat_pattern(inst)
    par_o, cc_o, Lp, Li = O.symbolic(inst.m, arp, aci, perm)
    assert np.array_equal(parent, par_o)
    assert np.array_equal(cc, cc_o)


def test_nd_shortens_the_staircase_elimination_tree():
    """On a chain of period blocks, dissection gives an elimination tree of
    logarithmic depth where minimum degree gives one about as deep as the
    chain, here (1000-row periods) with less fill."""
    inst = instances.staircase(40, 1000, 3000, 5, seed=13)
    amd = _analyze(inst).symbolic_info()
    nd = _analyze_nd(inst).symbolic_info()
    assert nd.n_levels * 3 <= amd.n_levels
    assert nd.nnz_l_exact <= amd.nnz_l_exact
