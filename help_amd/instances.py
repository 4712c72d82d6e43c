"""Deterministic LP generators for the benchmark workloads (C++, own RNG).

Workloads (sizes as named by the benchmark plan):
* random     - A=[B | I_m], B m x nb with k N(0,1) nnz per column at distinct
               uniform rows; x0/y0/s0 feasibility construction.  Optionally
               upper bounds on the first columns of B and free columns after
               the identity.
* mcf        - block-angular multicommodity flow on a side x side 4-connected
               grid with K commodities, 20 source-sink pairs each.
* staircase  - T periods of 1000 x 3000 blocks, 5 nnz per column in-block plus
               one linking nnz into the next period, plus I_m.
* zipf       - A=[B | I_m], 6 nnz per column with Zipf(s) row popularity.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .backend import HelpError, Instance


def _generate(params):
    lp = L.LP()
    st = L.lib().help_generate(C.byref(params), C.byref(lp))
    if st != L.HELP_OK:
        raise HelpError(st, "help_generate failed")
    return _from_lp(lp)


def _from_lp(lp):
    """numpy copy of a library-allocated help_lp, which is then freed"""
    try:
        n, m, nnz = lp.n, lp.m, lp.nnz
        colptr = np.ctypeslib.as_array(lp.colptr, shape=(n + 1,)).copy()
        rowind = np.ctypeslib.as_array(lp.rowind, shape=(max(nnz, 1),))[:nnz].copy()
        val = np.ctypeslib.as_array(lp.val, shape=(max(nnz, 1),))[:nnz].copy()
        b = np.ctypeslib.as_array(lp.b, shape=(m,)).copy()
        c = np.ctypeslib.as_array(lp.c, shape=(n,)).copy()
        ub = np.ctypeslib.as_array(lp.ub, shape=(lp.n_ub,)).copy() if lp.n_ub > 0 else None
        n_ub, n_fr, icpt = lp.n_ub, lp.n_fr, lp.c_intercept
    finally:
        L.lib().help_lp_free(C.byref(lp))
    return Instance(m, n, colptr, rowind, val, b, c, n_ub=n_ub, n_fr=n_fr, ub=ub, c_intercept=icpt)


def random_lp(m=2000, nb=4000, k=4, seed=0, ub_cols=0, free_cols=0):
    """ub_cols > 0 bounds the first ub_cols columns of B above; free_cols > 0
    appends that many free columns (HELP's [UB | LB-only | free] layout)."""
    p = L.GenParams(kind=L.GEN_RANDOM, seed=seed, m=m, nb=nb, k=k, ub_cols=ub_cols, free_cols=free_cols)
    return _generate(p)


def mcf(side=150, commodities=8, pairs=20, seed=0):
    p = L.GenParams(kind=L.GEN_MCF, seed=seed, side=side, commodities=commodities, pairs=pairs)
    return _generate(p)


def staircase(periods=200, block_rows=1000, block_cols=3000, block_k=5, seed=0):
    p = L.GenParams(kind=L.GEN_STAIRCASE, seed=seed, periods=periods, block_rows=block_rows,
                    block_cols=block_cols, block_k=block_k)
    return _generate(p)


def zipf(m=1_000_000, nb=3_000_000, k=6, s=1.1, seed=0):
    p = L.GenParams(kind=L.GEN_ZIPF, seed=seed, m=m, nb=nb, k=k, zipf_s=s)
    return _generate(p)


def write_mps(inst, path, name="HELPGEN"):
    s = inst.as_struct()
    st = L.lib().help_write_mps(C.byref(s), path.encode(), name.encode())
    if st != L.HELP_OK:
        raise HelpError(st, "help_write_mps failed")


def write_csc(inst, path):
    """Binary CSC instance file (help_write_csc_bin), for LPs beyond MPS."""
    s = inst.as_struct()
    st = L.lib().help_write_csc_bin(C.byref(s), path.encode())
    if st != L.HELP_OK:
        raise HelpError(st, "help_write_csc_bin failed")


def read_csc(path):
    lp = L.LP()
    st = L.lib().help_read_csc_bin(path.encode(), C.byref(lp))
    if st != L.HELP_OK:
        raise HelpError(st, f"help_read_csc_bin failed on {path}")
    return _from_lp(lp)


def by_config(cfg, seed=0):
    """The benchmark plan's configurations by index (0..4); 4 is a held-out
    instance of the warm-start family (grids of side 60-150 with 2-8
    commodities; the training instances use seeds 1000 and up, the held-out
    ones 2000 and up): the smallest of its first eight, side 76 with K=2."""
    if cfg == 0:
        return random_lp(2000, 4000, 4, seed)
    if cfg == 1:
        return mcf(150, 8, 20, seed)
    if cfg == 2:
        return staircase(200, 1000, 3000, 5, seed)
    if cfg == 3:
        return zipf(1_000_000, 3_000_000, 6, 1.1, seed)
    if cfg == 4:
        return mcf(76, 2, 20, 2002 + seed)
    raise ValueError(cfg)
