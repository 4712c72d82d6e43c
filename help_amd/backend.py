"""Python mirror of HELP's linear-solver backend interface (LP.h:157-160).

`LinearEquation` keeps the reference's four calls and their meaning:

* ``construct(A)``  - LinearEquation_Construct (UseCholmod.cpp:29-58): copy A
  (CSC, rows ascending per column) and run the symbolic analysis once.
* ``renew(d)``      - RenewLinearEquation (UseCholmod.cpp:84-149): assemble
  A D^-1 A^T + beta I and factor it.
* ``solve(d, b1, b2)`` - SolveLinearEquation (UseCholmod.cpp:157-238): returns
  (x1, x2) solving [D A^T; A 0][x1; x2] = [b1; b2].
* ``destruct()``    - LinearEquation_Destruct (UseCholmod.cpp:60-81); safe to
  call at any time, idempotent.

Errors surface as ``HelpError`` carrying the library status code (the
reference would CheckError -> exit; a library must not exit the process).
"""
import ctypes as C

import numpy as np

from . import _lib as L


class HelpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[help status {code}] {msg}")
        self.code = code


class Instance:
    """An LP in HELP's final-model layout, owned by numpy arrays."""

    def __init__(self, m, n, colptr, rowind, val, b, c, n_ub=0, n_fr=0, ub=None, c_intercept=0.0):
        self.m, self.n = int(m), int(n)
        self.colptr = np.ascontiguousarray(colptr, dtype=np.int64)
        self.rowind = np.ascontiguousarray(rowind, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.b = np.ascontiguousarray(b, dtype=np.float64)
        self.c = np.ascontiguousarray(c, dtype=np.float64)
        self.n_ub, self.n_fr = int(n_ub), int(n_fr)
        self.ub = None if ub is None else np.ascontiguousarray(ub, dtype=np.float64)
        self.c_intercept = float(c_intercept)

    @property
    def nnz(self):
        return int(self.colptr[-1])

    def as_struct(self):
        s = L.LP()
        s.m, s.n, s.nnz = self.m, self.n, self.nnz
        s.n_ub, s.n_lb, s.n_fr = self.n_ub, self.n - self.n_fr, self.n_fr
        s.colptr = L.i64ptr(self.colptr)
        s.rowind = L.i32ptr(self.rowind)
        s.val = L.dptr(self.val)
        s.b = L.dptr(self.b)
        s.c = L.dptr(self.c)
        s.ub = L.dptr(self.ub) if self.ub is not None else C.cast(None, C.POINTER(C.c_double))
        s.c_intercept = self.c_intercept
        return s

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.val, self.rowind, self.colptr), shape=(self.m, self.n))


class LinearEquation:
    """The MI355X backend behind HELP's four-function interface."""

    def __init__(self, mode=L.MODE_DIRECT_FP64, beta=1e-5, symbolic_only=False, **opts):
        self.opts = L.default_opts(mode=mode, beta=beta, symbolic_only=int(symbolic_only), **opts)
        self.h = None
        self.m = self.n = 0
        self.renew_stats = L.RenewStats()
        self.solve_stats = L.SolveStats()

    def _check(self, st, what):
        if st != L.HELP_OK:
            msg = (L.lib().help_gpu_last_error(self.h) if self.h else L.lib().help_last_create_error())
            raise HelpError(st, f"{what}: {msg.decode() if msg else ''}")

    def construct(self, inst):
        self.destruct()
        h = C.c_void_p()
        st = L.lib().help_gpu_create(inst.m, inst.n, inst.nnz, L.i64ptr(inst.colptr),
                                     L.i32ptr(inst.rowind), L.dptr(inst.val), C.byref(self.opts),
                                     C.byref(h))
        self._check(st, "construct")
        self.h = h
        self.m, self.n = inst.m, inst.n
        return 0

    def destruct(self):
        if self.h:
            L.lib().help_gpu_destroy(self.h)
        self.h = None
        return 0

    def renew(self, d):
        d = np.ascontiguousarray(d, dtype=np.float64)
        self._check(L.lib().help_gpu_renew(self.h, L.dptr(d), C.byref(self.renew_stats)), "renew")

    def solve(self, d, b1, b2):
        d = np.ascontiguousarray(d, dtype=np.float64)
        b1 = np.ascontiguousarray(b1, dtype=np.float64)
        b2 = np.ascontiguousarray(b2, dtype=np.float64)
        x1 = np.empty(self.n)
        x2 = np.empty(self.m)
        self._check(L.lib().help_gpu_solve(self.h, L.dptr(d), L.dptr(b1), L.dptr(b2), L.dptr(x1),
                                           L.dptr(x2), C.byref(self.solve_stats)), "solve")
        return x1, x2

    def solve2(self, d, b1a, b2a, b1b, b2b):
        """Two solves with the same d in one call (help_gpu_solve2)."""
        d = np.ascontiguousarray(d, dtype=np.float64)
        bs = [np.ascontiguousarray(v, dtype=np.float64) for v in (b1a, b2a, b1b, b2b)]
        x1a, x1b = np.empty(self.n), np.empty(self.n)
        x2a, x2b = np.empty(self.m), np.empty(self.m)
        sb = L.SolveStats()
        self._check(L.lib().help_gpu_solve2(self.h, L.dptr(d), L.dptr(bs[0]), L.dptr(bs[1]), L.dptr(x1a),
                                            L.dptr(x2a), L.dptr(bs[2]), L.dptr(bs[3]), L.dptr(x1b),
                                            L.dptr(x2b), C.byref(self.solve_stats), C.byref(sb)), "solve2")
        return (x1a, x2a), (x1b, x2b)

    # ---- introspection for parity tests ----
    def symbolic_info(self):
        info = L.SymbolicInfo()
        self._check(L.lib().help_gpu_symbolic_info(self.h, C.byref(info)), "symbolic_info")
        return info

    def symbolic_arrays(self):
        perm = np.empty(self.m, np.int32)
        parent = np.empty(self.m, np.int32)
        cc = np.empty(self.m, np.int32)
        self._check(L.lib().help_gpu_symbolic_arrays(self.h, L.i32ptr(perm), L.i32ptr(parent),
                                                     L.i32ptr(cc)), "symbolic_arrays")
        return perm, parent, cc

    def supernodes(self):
        info = self.symbolic_info()
        ns = info.n_supernodes
        sf = np.empty(ns + 1, np.int32)
        nr = np.empty(ns, np.int32)
        lv = np.empty(ns, np.int32)
        tl = np.empty(ns, np.int32)
        self._check(L.lib().help_gpu_supernodes(self.h, L.i32ptr(sf), L.i32ptr(nr), L.i32ptr(lv),
                                                L.i32ptr(tl)), "supernodes")
        return sf, nr, lv, tl

    def set_warmstart(self, net):
        """HELP_MODE_PCG_LEARNED: start PCG from ``net``'s guess (a Mpnn)."""
        self._check(L.lib().help_gpu_set_warmstart(self.h, net.ptr if net is not None else None), "set_warmstart")
        self._net = net

    def pcg_factor(self):
        """PCG mode: (perm, D, S, L) of the partial-Cholesky preconditioner,
        L as an m x K array by original row."""
        K = C.c_int32()
        self._check(L.lib().help_gpu_pcg_factor(self.h, C.byref(K), None, None, None, None), "pcg_factor")
        k = K.value
        perm = np.empty(self.m, np.int32)
        D = np.empty(max(k, 1))
        S = np.empty(self.m)
        Lf = np.empty(max(self.m * k, 1))
        self._check(L.lib().help_gpu_pcg_factor(self.h, C.byref(K), L.i32ptr(perm), L.dptr(D), L.dptr(S),
                                                L.dptr(Lf)), "pcg_factor")
        return perm, D[:k], S, Lf[:self.m * k].reshape(k, self.m).T.copy()

    def pcg_apply(self, r):
        """PCG mode: M^-1 r with the preconditioner of the last renew."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty(self.m)
        self._check(L.lib().help_gpu_pcg_apply(self.h, L.dptr(r), L.dptr(z)), "pcg_apply")
        return z

    def adat_pattern(self):
        info = self.symbolic_info()
        rp = np.empty(self.m + 1, np.int64)
        ci = np.empty(info.nnz_adat_upper, np.int32)
        self._check(L.lib().help_gpu_adat_pattern(self.h, L.i64ptr(rp), L.i32ptr(ci)), "adat_pattern")
        return rp, ci

    def adat_values(self, d):
        info = self.symbolic_info()
        d = np.ascontiguousarray(d, dtype=np.float64)
        v = np.empty(info.nnz_adat_upper)
        self._check(L.lib().help_gpu_adat_values(self.h, L.dptr(d), L.dptr(v)), "adat_values")
        return v

    def spmv(self, x, trans=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n if trans else self.m)
        self._check(L.lib().help_gpu_spmv(self.h, int(trans), L.dptr(x), L.dptr(y)), "spmv")
        return y

    def factor_column(self, j):
        rows = np.empty(self.m, np.int32)
        vals = np.empty(self.m)
        nr = C.c_int32()
        self._check(L.lib().help_gpu_factor_column(self.h, j, L.i32ptr(rows), L.dptr(vals),
                                                   C.byref(nr)), "factor_column")
        return rows[:nr.value].copy(), vals[:nr.value].copy()

    def __del__(self):
        try:
            self.destruct()
        except Exception:
            pass


def device_count():
    return L.lib().help_device_count()
