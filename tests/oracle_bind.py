"""ctypes binding of oracle/liboracle.so (the CPU FP64 checker; tests only)."""
import ctypes as C
import os

import numpy as np

from help_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)


class PChol(C.Structure):
    _fields_ = [("m", C.c_int32), ("k", C.c_int32), ("perm", _I32), ("L", _D), ("D", _D),
                ("diag_s", _D), ("l11inv", _D)]


_sig = {
    "or_csc_to_csr": (None, [C.c_int32, C.c_int32, _I64, _I32, _I64, _I32, _I64]),
    "or_adat_count": (C.c_int64, [C.c_int32, C.c_int32, _I64, _I32, _I64, _I32, _I64, _I64]),
    "or_adat_fill": (None, [C.c_int32, C.c_int32, _I64, _I32, _I64, _I32, _I64, _I64, _I32]),
    "or_adat_values": (None, [C.c_int32, C.c_int32, _I64, _I32, _D, _I64, _I32, _I64, _D, _I64,
                              _I32, _D]),
    "or_csrmv_n": (None, [C.c_int32, C.c_double, _D, _I32, _I64, _D, C.c_double, _D]),
    "or_csrmv_t": (None, [C.c_int32, C.c_int32, C.c_double, _D, _I32, _I64, _D, C.c_double, _D]),
    "or_a_times": (None, [C.c_int32, C.c_int32, _I64, _I32, _D, C.c_int32, _D, _D]),
    "or_at_times": (None, [C.c_int32, C.c_int32, _I64, _I32, _D, C.c_int32, _D, _D]),
    "or_symbolic": (C.c_int64, [C.c_int32, _I64, _I32, _I32, _I32, _I32, _I64, _I32]),
    "or_cholesky": (C.c_int, [C.c_int32, _I64, _I32, _D, C.c_double, _I32, _I64, _I32, _D]),
    "or_chol_solve": (None, [C.c_int32, _I32, _I64, _I32, _D, _D, _D, _D]),
    "or_solve_kkt": (None, [C.c_int32, C.c_int32, _I64, _I32, _D, _I32, _I64, _I32, _D, _D, _D,
                            _D, _D, _D, _D]),
    "or_pchol_build": (C.c_int, [C.c_int32, C.c_int32, _I64, _I32, _D, _D, C.c_double,
                                 C.c_double, C.c_int32, C.POINTER(PChol)]),
    "or_pchol_free": (None, [C.POINTER(PChol)]),
    "or_pchol_apply": (None, [C.POINTER(PChol), _D, _D, _D]),
    "or_pcg": (C.c_int, [C.c_int32, C.c_int32, _I64, _I32, _D, _D, C.c_double, C.c_double,
                         C.POINTER(PChol), _D, _D, _D, C.c_double, C.c_int32, _D]),
    "or_backend_new": (C.c_void_p, [C.c_double, _I32, C.c_int32]),
    "or_backend_free": (None, [C.c_void_p]),
    "or_be_seconds_renew": (C.c_double, [C.c_void_p]),
    "or_be_seconds_solve": (C.c_double, [C.c_void_p]),
}

_o = None


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run __graft_entry__.build()")
        o = C.CDLL(ORACLE_SO)
        for k, (r, a) in _sig.items():
            f = getattr(o, k)
            f.restype = r
            f.argtypes = a
        _o = o
    return _o


def d(a):
    return L.dptr(a)


def i32(a):
    return L.i32ptr(a)


def i64(a):
    return L.i64ptr(a)


def csr(inst):
    rp = np.empty(inst.m + 1, np.int64)
    ci = np.empty(inst.nnz, np.int32)
    pos = np.empty(inst.nnz, np.int64)
    lib().or_csc_to_csr(inst.m, inst.n, i64(inst.colptr), i32(inst.rowind), i64(rp), i32(ci), i64(pos))
    return rp, ci, pos


def adat_pattern(inst):
    rp, ci, pos = csr(inst)
    arp = np.empty(inst.m + 1, np.int64)
    nup = lib().or_adat_count(inst.m, inst.n, i64(inst.colptr), i32(inst.rowind), i64(rp), i32(ci),
                              i64(pos), i64(arp))
    aci = np.empty(nup, np.int32)
    lib().or_adat_fill(inst.m, inst.n, i64(inst.colptr), i32(inst.rowind), i64(rp), i32(ci), i64(pos),
                       i64(arp), i32(aci))
    return arp, aci


def adat_values(inst, dvec, arp, aci):
    rp, ci, pos = csr(inst)
    out = np.empty(len(aci))
    dvec = np.ascontiguousarray(dvec, dtype=np.float64)
    lib().or_adat_values(inst.m, inst.n, i64(inst.colptr), i32(inst.rowind), d(inst.val), i64(rp),
                         i32(ci), i64(pos), d(dvec), i64(arp), i32(aci), d(out))
    return out


def symbolic(m, arp, aci, perm):
    parent = np.empty(m, np.int32)
    cc = np.empty(m, np.int32)
    Lp = np.empty(m + 1, np.int64)
    nl = lib().or_symbolic(m, i64(arp), i32(aci), i32(perm), i32(parent), i32(cc), i64(Lp),
                           C.cast(None, _I32))
    Li = np.empty(nl, np.int32)
    lib().or_symbolic(m, i64(arp), i32(aci), i32(perm), i32(parent), i32(cc), i64(Lp), i32(Li))
    return parent, cc, Lp, Li


def cholesky(m, arp, aci, aval, beta, perm, Lp, Li):
    Lx = np.empty(len(Li))
    st = lib().or_cholesky(m, i64(arp), i32(aci), d(aval), beta, i32(perm), i64(Lp), i32(Li), d(Lx))
    return st, Lx


def solve_kkt(inst, perm, Lp, Li, Lx, dvec, b1, b2):
    x1 = np.empty(inst.n)
    x2 = np.empty(inst.m)
    work = np.empty(inst.m + inst.n)
    lib().or_solve_kkt(inst.m, inst.n, i64(inst.colptr), i32(inst.rowind), d(inst.val), i32(perm),
                       i64(Lp), i32(Li), d(Lx), d(dvec), d(b1), d(b2), d(x1), d(x2), d(work))
    return x1, x2


class OracleBackend:
    """The CPU FP64 oracle as a help_backend table (CHOLMOD semantics)."""

    def __init__(self, beta=1e-5, perm=None):
        if perm is not None:
            self._perm = np.ascontiguousarray(perm, dtype=np.int32)
            self.ptr = lib().or_backend_new(beta, i32(self._perm), len(self._perm))
        else:
            self.ptr = lib().or_backend_new(beta, C.cast(None, _I32), 0)
        o = lib()
        self.table = L.Backend()
        self.table.ctx = self.ptr
        self.table.construct = C.cast(o.or_be_construct, C.c_void_p)
        self.table.destruct = C.cast(o.or_be_destruct, C.c_void_p)
        self.table.renew = C.cast(o.or_be_renew, C.c_void_p)
        self.table.solve = C.cast(o.or_be_solve, C.c_void_p)

    def seconds(self):
        return lib().or_be_seconds_renew(self.ptr), lib().or_be_seconds_solve(self.ptr)

    def close(self):
        if self.ptr:
            lib().or_backend_free(self.ptr)
            self.ptr = None
