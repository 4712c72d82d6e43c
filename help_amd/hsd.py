"""HSD interior-point driver (C++ restatement of HELP HSDSolver.cpp) over a backend.

``solve_lp(inst, backend)`` runs the driver; ``backend`` is a ``help_backend``
table: the MI355X backend (``gpu_backend(mode)``) or any other C table with
the same four callbacks (construct/destruct/renew/solve).
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .backend import HelpError


class GpuBackend:
    """The MI355X backend as a help_backend callback table."""

    def __init__(self, mode=L.MODE_DIRECT_FP64, **opts):
        self.opts = L.default_opts(mode=mode, **opts)
        self.table = L.Backend()
        st = L.lib().help_gpu_backend_new(C.byref(self.opts), C.byref(self.table))
        if st != L.HELP_OK:
            raise HelpError(st, "help_gpu_backend_new failed")

    def set_warmstart(self, net):
        """HELP_MODE_PCG_LEARNED: start every PCG solve from ``net``'s guess."""
        self._net = net
        L.lib().help_gpu_backend_set_warmstart(C.byref(self.table), net.ptr if net is not None else None)

    def pcg_iterations(self):
        return L.lib().help_gpu_backend_pcg_iterations(C.byref(self.table))

    def solve_stats(self):
        """(solves, FP64 refinement iterations, largest final relative residual)"""
        n, it, rel = C.c_int64(), C.c_int64(), C.c_double()
        L.lib().help_gpu_backend_solve_stats(C.byref(self.table), C.byref(n), C.byref(it), C.byref(rel))
        return n.value, it.value, rel.value

    def close(self):
        if self.table.ctx:
            L.lib().help_gpu_backend_free(C.byref(self.table))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_lp(inst, table, max_iter=0, log=None, exact=False, device=False):
    """Run the HSD driver. ``table`` is an L.Backend. Returns (result, log_rows).

    exact=True sums every reduction in the reference's sequential order, so
    the iterates are bit-identical to HELP's own driver with the same backend.
    device=True runs the GPU-resident driver (table must be a GpuBackend's).
    """
    rows = []

    def _cb(user, it, pobj, dobj, gap, pinf, dinf):
        rows.append((it, pobj, dobj, gap, pinf, dinf))
        if log is not None:
            log(it, pobj, dobj, gap, pinf, dinf)

    cb = L.LOG_FN(_cb)
    res = L.HsdResult()
    s = inst.as_struct()
    if device:
        st = L.lib().help_hsd_solve_gpu(C.byref(s), C.byref(table), int(max_iter), cb, None,
                                        C.byref(res))
        if st != L.HELP_OK:
            raise HelpError(st, f"help_hsd_solve_gpu failed (backend error {res.backend_error})")
        return res, rows
    st = L.lib().help_hsd_solve(C.byref(s), C.byref(table), int(max_iter), 1 if exact else 0, cb,
                                None, C.byref(res))
    if st != L.HELP_OK:
        raise HelpError(st, f"help_hsd_solve failed (backend error {res.backend_error})")
    return res, rows


class RecordingBackend:
    """A help_backend table over a ``LinearEquation`` that hands every solve
    to ``on_solve(renew_count, k, d, b1, b2, x1, x2)`` (k = solves since the
    last renew; the arrays are views valid during the call).  HELP's driver
    calls SolveLinearEquation first for the bottom system of an iteration
    (HSDSolver.cpp:770), so k == 0 picks that system.  Used to harvest the
    (d, b1, b2) of a real interior-point trajectory.  The backend's statuses
    go back to the driver; the handle's last solve statistics are in
    ``le.solve_stats`` during the call."""

    def __init__(self, inst, on_solve, **opts):
        from .backend import LinearEquation
        self.inst = inst
        self.le = LinearEquation(**opts)
        self.renews = 0
        self._k = 0
        n, m = inst.n, inst.m

        def construct(ctx, lp):
            self.le.construct(self.inst)
            return 0

        def destruct(ctx):
            self.le.destruct()
            return 0

        def renew(ctx, dp):
            # the library's status goes back to the driver (no exception
            # crosses the C callback)
            st = L.lib().help_gpu_renew(self.le.h, dp, C.byref(self.le.renew_stats))
            self.renews += 1
            self._k = 0
            return st

        def solve(ctx, dp, b1p, b2p, x1p, x2p):
            st = L.lib().help_gpu_solve(self.le.h, dp, b1p, b2p, x1p, x2p, C.byref(self.le.solve_stats))
            d = np.ctypeslib.as_array(dp, shape=(n,))
            b1 = np.ctypeslib.as_array(b1p, shape=(n,))
            b2 = np.ctypeslib.as_array(b2p, shape=(m,))
            x1 = np.ctypeslib.as_array(x1p, shape=(n,))
            x2 = np.ctypeslib.as_array(x2p, shape=(m,))
            on_solve(self.renews, self._k, d, b1, b2, x1, x2)
            self._k += 1
            return st

        self._fns = (L.CONSTRUCT_FN(construct), L.DESTRUCT_FN(destruct), L.RENEW_FN(renew),
                     L.SOLVE_FN(solve))
        self.table = L.Backend()
        self.table.ctx = None
        for name, f in zip(("construct", "destruct", "renew", "solve"), self._fns):
            setattr(self.table, name, C.cast(f, C.c_void_p))
