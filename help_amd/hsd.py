"""HSD interior-point driver (C++ restatement of HELP HSDSolver.cpp) over a backend.

``solve_lp(inst, backend)`` runs the driver; ``backend`` is a ``help_backend``
table: the MI355X backend (``gpu_backend(mode)``) or any other C table with
the same four callbacks (construct/destruct/renew/solve).
"""
import ctypes as C

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
