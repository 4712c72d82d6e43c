"""Learned PCG warm start: the message-passing network of the backend
(help_mpnn_*), data harvested from the plain solver, training and evaluation.

* ``Mpnn``: the network (BF16 or FP32 activations) with predict / loss_grad
  / adam_step running on the GPU through the backend's own kernels.
* ``harvest(inst)``: runs HELP's HSD driver with the plain FP64 direct
  backend and keeps (d, r, dy) tuples of the normal-equation solves, with
  r = A D^-1 b1 - b2 and dy = x2.
* ``train`` / ``pcg_iterations``: fit on harvested tuples; count PCG
  iterations from zero and from the network's guess to the same FP64 target.
"""
import ctypes as C
import time

import numpy as np

from . import _lib as L
from .backend import HelpError, LinearEquation


class Mpnn:
    """The warm-start network of the backend (parameters live on the GPU)."""

    def __init__(self, bf16=True, seed=0):
        p = C.c_void_p()
        st = L.lib().help_mpnn_new(1 if bf16 else 0, seed, C.byref(p))
        if st != L.HELP_OK:
            raise HelpError(st, "help_mpnn_new failed")
        self.ptr = p.value
        self.bf16 = bf16
        self.nparams = L.lib().help_mpnn_num_params()

    def params(self):
        out = np.empty(self.nparams, np.float32)
        L.lib().help_mpnn_get_params(self.ptr, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        st = L.lib().help_mpnn_set_params(self.ptr, p.ctypes.data_as(C.POINTER(C.c_float)))
        if st != L.HELP_OK:
            raise HelpError(st, "help_mpnn_set_params failed")

    def predict(self, be, d, r):
        x0 = np.empty(be.m)
        d = np.ascontiguousarray(d, np.float64)
        r = np.ascontiguousarray(r, np.float64)
        st = L.lib().help_mpnn_predict(self.ptr, be.h, L.dptr(d), L.dptr(r), L.dptr(x0))
        if st != L.HELP_OK:
            raise HelpError(st, "help_mpnn_predict failed")
        return x0

    def loss_grad(self, be, d, r, dy, want_grad=False):
        loss = C.c_double()
        g = np.empty(self.nparams, np.float32) if want_grad else None
        d = np.ascontiguousarray(d, np.float64)
        r = np.ascontiguousarray(r, np.float64)
        dy = np.ascontiguousarray(dy, np.float64)
        st = L.lib().help_mpnn_loss_grad(self.ptr, be.h, L.dptr(d), L.dptr(r), L.dptr(dy), C.byref(loss),
                                         g.ctypes.data_as(C.POINTER(C.c_float)) if want_grad else None)
        if st != L.HELP_OK:
            raise HelpError(st, "help_mpnn_loss_grad failed")
        return (loss.value, g) if want_grad else loss.value

    def adam_step(self, lr):
        st = L.lib().help_mpnn_adam_step(self.ptr, float(lr))
        if st != L.HELP_OK:
            raise HelpError(st, "help_mpnn_adam_step failed")

    def save(self, path):
        np.savez(path, params=self.params(), bf16=np.int32(self.bf16))

    @classmethod
    def load(cls, path):
        z = np.load(path)
        net = cls(bf16=bool(z["bf16"]))
        net.set_params(z["params"])
        return net

    def close(self):
        if self.ptr:
            L.lib().help_mpnn_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_CONSTRUCT = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)
_DESTRUCT = C.CFUNCTYPE(C.c_int, C.c_void_p)
_RENEW = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double))
_SOLVE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                     C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double))


def harvest(inst, every=3, max_tuples=64, max_iter=0):
    """(d, r, dy) tuples from every ``every``-th normal-equation solve of an
    HSD run with the plain FP64 direct backend (HELP's driver on the host)."""
    from .hsd import solve_lp
    A = inst.to_scipy().tocsr()
    be = LinearEquation(mode=L.MODE_DIRECT_FP64)
    tuples = []
    count = [0]
    m, n = inst.m, inst.n

    def construct(ctx, lp):
        be.construct(inst)
        return 0

    def destruct(ctx):
        be.destruct()
        return 0

    def renew(ctx, d):
        be.renew(np.ctypeslib.as_array(d, shape=(n,)))
        return 0

    def solve(ctx, d, b1, b2, x1, x2):
        D = np.ctypeslib.as_array(d, shape=(n,)).copy()
        B1 = np.ctypeslib.as_array(b1, shape=(n,)).copy()
        B2 = np.ctypeslib.as_array(b2, shape=(m,)).copy()
        X1, X2 = be.solve(D, B1, B2)
        np.ctypeslib.as_array(x1, shape=(n,))[:] = X1
        np.ctypeslib.as_array(x2, shape=(m,))[:] = X2
        if count[0] % every == 0 and len(tuples) < max_tuples:
            tuples.append((D, A @ (B1 / D) - B2, X2.copy()))
        count[0] += 1
        return 0

    fns = (_CONSTRUCT(construct), _DESTRUCT(destruct), _RENEW(renew), _SOLVE(solve))
    table = L.Backend()
    table.ctx = None
    table.construct, table.destruct, table.renew, table.solve = [C.cast(f, C.c_void_p) for f in fns]
    res, _ = solve_lp(inst, table, max_iter=max_iter)
    return tuples, res


def graph_handle(inst, pcg_k=0, mode=L.MODE_PCG, **opts):
    """A PCG-mode handle on the instance's matrix (for training and counting)."""
    be = LinearEquation(mode=mode, pcg_k=pcg_k, **opts)
    be.construct(inst)
    return be


def train(net, data, epochs=10, lr=2e-3, log=None, seed=0):
    """data: list of (handle, (d, r, dy)).  One Adam step per tuple."""
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(data))
        tot = 0.0
        t0 = time.time()
        for i in order:
            be, (d, r, dy) = data[i]
            tot += net.loss_grad(be, d, r, dy)
            net.adam_step(lr)
        hist.append(tot / max(1, len(data)))
        if log:
            log(f"epoch {ep}: mean loss {hist[-1]:.4f} ({time.time() - t0:.1f}s)")
    return hist


def pcg_iterations(be_plain, be_learned, d, r):
    """PCG iterations for A D^-1 A^T x = r from zero (be_plain, MODE_PCG) and
    from the network's guess (be_learned, MODE_PCG_LEARNED), same target."""
    n = be_plain.n
    b1 = np.zeros(n)
    b2 = -np.asarray(r, np.float64)
    be_plain.renew(d)
    t = time.time()
    _, x_zero = be_plain.solve(d, b1, b2)
    t_zero = time.time() - t
    it_zero = be_plain.solve_stats.pcg_iters
    be_learned.renew(d)
    t = time.time()
    _, x_net = be_learned.solve(d, b1, b2)
    t_net = time.time() - t
    it_net = be_learned.solve_stats.pcg_iters
    return dict(it_zero=it_zero, it_net=it_net, t_zero=t_zero, t_net=t_net,
                fallback=be_learned.solve_stats.fallback,
                agree=float(np.linalg.norm(x_net - x_zero) / max(np.linalg.norm(x_zero), 1e-300)))
