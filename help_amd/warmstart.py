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


def harvest(inst, every=3, max_tuples=64, max_iter=0, log=None):
    """(d, r, dy) tuples from every ``every``-th normal-equation solve of an
    HSD run with the plain FP64 direct backend (HELP's driver on the host,
    the backend called through the C-ABI), r = A D^-1 b1 - b2, dy = x2.
    ``log`` gets a progress line every 25 interior-point iterations."""
    from .hsd import RecordingBackend, solve_lp
    A = inst.to_scipy().tocsr()
    tuples = []
    count = [0]
    t0 = time.time()

    def on_solve(renews, k, d, b1, b2, x1, x2):
        if count[0] % every == 0 and len(tuples) < max_tuples:
            D = d.copy()
            tuples.append((D, A @ (b1 / D) - b2, x2.copy()))
        count[0] += 1

    def progress(it, *_):
        if log is not None and it % 25 == 0:
            log(f"  harvest m={inst.m}: iteration {it}, {count[0]} solves, {time.time() - t0:.1f}s")

    rb = RecordingBackend(inst, on_solve, mode=L.MODE_DIRECT_FP64)
    res, _ = solve_lp(inst, rb.table, max_iter=max_iter, log=progress)
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


def _pcg_solve(be, d, b1, b2):
    """One PCG-mode solve; a solve that reaches the iteration cap
    (HELP_ERR_NOCONV) still returns its iterate, flagged."""
    x1 = np.empty(be.n)
    x2 = np.empty(be.m)
    t = time.time()
    st = L.lib().help_gpu_solve(be.h, L.dptr(d), L.dptr(b1), L.dptr(b2), L.dptr(x1), L.dptr(x2),
                                C.byref(be.solve_stats))
    elapsed = time.time() - t
    if st not in (L.HELP_OK, L.HELP_ERR_NOCONV):
        be._check(st, "solve")
    return x2, elapsed, st == L.HELP_ERR_NOCONV


def pcg_iterations(be_plain, be_learned, d, r):
    """PCG iterations for A D^-1 A^T x = r from zero (be_plain, MODE_PCG) and
    from the network's guess (be_learned, MODE_PCG_LEARNED), same target; a
    solve at the iteration cap counts the cap and is flagged."""
    n = be_plain.n
    d = np.ascontiguousarray(d, np.float64)
    b1 = np.zeros(n)
    b2 = np.ascontiguousarray(-np.asarray(r, np.float64))
    be_plain.renew(d)
    x_zero, t_zero, cap_zero = _pcg_solve(be_plain, d, b1, b2)
    it_zero = be_plain.solve_stats.pcg_iters
    be_learned.renew(d)
    x_net, t_net, cap_net = _pcg_solve(be_learned, d, b1, b2)
    it_net = be_learned.solve_stats.pcg_iters
    return dict(it_zero=it_zero, it_net=it_net, t_zero=t_zero, t_net=t_net, capped_zero=cap_zero,
                capped_net=cap_net, fallback=be_learned.solve_stats.fallback,
                agree=float(np.linalg.norm(x_net - x_zero) / max(np.linalg.norm(x_zero), 1e-300)))


def _instances(count, seed0, side, ks):
    from . import instances
    rng = np.random.default_rng(seed0)
    out = []
    for i in range(count):
        s = int(rng.integers(side[0], side[1] + 1))
        k = int(rng.integers(ks[0], ks[1] + 1))
        out.append(((s, k, seed0 + i), instances.mcf(side=s, commodities=k, pairs=20, seed=seed0 + i)))
    return out


def main(argv=None):
    """Harvest, train and evaluate on the multicommodity-flow family:

        python -m help_amd.warmstart --train 24 --test 8 --out DIR

    writes DIR/warmstart_net.npz and DIR/warmstart_eval.json."""
    import argparse
    import json
    import os
    from .hsd import GpuBackend, solve_lp
    ap = argparse.ArgumentParser(prog="python -m help_amd.warmstart")
    ap.add_argument("--train", type=int, default=24)
    ap.add_argument("--test", type=int, default=8)
    ap.add_argument("--side", type=int, nargs=2, default=[20, 40])
    ap.add_argument("--k", type=int, nargs=2, default=[2, 4])
    ap.add_argument("--every", type=int, default=3)
    ap.add_argument("--tuples", type=int, default=12)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--pcg-k", type=int, default=50)
    ap.add_argument("--eval-per-instance", type=int, default=3)
    ap.add_argument("--eval-cap", type=int, default=20000,
                    help="PCG iteration cap of the held-out comparisons")
    ap.add_argument("--harvest-iter", type=int, default=0,
                    help="interior-point iteration cap of each harvesting run (0: the driver's own)")
    ap.add_argument("--e2e", type=int, default=1,
                    help="1: HSD end to end in PCG and learned-PCG mode on the smallest held-out instance")
    ap.add_argument("--e2e-iter", type=int, default=0,
                    help="interior-point iteration cap of the end-to-end runs (0: the driver's own)")
    ap.add_argument("--out", default=".")
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    log = lambda s: print(s, flush=True)
    t0 = time.time()
    train_inst = _instances(args.train, 1000, args.side, args.k)
    test_inst = _instances(args.test, 2000, args.side, args.k)
    data, handles = [], []
    for key, inst in train_inst:
        t_i = time.time()
        tuples, res = harvest(inst, every=args.every, max_tuples=args.tuples, max_iter=args.harvest_iter,
                              log=log)
        be = graph_handle(inst)
        handles.append(be)
        data += [(be, t) for t in tuples]
        log(f"train {key}: m={inst.m} n={inst.n} hsd status {res.status} iters {res.iterations} "
            f"tuples {len(tuples)} ({time.time() - t_i:.1f}s)")
    t_harvest = time.time() - t0
    log(f"{len(data)} training tuples ({t_harvest:.1f}s)")
    net = Mpnn(bf16=True, seed=0)
    t1 = time.time()
    hist = train(net, data, epochs=args.epochs, lr=args.lr, log=log)
    t_train = time.time() - t1
    net.save(os.path.join(args.out, "warmstart_net.npz"))
    for be in handles:
        be.destruct()
    rows = []
    for key, inst in test_inst:
        tuples, res = harvest(inst, every=args.every, max_tuples=args.tuples, max_iter=args.harvest_iter,
                              log=log)
        g = graph_handle(inst)
        losses = [net.loss_grad(g, *t) for t in tuples]
        g.destruct()
        plain = graph_handle(inst, pcg_k=args.pcg_k, pcg_tol=1e-10, pcg_max_iter=args.eval_cap)
        learned = graph_handle(inst, pcg_k=args.pcg_k, pcg_tol=1e-10, pcg_max_iter=args.eval_cap,
                               mode=L.MODE_PCG_LEARNED)
        learned.set_warmstart(net)
        pick = np.linspace(0, len(tuples) - 1, min(args.eval_per_instance, len(tuples))).astype(int)
        for i in pick:
            out = pcg_iterations(plain, learned, tuples[i][0], tuples[i][1])
            out.update(instance=key, tuple=int(i), loss=float(losses[i]), m=inst.m, n=inst.n)
            rows.append(out)
            log(f"test {key} tuple {i}: loss {losses[i]:.3f} pcg iterations zero {out['it_zero']} "
                f"net {out['it_net']} ({out['t_zero']:.3f}s / {out['t_net']:.3f}s) agree {out['agree']:.1e}"
                + (" (cap reached)" if out["capped_zero"] or out["capped_net"] else ""))
        plain.destruct()
        learned.destruct()
    it0 = np.array([r["it_zero"] for r in rows], float)
    it1 = np.array([r["it_net"] for r in rows], float)
    tz = np.array([r["t_zero"] for r in rows], float)
    tn = np.array([r["t_net"] for r in rows], float)
    # end to end on the smallest held-out instance: PCG vs learned PCG
    key, inst = min(test_inst, key=lambda ki: ki[1].m)
    e2e = {"instance": key}
    for name, mode in (("pcg", L.MODE_PCG), ("pcg_learned", L.MODE_PCG_LEARNED)):
        if not args.e2e:
            break
        gb = GpuBackend(mode=mode, pcg_k=args.pcg_k)
        if mode == L.MODE_PCG_LEARNED:
            gb.set_warmstart(net)
        t = time.time()
        res, _ = solve_lp(inst, gb.table, device=True, max_iter=args.e2e_iter,
                          log=lambda it, *_: log(f"  {name} iteration {it} ({time.time() - t:.1f}s)"))
        e2e[name] = dict(status=L.HSD_STATUS.get(res.status, str(res.status)), iterations=res.iterations,
                         rel_gap=res.rel_gap, primal_obj=res.primal_obj, seconds=time.time() - t,
                         pcg_iterations=int(gb.pcg_iterations()))
        gb.close()
        log(f"end to end {name}: {e2e[name]}")
    summary = dict(
        family="multicommodity flow on a 4-connected grid (instances.mcf)", side=args.side, commodities=args.k,
        n_train=args.train, n_test=args.test, training_tuples=len(data), epochs=args.epochs,
        params=net.nparams, precision="bf16 activations and weights, fp32 accumulation",
        loss_history=hist, harvest_seconds=t_harvest, train_seconds=t_train, pcg_k=args.pcg_k, pcg_tol=1e-10,
        eval_pcg_cap=args.eval_cap,
        capped_solves=int(sum(1 for r in rows if r["capped_zero"] or r["capped_net"])),
        pcg_iterations_zero_mean=float(it0.mean()), pcg_iterations_net_mean=float(it1.mean()),
        iteration_ratio_mean=float((it1 / np.maximum(it0, 1)).mean()),
        time_zero_mean=float(tz.mean()), time_net_mean=float(tn.mean()),
        end_to_end=e2e, per_solve=rows)
    with open(os.path.join(args.out, "warmstart_eval.json"), "w") as f:
        json.dump(summary, f, indent=1)
    log(json.dumps({k: v for k, v in summary.items() if k not in ("per_solve", "loss_history")}))
    net.close()
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
