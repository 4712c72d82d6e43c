"""GPU checks of the learned PCG warm start: the hand-written forward and
backward kernels against the CPU oracle (torch float64 autograd), training
reduces the loss, and the learned PCG mode returns the same answer as PCG
from zero (the guess can only change the iteration count)."""
import os
import sys

import numpy as np
import pytest

from help_amd import _lib as L
from help_amd import instances
from help_amd.warmstart import Mpnn, graph_handle, harvest, pcg_iterations, train

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def _flush(msg):
    print(msg, flush=True)


def _case(seed=0, side=6):
    inst = instances.mcf(side=side, commodities=2, pairs=3, seed=31 + seed)
    rng = np.random.default_rng(seed)
    d = np.exp(rng.uniform(-2, 2, inst.n) * np.log(10.0))
    r = rng.standard_normal(inst.m)
    dy = rng.standard_normal(inst.m) * 1e-2
    return inst, d, r, dy


def test_mfma_bf16_lane_layout():
    """In-bounds probe of V_MFMA_F32_16X16X32_BF16, the BF16 dense layers'
    instruction: A[i][k] in lane i+16(k/8) element k%8, B[k][j] in lane
    j+16(k/8) element k%8, D[i][j] in lane j+16(i/4) element i%4 (small
    integers: every product and sum is exact)."""
    rng = np.random.default_rng(12)
    A = rng.integers(-4, 5, (16, 32)).astype(np.float64)
    B = rng.integers(-4, 5, (32, 16)).astype(np.float64)
    a = np.array([A[l % 16, 8 * (l // 16) + e] for l in range(64) for e in range(8)])
    b = np.array([B[8 * (l // 16) + e, l % 16] for l in range(64) for e in range(8)])
    out = np.zeros(256)
    assert L.lib().help_selftest_mfma_bf16(L.dptr(a), L.dptr(b), L.dptr(out)) == 0
    D = A @ B
    exp = np.array([D[4 * (l // 16) + r, l % 16] for l in range(64) for r in range(4)])
    _flush(f"mfma bf16 16x16x32 max|diff|={np.abs(out - exp).max()}")
    assert np.array_equal(out, exp)


def test_gradients_fp32_match_autograd():
    import mpnn_oracle as MO
    inst, d, r, dy = _case()
    _flush(f"mpnn fp32 gradient check m={inst.m} n={inst.n}")
    be = graph_handle(inst)
    net = Mpnn(bf16=False, seed=3)
    p = net.params()
    loss, g = net.loss_grad(be, d, r, dy, want_grad=True)
    x0 = net.predict(be, d, r)
    lo, go, xo = MO.loss_and_grad(p, inst.to_scipy(), d, r, dy)
    rel_g = np.linalg.norm(g - go) / np.linalg.norm(go)
    _flush(f"  loss {loss:.8f} vs {lo:.8f}; grad rel err {rel_g:.2e}")
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    assert rel_g <= 1e-4
    assert np.linalg.norm(x0 - xo) <= 1e-5 * np.linalg.norm(xo)
    net.close()
    be.destruct()


def test_gradients_bf16_close_to_autograd():
    import mpnn_oracle as MO
    inst, d, r, dy = _case(seed=1)
    be = graph_handle(inst)
    net = Mpnn(bf16=True, seed=4)
    p = net.params()
    loss, g = net.loss_grad(be, d, r, dy, want_grad=True)
    lo, go, _ = MO.loss_and_grad(p, inst.to_scipy(), d, r, dy, bf16_weights=True)
    cos = float(g @ go / (np.linalg.norm(g) * np.linalg.norm(go)))
    rel_g = np.linalg.norm(g - go) / np.linalg.norm(go)
    _flush(f"mpnn bf16: loss {loss:.6f} vs {lo:.6f}; grad cos {cos:.5f} rel {rel_g:.2e}")
    assert abs(loss - lo) <= 2e-2 * abs(lo)
    assert cos >= 0.995 and rel_g <= 0.1
    net.close()
    be.destruct()


def test_training_reduces_loss():
    inst = instances.mcf(side=8, commodities=2, pairs=4, seed=40)
    tuples, res = harvest(inst, every=3, max_tuples=12)
    _flush(f"harvested {len(tuples)} tuples (hsd status {res.status}, {res.iterations} iterations)")
    assert len(tuples) >= 6
    be = graph_handle(inst)
    net = Mpnn(bf16=True, seed=5)
    data = [(be, t) for t in tuples]
    hist = train(net, data, epochs=40, lr=3e-3, log=_flush)
    assert min(hist[-5:]) < 0.8 * hist[0]
    net.close()
    be.destruct()


def test_learned_pcg_same_answer_and_fallback():
    inst = instances.mcf(side=10, commodities=2, pairs=5, seed=41)
    tuples, _ = harvest(inst, every=3, max_tuples=10)
    g = graph_handle(inst)
    net = Mpnn(bf16=True, seed=6)
    train(net, [(g, t) for t in tuples], epochs=10, lr=3e-3)
    g.destruct()
    plain = graph_handle(inst, pcg_k=20, pcg_tol=1e-10)
    learned = graph_handle(inst, pcg_k=20, pcg_tol=1e-10, mode=L.MODE_PCG_LEARNED)
    learned.set_warmstart(net)
    d, r, _ = tuples[len(tuples) // 2]
    out = pcg_iterations(plain, learned, d, r)
    _flush(f"pcg iterations zero {out['it_zero']} net {out['it_net']} agree {out['agree']:.2e}")
    assert out["fallback"] == 0
    assert out["agree"] <= 1e-7
    # a non-finite guess (forced from the host) starts from zero instead
    forced = graph_handle(inst, pcg_k=20, pcg_tol=1e-10, mode=L.MODE_PCG_LEARNED,
                          fault_flags=L.FAULT_NET_NONFINITE)
    forced.set_warmstart(net)
    out2 = pcg_iterations(plain, forced, d, r)
    assert out2["fallback"] == 1 and out2["it_net"] == out2["it_zero"]
    for b in (plain, learned, forced):
        b.destruct()
    net.close()


def test_committed_network_on_heldout_instance():
    """The committed network (help_amd/data) in learned-PCG mode on a held-out
    instance of its family at side >= 60 (side 76, K=2, seed 2002; training
    used seeds 1000 and up): same answer as PCG from zero, no fallback."""
    path = os.path.join(ROOT, "help_amd", "data", "warmstart_net_mcf.npz")
    inst = instances.by_config(4)
    _flush(f"held-out mcf side 76 K=2 m={inst.m} n={inst.n}")
    tuples, res = harvest(inst, every=30, max_tuples=2, max_iter=40)
    assert len(tuples) == 2
    net = Mpnn.load(path)
    plain = graph_handle(inst, pcg_k=50, pcg_tol=1e-10)
    learned = graph_handle(inst, pcg_k=50, pcg_tol=1e-10, mode=L.MODE_PCG_LEARNED)
    learned.set_warmstart(net)
    for d, r, _ in tuples:
        out = pcg_iterations(plain, learned, d, r)
        _flush(f"  pcg iterations zero {out['it_zero']} net {out['it_net']} agree {out['agree']:.2e}")
        assert out["fallback"] == 0
        assert out["agree"] <= 1e-7
    for b in (plain, learned):
        b.destruct()
    net.close()
