"""CPU checks that pin the warm-start network oracle (oracle/mpnn_oracle.py)
by properties the mathematics fixes, independently of its code:

* all weights zero except the output bias: y = bo everywhere, so the loss is
  sum (bo - t)^2 / sum t^2 and d loss / d bo = 2 sum (bo - t) / sum t^2, with
  t = dy g / rms(r) and g = diag(A D^-1 A^T);
* permuting the rows (or the columns) of A with r, dy (or d) permutes the
  guess and leaves the loss unchanged (mean aggregation is order-free);
* scaling r and dy by a > 0 scales the guess by a and leaves the loss
  unchanged; scaling d by a scales g by 1/a and the guess by a;
* the autograd gradient equals central finite differences.
A wrong axis in an aggregation, a dropped normalisation or a wrong target
definition fails one of these.
"""
import os
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mpnn_oracle as MO  # noqa: E402


def _graph(seed=0, m=7, n=11):
    rng = np.random.default_rng(seed)
    A = np.zeros((m, n))
    for j in range(n):
        rows = rng.choice(m, size=2, replace=False)
        A[rows, j] = rng.standard_normal(2)
    for i in range(m):          # every row has an entry
        if not A[i].any():
            A[i, rng.integers(n)] = 1.0
    d = np.exp(rng.uniform(-1, 1, n))
    r = rng.standard_normal(m)
    dy = rng.standard_normal(m)
    return A, d, r, dy


def _params(seed=1, scale=0.3):
    _, total = MO.layout()
    rng = np.random.default_rng(seed)
    return (scale * rng.standard_normal(total)).astype(np.float32)


def test_layout_size():
    _, total = MO.layout()
    assert total == 37761


def test_zero_network_closed_form():
    A, d, r, dy = _graph()
    off, total = MO.layout()
    p = np.zeros(total, np.float32)
    bo = 0.625
    p[off["bo"][0]] = bo
    loss, grad, x0 = MO.loss_and_grad(p, A, d, r, dy)
    g = (A * A) @ (1.0 / d)
    rs = np.sqrt((r * r).sum() / len(r))
    t = dy * g / rs
    T = (t * t).sum()
    assert loss == pytest.approx(((bo - t) ** 2).sum() / T, rel=1e-12)
    assert grad[off["bo"][0]] == pytest.approx(2 * (bo - t).sum() / T, rel=1e-12)
    assert np.allclose(x0, bo * rs / g, rtol=1e-12)


def test_row_and_column_permutation_equivariance():
    A, d, r, dy = _graph(2)
    p = _params()
    loss, _, x0 = MO.loss_and_grad(p, A, d, r, dy)
    rng = np.random.default_rng(3)
    pr = rng.permutation(A.shape[0])
    pc = rng.permutation(A.shape[1])
    loss2, _, x2 = MO.loss_and_grad(p, A[pr][:, pc], d[pc], r[pr], dy[pr])
    assert loss2 == pytest.approx(loss, rel=1e-10)
    assert np.allclose(x2, x0[pr], rtol=1e-9, atol=1e-12)


def test_scale_invariance():
    A, d, r, dy = _graph(4)
    p = _params(5)
    loss, _, x0 = MO.loss_and_grad(p, A, d, r, dy)
    # the features are stored in FP32 by the kernels, so allow FP32-level drift
    loss_r, _, x_r = MO.loss_and_grad(p, A, d, 4.0 * r, 4.0 * dy)
    assert loss_r == pytest.approx(loss, rel=1e-5)
    assert np.allclose(x_r, 4.0 * x0, rtol=1e-5)
    loss_d, _, x_d = MO.loss_and_grad(p, A, 8.0 * d, r, 8.0 * dy)
    assert loss_d == pytest.approx(loss, rel=1e-5)
    assert np.allclose(x_d, 8.0 * x0, rtol=1e-5)


def test_gradient_matches_finite_differences():
    A, d, r, dy = _graph(6)
    p = _params(7, scale=0.2).astype(np.float64)
    _, grad, _ = MO.loss_and_grad(p, A, d, r, dy)
    off, _ = MO.layout()
    rng = np.random.default_rng(8)
    picks = [off["bo"][0], off["wo"][0] + 3, off["W4"][0] + 77, off["W1"][0] + 200, off["Wr0"][0] + 5]
    picks += list(rng.integers(0, len(p), 5))
    for k in picks:
        h = 1e-3 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        lp, _, _ = MO.loss_and_grad(pp, A, d, r, dy)
        lm, _, _ = MO.loss_and_grad(pm, A, d, r, dy)
        fd = (lp - lm) / (2 * h)
        assert grad[k] == pytest.approx(fd, rel=2e-3, abs=1e-6), k


def _path_params(route):
    """Weights that carry one unit through the network.  route 'row': a
    constant column state through mean() in layer 2; 'col': a constant row
    state through meanT() in layer 1 and then mean() in layer 2."""
    off, total = MO.layout()
    p = np.zeros(total, np.float32)
    H = MO.H

    def put(name, idx, v):
        p[off[name][0] + idx] = v

    if route == "row":
        put("bc0", 0, 1.0)                 # hc0[:, 0] = 1
        put("W1", 0 * H + 0, 1.0)          # hc1[:, 0] = hc0[:, 0]
        put("W2", (H + 0) * H + 0, 1.0)    # hr1[:, 0] = mean(hc1)[:, 0]
    else:
        put("br0", 0, 1.0)                 # hr0[:, 0] = 1
        put("W1", (H + 0) * H + 0, 1.0)    # hc1[:, 0] = meanT(hr0)[:, 0]
        put("W2", (H + 0) * H + 0, 1.0)    # hr1[:, 0] = mean(hc1)[:, 0]
    put("W4", 0 * H + 0, 1.0)              # hr2[:, 0] = hr1[:, 0]
    put("Wh", 0 * H + 0, 1.0)              # z[:, 0] = hr2[:, 0]
    put("wo", 0, 1.0)                      # y = z[:, 0]
    return p


@pytest.mark.parametrize("route", ["row", "col"])
def test_aggregations_closed_form(route):
    A, d, r, dy = _graph(9)
    A = np.abs(A)                          # positive: the ReLUs pass everything
    p = _path_params(route)
    _, _, x0 = MO.loss_and_grad(p, A, d, r, dy)
    g = (A * A) @ (1.0 / d)
    rs = np.sqrt((r * r).sum() / len(r))
    y = x0 * g / rs
    degr = (A != 0).sum(1)
    degc = (A != 0).sum(0)
    if route == "row":
        want = A.sum(1) / degr
    else:
        want = (A @ (A.sum(0) / degc)) / degr
    assert np.allclose(y, want, rtol=1e-12)
