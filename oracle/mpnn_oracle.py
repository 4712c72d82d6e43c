"""CPU oracle of the warm-start network (TEST INFRASTRUCTURE ONLY).

The same function as the backend's HIP kernels (help_amd/csrc/mpnn.h states
the math) written densely in torch float64 on the CPU; gradients come from
autograd, so they check the hand-written backward kernels independently.
Only tests import this module.  There is no reference counterpart: HELP
starts PCG from zero (PCG.cpp:309-310).  Pinned by tests/test_mpnn_oracle.py
(closed forms of the zero network and of both aggregations, permutation
equivariance, scale invariances, finite-difference gradients).
"""
import numpy as np
import torch

H, FR, FC = 64, 4, 2


def layout():
    off = {}
    o = 0
    for name, size in [("Wr0", FR * H), ("br0", H), ("Wc0", FC * H), ("bc0", H)]:
        off[name] = (o, size)
        o += size
    for l in range(1, 5):
        off[f"W{l}"] = (o, 2 * H * H)
        o += 2 * H * H
        off[f"b{l}"] = (o, H)
        o += H
    for name, size in [("Wh", H * H), ("bh", H), ("wo", H), ("bo", 1)]:
        off[name] = (o, size)
        o += size
    return off, o


def bf16_round(x):
    """Round float32 values to the nearest bfloat16 (ties to even)."""
    a = np.ascontiguousarray(x, dtype=np.float32).copy()
    u = a.view(np.uint32)
    finite = (u & 0x7F800000) != 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    u[finite] = r[finite]
    return a


def loss_and_grad(params, A, d, r, dy, bf16_weights=False):
    """Loss, gradient (float64 numpy) and the output guess x0 for a dense or
    scipy-sparse A (m x n), float64 d, r, dy."""
    off, total = layout()
    p_np = np.asarray(params, np.float32)
    if bf16_weights:
        p_np = bf16_round(p_np)
    p = torch.tensor(p_np.astype(np.float64), requires_grad=True)
    Ad = torch.tensor(np.asarray(A.todense() if hasattr(A, "todense") else A, np.float64))
    m, n = Ad.shape
    d = torch.tensor(np.asarray(d, np.float64))
    r = torch.tensor(np.asarray(r, np.float64))
    dy = torch.tensor(np.asarray(dy, np.float64))

    def P(name, shape):
        o, s = off[name]
        return p[o:o + s].reshape(shape)

    dinv = 1.0 / d
    g = (Ad * Ad) @ dinv
    gi = torch.where(g > 0, g, torch.ones_like(g))
    pos = g > 0
    mu = torch.log(g[pos]).mean() if pos.any() else torch.tensor(0.0, dtype=torch.float64)
    nu = torch.log(dinv).mean()
    rs = torch.sqrt((r * r).sum() / m)
    rs = rs if rs > 0 else torch.tensor(1.0, dtype=torch.float64)
    jac = torch.clamp(torch.exp(mu) * r / (rs * gi), -10.0, 10.0)
    xr = torch.stack([r / rs, (torch.log(gi) - mu) / 10.0, jac / 10.0, torch.ones_like(r)], 1)
    xc = torch.stack([(torch.log(dinv) - nu) / 10.0, torch.ones_like(d)], 1)
    # the kernels store the features in FP32
    xr = xr.float().double()
    xc = xc.float().double()
    deg_r = torch.clamp((Ad != 0).sum(1).double(), min=1.0)[:, None]
    deg_c = torch.clamp((Ad != 0).sum(0).double(), min=1.0)[:, None]
    relu = torch.relu

    def mean_rows(hc):
        return (Ad @ hc) / deg_r

    def mean_cols(hr):
        return (Ad.T @ hr) / deg_c

    hr0 = relu(xr @ P("Wr0", (FR, H)) + P("br0", (H,)))
    hc0 = relu(xc @ P("Wc0", (FC, H)) + P("bc0", (H,)))
    hc1 = relu(torch.cat([hc0, mean_cols(hr0)], 1) @ P("W1", (2 * H, H)) + P("b1", (H,)))
    hr1 = relu(torch.cat([hr0, mean_rows(hc1)], 1) @ P("W2", (2 * H, H)) + P("b2", (H,)))
    hc2 = relu(torch.cat([hc1, mean_cols(hr1)], 1) @ P("W3", (2 * H, H)) + P("b3", (H,)))
    hr2 = relu(torch.cat([hr1, mean_rows(hc2)], 1) @ P("W4", (2 * H, H)) + P("b4", (H,)))
    z = relu(hr2 @ P("Wh", (H, H)) + P("bh", (H,)))
    y = z @ P("wo", (H,)) + P("bo", (1,))
    t = dy * g / rs
    T = (t * t).sum()
    loss = ((y - t) ** 2).sum() / (T if T > 0 else 1.0)
    loss.backward()
    x0 = torch.where(g > 0, y.detach() * rs / gi, torch.zeros_like(y))
    return float(loss.detach()), p.grad.numpy().copy(), x0.numpy()
