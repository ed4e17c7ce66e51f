"""CPU oracle of the learned edge prior (MLP 8 -> 64 -> 64 -> 1, tanh GELU).

TEST INFRASTRUCTURE ONLY: the checker of the HIP prior kernels
(gplsaco_amd/csrc/prior_kernels.hip); only tests/, __graft_entry__.smoke()
and bench.py's baseline leg may use it. The reference has no learned prior,
so this file and the oracle's edge features (oracle_prior_features in
faco_oracle.cpp) are the definition. Parity of the device network is to a
tolerance: the device sums bf16 products on matrix cores in FP32 in its own
order; this oracle applies the same bf16 rounding points and sums in FP64.

Parameter layout (4801 floats): W1[64][8], b1[64], W2[64][64], b2[64],
w3[64], b3. Loss: mean over edges of w_e * (softplus(s) - y s), w_e =
pos_weight for positives, 1 otherwise.

`rounding=False` runs the same network in plain FP64 (no bf16, no FP32):
its gradient is checked against finite differences (tests/test_oracle_prior.py).
"""
from __future__ import annotations

import numpy as np

P = 4801
O_W1, O_B1, O_W2, O_B2, O_W3, O_B3 = 0, 512, 576, 4672, 4736, 4800
C_GELU = 0.7978845608028654


def bf16(x):
    """Round FP32 values to bf16 (round to nearest even), returned as FP32."""
    x = np.asarray(x, dtype=np.float32)
    bits = x.view(np.uint32).astype(np.uint64)
    rounded = (bits + 0x7FFF + ((bits >> 16) & 1)) & 0xFFFF0000
    return rounded.astype(np.uint32).view(np.float32)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    return (p[O_W1:O_B1].reshape(64, 8), p[O_B1:O_W2], p[O_W2:O_B2].reshape(64, 64), p[O_B2:O_W3],
            p[O_W3:O_B3], p[O_B3])


def gelu(z):
    u = C_GELU * (z + 0.044715 * z * z * z)
    return 0.5 * z * (1.0 + np.tanh(u))


def gelu_grad(z):
    u = C_GELU * (z + 0.044715 * z * z * z)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 0.134145 * z * z)


def _r(x, rounding):
    """bf16 operand of a matrix-core product (FP64 value of the bf16)."""
    return bf16(np.asarray(x, dtype=np.float32)).astype(np.float64) if rounding else np.asarray(x, dtype=np.float64)


def _f32(x, rounding):
    return np.asarray(x, dtype=np.float32).astype(np.float64) if rounding else np.asarray(x, dtype=np.float64)


def forward(params, features, rounding=True):
    W1, b1, W2, b2, w3, b3 = unpack(params)
    if rounding:
        W1, W2, w3 = (bf16(np.asarray(W, dtype=np.float32)).astype(np.float64) for W in (W1, W2, w3))
        b1, b2, b3 = (np.asarray(b, dtype=np.float32).astype(np.float64) for b in (b1, b2, b3))
    x = _r(features, rounding)
    z1 = _f32(x @ W1.T + b1, rounding)
    h1 = _r(_f32(gelu(z1), rounding), rounding)
    z2 = _f32(h1 @ W2.T + b2, rounding)
    h2 = _r(_f32(gelu(z2), rounding), rounding)
    s = _f32(h2 @ w3 + b3, rounding)
    return {"x": x, "z1": z1, "h1": h1, "z2": z2, "h2": h2, "s": s, "W2": W2, "w3": w3}


def loss_and_grad(params, features, labels, pos_weight, rounding=True):
    """Loss and gradient [4801] of the weighted BCE, mean over edges."""
    f = forward(params, features, rounding)
    s = f["s"]
    y = np.asarray(labels, dtype=np.float64)
    w = np.where(y > 0, float(pos_weight), 1.0)
    E = s.size
    softplus = np.where(s > 0, s + np.log1p(np.exp(-np.abs(s))), np.log1p(np.exp(s)))
    loss = float(np.sum(w * (softplus - y * s)) / E)
    ds = _f32(w * (1.0 / (1.0 + np.exp(-s)) - y) / E, rounding)
    g = np.zeros(P)
    g[O_B3] = ds.sum()
    g[O_W3:O_B3] = f["h2"].T @ ds
    dz2 = _f32(ds[:, None] * f["w3"][None, :] * gelu_grad(f["z2"]), rounding)
    g[O_B2:O_W3] = dz2.sum(axis=0)
    dz2r = _r(dz2, rounding)
    g[O_W2:O_B2] = (dz2r.T @ f["h1"]).reshape(-1)
    dz1 = _f32((dz2r @ f["W2"]) * gelu_grad(f["z1"]), rounding)
    g[O_B1:O_W2] = dz1.sum(axis=0)
    g[O_W1:O_B1] = (_r(dz1, rounding).T @ f["x"]).reshape(-1)
    return loss, g


def labels_from_tour(primary, tour):
    """1 for candidate edges (a, primary[a][j]) that are edges of `tour`."""
    primary = np.asarray(primary)
    n = primary.shape[0]
    tour = np.asarray(tour)
    succ = np.empty(n, dtype=np.int64)
    pred = np.empty(n, dtype=np.int64)
    succ[tour] = np.roll(tour, -1)
    pred[tour] = np.roll(tour, 1)
    return ((primary == succ[:, None]) | (primary == pred[:, None])).astype(np.uint8).reshape(-1)
