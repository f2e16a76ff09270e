"""Float64 CPU restatement of the strategy/capacity MLP (test infrastructure only).

Network: 16 -> 64 (ReLU) -> 64 (ReLU) -> 4, parameters flattened as W1[16x64], b1[64], W2[64x64], b2[64],
W3[64x4], b3[4] (row-major, x @ W).  Output 0 regresses log2 of the group count; outputs 1..3 are the logits
of the strategies (LDS-local, global-atomic, radix-partitioned).
Loss = mean (y0 - t)^2 + mean cross-entropy(softmax(y1..3), c).
The analytic gradient here is itself pinned by central finite differences in tests/test_mlp.py.
"""
from __future__ import annotations

import numpy as np

SIZES = [(16, 64), (64, 64), (64, 4)]
N_PARAMS = sum(i * o + o for i, o in SIZES)  # 5508


def unpack(p: np.ndarray):
    out, off = [], 0
    for i, o in SIZES:
        W = p[off:off + i * o].reshape(i, o)
        off += i * o
        b = p[off:off + o]
        off += o
        out.append((W, b))
    return out


def forward(p: np.ndarray, X: np.ndarray):
    (W1, b1), (W2, b2), (W3, b3) = unpack(p.astype(np.float64))
    X = X.astype(np.float64)
    a1 = X @ W1 + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0)
    y = h2 @ W3 + b3
    return y, (X, a1, h1, a2, h2)


def loss(p, X, t, c):
    y, _ = forward(p, X)
    B = X.shape[0]
    mse = np.mean((y[:, 0] - t) ** 2)
    z = y[:, 1:4]
    z = z - z.max(axis=1, keepdims=True)
    lse = np.log(np.exp(z).sum(axis=1))
    ce = np.mean(lse - z[np.arange(B), c])
    return mse + ce


def grad(p, X, t, c):
    (W1, b1), (W2, b2), (W3, b3) = unpack(p.astype(np.float64))
    y, (X, a1, h1, a2, h2) = forward(p, X)
    B = X.shape[0]
    dy = np.zeros_like(y)
    dy[:, 0] = 2.0 * (y[:, 0] - t) / B
    z = y[:, 1:4] - y[:, 1:4].max(axis=1, keepdims=True)
    sm = np.exp(z) / np.exp(z).sum(axis=1, keepdims=True)
    sm[np.arange(B), c] -= 1.0
    dy[:, 1:4] = sm / B
    dW3 = h2.T @ dy
    db3 = dy.sum(0)
    dh2 = (dy @ W3.T) * (a2 > 0)
    dW2 = h1.T @ dh2
    db2 = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (a1 > 0)
    dW1 = X.T @ dh1
    db1 = dh1.sum(0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])


def finite_difference(p, X, t, c, idx, eps=1e-6):
    g = np.zeros(len(idx))
    for k, i in enumerate(idx):
        q = p.astype(np.float64).copy()
        q[i] += eps
        lp = loss(q, X, t, c)
        q[i] -= 2 * eps
        lm = loss(q, X, t, c)
        g[k] = (lp - lm) / (2 * eps)
    return g
