"""CPU (numpy, float64) forward/backward of the fusion cost model's MLP.

TEST INFRASTRUCTURE ONLY: the checker of the engine's MLP kernels
(qvm_amd/csrc/mlp.hip).  The network is din -> h -> h -> 1 with ReLU; the
flat parameter vector is W1[h][din], b1[h], W2[h][h], b2[h], w3[h], b3; the
loss is mean_b 0.5 (f(x_b) - t_b)^2.  The cost model has no counterpart in
the reference; this restatement is pinned by central finite differences in
tests/test_costmodel.py.
"""
from __future__ import annotations

import numpy as np


def unpack(p: np.ndarray, din: int, h: int):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + h * din].reshape(h, din); o += h * din
    b1 = p[o:o + h]; o += h
    W2 = p[o:o + h * h].reshape(h, h); o += h * h
    b2 = p[o:o + h]; o += h
    w3 = p[o:o + h]; o += h
    b3 = p[o]; o += 1
    assert o == p.size
    return W1, b1, W2, b2, w3, b3


def num_params(din: int, h: int) -> int:
    return h * din + h + h * h + h + h + 1


def forward(p, X, din, h):
    W1, b1, W2, b2, w3, b3 = unpack(p, din, h)
    X = np.asarray(X, dtype=np.float64)
    a1 = np.maximum(X @ W1.T + b1, 0.0)
    a2 = np.maximum(a1 @ W2.T + b2, 0.0)
    return a2 @ w3 + b3


def loss_and_grad(p, X, T, din, h):
    W1, b1, W2, b2, w3, b3 = unpack(p, din, h)
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    nb = X.shape[0]
    z1 = X @ W1.T + b1
    a1 = np.maximum(z1, 0.0)
    z2 = a1 @ W2.T + b2
    a2 = np.maximum(z2, 0.0)
    y = a2 @ w3 + b3
    e = (y - T) / nb
    loss = 0.5 * np.sum((y - T) ** 2) / nb
    g_w3 = a2.T @ e
    g_b3 = e.sum()
    d2 = np.outer(e, w3) * (a2 > 0)
    g_W2 = d2.T @ a1
    g_b2 = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    g_W1 = d1.T @ X
    g_b1 = d1.sum(0)
    g = np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3]])
    return loss, g
