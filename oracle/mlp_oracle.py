"""CPU float64 restatement of the V' surrogate network -- TEST INFRASTRUCTURE ONLY.

Same parameter layout as the device kernels (W1[32][2], b1[32], W2[32][32], b2[32], w3[32], b3)
and the same loss L = (1/2n) sum (y - t)^2.  Used by the tests to check the device forward
pass and gradients; pinned itself by central finite differences (tests/test_oracle_mlp.py).
"""
from __future__ import annotations

import numpy as np

H = 32
NP = 2 * H + H + H * H + H + H + 1  # 1185


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + 2 * H].reshape(H, 2); o += 2 * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64).reshape(-1, 2)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ w3 + b3, (X, h1, h2)


def loss_grad(p, X, Y, W=None):
    """(L, dL/dp) for L = (1/2n) sum w (y - t)^2 (w = 1 if W is None)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, (X, h1, h2) = forward(p, X)
    Y = np.asarray(Y, dtype=np.float64).ravel()
    n = y.size
    r = y - Y
    w = np.ones(n) if W is None else np.asarray(W, dtype=np.float64).ravel()
    L = 0.5 * np.dot(w * r, r) / n
    dy = w * r / n
    g_w3 = h2.T @ dy
    g_b3 = dy.sum()
    d2 = np.outer(dy, w3) * (1.0 - h2 ** 2)
    g_W2 = d2.T @ h1
    g_b2 = d2.sum(0)
    d1 = (d2 @ W2) * (1.0 - h1 ** 2)
    g_W1 = d1.T @ X
    g_b1 = d1.sum(0)
    g = np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3]])
    return L, g


def random_params(seed=0, scale=0.3):
    """Random test parameters (float32)."""
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, scale, NP).astype(np.float32)
