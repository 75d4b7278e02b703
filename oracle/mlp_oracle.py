"""NumPy restatement of the fusion cost model's network (TEST INFRASTRUCTURE ONLY).

Only tests/ import this module, to check the HIP kernels of
dm-sim_amd/csrc/mlp.hip (forward, backward, Adam) against it; the product
path never does.  The network has no counterpart in the reference (whose
executor applies every gate as two sweeps and plans nothing), so this file
follows the model's own definition:

  params (float32[P], row-major):  W1[H][F], b1[H], W2[H][H], b2[H], w3[H], b3
  f(x)  = w3 . relu(W2 relu(W1 x + b1) + b2) + b3
  loss  = mean_s (f(x_s) - y_s)^2
  Adam  (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected)

Everything is computed in float64.  `loss_grad` is pinned by central finite
differences of `loss` in the tests (a dropped term, a transposed weight or a
wrong ReLU mask fails that check); `forward` is pinned by hand-built weights
with closed-form outputs.
"""
from __future__ import annotations

import numpy as np

F = 14
H = 64
P = H * F + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    assert p.size == P
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]]).astype(np.float64)


def forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    h1 = np.maximum(X @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ w3 + b3


def loss(p, X, y):
    d = forward(p, X) - np.asarray(y, dtype=np.float64)
    return float(np.mean(d * d))


def loss_grad(p, X, y):
    """Gradient of `loss` with respect to the parameter vector (backprop)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0.0)
    out = h2 @ w3 + b3
    dy = 2.0 * (out - y) / len(y)              # [S]
    g_w3 = h2.T @ dy
    g_b3 = dy.sum()
    d2 = np.outer(dy, w3) * (a2 > 0)           # [S, H]
    g_W2 = d2.T @ h1
    g_b2 = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)                  # [S, H]
    g_W1 = d1.T @ X
    g_b1 = d1.sum(0)
    return pack(g_W1, g_b1, g_W2, g_b2, g_w3, g_b3)


def adam_step(p, g, m, v, lr, step, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update; returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
