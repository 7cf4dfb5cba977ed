"""CPU reference of the con2prim guess network -- TEST INFRASTRUCTURE ONLY.

float64 numpy forward pass and mean-squared-error gradient of the 4-32-32-1
tanh network whose weight-vector layout is gpluto_dev_amd/csrc/c2p_net.hpp's
(mu[4], sd[4], W1[32][4], b1[32], W2[32][32], b2[32], W3[32], b3, ysc, 0).
Used by tests/ to check the GPU training kernel's gradients and the
matrix-core inference.  The gradient is itself pinned by central finite
differences in tests/test_net_reference.py.
"""
from __future__ import annotations

import numpy as np

H = 32
MU, SD, W1, B1 = 0, 4, 8, 8 + 4 * H
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H
YSC = B3 + 1
SIZE = YSC + 2
TRAIN = slice(W1, B3 + 1)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return (w[MU:MU + 4], w[SD:SD + 4], w[W1:B1].reshape(H, 4), w[B1:W2],
            w[W2:B2].reshape(H, H), w[B2:W3], w[W3:B3], w[B3], w[YSC])


def forward(w, x):
    """x[n, 4] raw features -> normalised network output[n]."""
    mu, sd, w1, b1, w2, b2, w3, b3, _ = unpack(w)
    xn = (np.asarray(x, dtype=np.float64) - mu) / sd
    a1 = np.tanh(xn @ w1.T + b1)
    a2 = np.tanh(a1 @ w2.T + b2)
    return a2 @ w3 + b3


def loss_and_grad(w, x, y):
    """Mean over samples of (out - y/ysc)^2 and its gradient (SIZE vector,
    zero outside the trainable range)."""
    mu, sd, w1, b1, w2, b2, w3, b3, ysc = unpack(w)
    xn = (np.asarray(x, dtype=np.float64) - mu) / sd
    yt = np.asarray(y, dtype=np.float64) / ysc
    n = xn.shape[0]
    a1 = np.tanh(xn @ w1.T + b1)
    a2 = np.tanh(a1 @ w2.T + b2)
    out = a2 @ w3 + b3
    err = out - yt
    g = 2.0 * err / n
    d2 = (g[:, None] * w3[None, :]) * (1.0 - a2 * a2)
    d1 = (d2 @ w2) * (1.0 - a1 * a1)
    grad = np.zeros(SIZE)
    grad[W1:B1] = (d1.T @ xn).ravel()
    grad[B1:W2] = d1.sum(0)
    grad[W2:B2] = (d2.T @ a1).ravel()
    grad[B2:W3] = d2.sum(0)
    grad[W3:B3] = g @ a2
    grad[B3] = g.sum()
    return float(np.mean(err * err)), grad


def random_weights(x, y, seed=0):
    """Glorot-scaled weights with the data's normalisation (for checks)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(SIZE)
    x = np.asarray(x, dtype=np.float64)
    w[MU:MU + 4] = x.mean(0)
    w[SD:SD + 4] = np.maximum(x.std(0), 1e-6)
    w[W1:B1] = rng.normal(0, 0.5, 4 * H)
    w[B1:W2] = rng.normal(0, 0.1, H)
    w[W2:B2] = rng.normal(0, 1 / np.sqrt(H), H * H)
    w[B2:W3] = rng.normal(0, 0.1, H)
    w[W3:B3] = rng.normal(0, 1 / np.sqrt(H), H)
    w[B3] = 0.05
    w[YSC] = max(float(np.sqrt(np.mean(np.asarray(y, dtype=np.float64) ** 2))), 1e-12)
    return w.astype(np.float32)
