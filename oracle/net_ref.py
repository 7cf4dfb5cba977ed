"""CPU reference of the con2prim guess network -- TEST INFRASTRUCTURE ONLY.

float64 numpy forward pass and mean-squared-error gradient of the 8-64-64-1
ReLU network whose weight-vector layout is gpluto_dev_amd/csrc/c2p_net.hpp's
(mu[8], sd[8], W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3, ysc, 0).
Used by tests/ to check the GPU training kernel's gradients and the
matrix-core inference.  The gradient is itself pinned by central finite
differences in tests/test_net_reference.py.
"""
from __future__ import annotations

import numpy as np

F = 8
H = 64
MU, SD, W1 = 0, F, 2 * F
B1 = W1 + F * H
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H
YSC = B3 + 1
SIZE = YSC + 2
TRAIN = slice(W1, B3 + 1)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return (w[MU:MU + F], w[SD:SD + F], w[W1:B1].reshape(H, F), w[B1:W2],
            w[W2:B2].reshape(H, H), w[B2:W3], w[W3:B3], w[B3], w[YSC])


def _layers(w, x):
    mu, sd, w1, b1, w2, b2, w3, b3, ysc = unpack(w)
    xn = (np.asarray(x, dtype=np.float64) - mu) / sd
    z1 = xn @ w1.T + b1
    a1 = np.maximum(z1, 0.0)
    z2 = a1 @ w2.T + b2
    a2 = np.maximum(z2, 0.0)
    return xn, z1, a1, z2, a2, a2 @ w3 + b3


def forward(w, x):
    """x[n, 8] raw features -> normalised network output[n] (the correction
    of log W is ysc times it)."""
    return _layers(w, x)[-1]


def loss_and_grad(w, x, y):
    """Mean over samples of (out - y/ysc)^2 and its gradient (SIZE vector,
    zero outside the trainable range)."""
    _, _, _, _, w2, _, w3, _, ysc = unpack(w)
    xn, z1, a1, z2, a2, out = _layers(w, x)
    yt = np.asarray(y, dtype=np.float64) / ysc
    n = xn.shape[0]
    err = out - yt
    g = 2.0 * err / n
    d2 = (g[:, None] * w3[None, :]) * (z2 > 0)
    d1 = (d2 @ w2) * (z1 > 0)
    grad = np.zeros(SIZE)
    grad[W1:B1] = (d1.T @ xn).ravel()
    grad[B1:W2] = d1.sum(0)
    grad[W2:B2] = (d2.T @ a1).ravel()
    grad[B2:W3] = d2.sum(0)
    grad[W3:B3] = g @ a2
    grad[B3] = g.sum()
    return float(np.mean(err * err)), grad


def random_weights(x, y, seed=0):
    """He-scaled weights with the data's normalisation (for checks)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(SIZE)
    x = np.asarray(x, dtype=np.float64)
    w[MU:MU + F] = x.mean(0)
    w[SD:SD + F] = np.maximum(x.std(0), 1e-6)
    w[W1:B1] = rng.normal(0, np.sqrt(2 / F), F * H)
    w[B1:W2] = rng.normal(0, 0.1, H)
    w[W2:B2] = rng.normal(0, np.sqrt(2 / H), H * H)
    w[B2:W3] = rng.normal(0, 0.1, H)
    w[W3:B3] = rng.normal(0, 1 / np.sqrt(H), H)
    w[B3] = 0.05
    w[YSC] = max(float(np.sqrt(np.mean(np.asarray(y, dtype=np.float64) ** 2))), 1e-12)
    return w.astype(np.float32)
