"""Oracle: numpy forward/backward of the proposal MLP.  Test infrastructure only.

Follows the network and loss defined for the learned importance proposal (no
reference counterpart): s(x) = w3 . silu(W2 silu(W1 x + b1) + b2) + b3 with bf16
(round-to-nearest-even) weights, loss l = y exp(-s) + s, dl/ds clipped to [-clip, 1].
"""

import numpy as np

H, IN = 64, 4


def bf16_round(w):
    b = np.asarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    o = 0
    W1 = w[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = w[o:o + H]; o += H
    W2 = w[o:o + H * H].reshape(H, H); o += H * H
    b2 = w[o:o + H]; o += H
    w3 = w[o:o + H]; o += H
    b3 = w[o]
    return W1, b1, W2, b2, w3, b3


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(w_bf16, x):
    W1, b1, W2, b2, w3, b3 = unpack(w_bf16)
    z1 = x @ W1.T + b1
    h1 = silu(z1)
    z2 = h1 @ W2.T + b2
    h2 = silu(z2)
    return h2 @ w3 + b3, (z1, h1, z2, h2)


def grad(w_bf16, x, y, clip):
    """Mean gradient (flat, same layout) and mean loss."""
    W1, b1, W2, b2, w3, b3 = unpack(w_bf16)
    x = np.asarray(x, dtype=np.float64)
    s, (z1, h1, z2, h2) = forward(w_bf16, x)
    ye = y * np.exp(-s)
    g = np.maximum(1.0 - ye, -clip)
    n = x.shape[0]
    d2 = (g[:, None] * w3[None, :]) * dsilu(z2)
    d1 = (d2 @ W2) * dsilu(z1)
    out = np.concatenate([(d1.T @ x).ravel(), d1.sum(0), (d2.T @ h1).ravel(), d2.sum(0), h2.T @ g, [g.sum()]]) / n
    return out, float(np.mean(ye + s))
