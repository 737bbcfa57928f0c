"""FP64 numpy restatement of the router MLP — TEST ONLY.

The network of sable_amd/csrc/sable_mlp.h: hidden = relu(W1 x + b1),
out = W2 hidden + b2, loss = scale * sum(mask * (out - target)^2).  The
gradients are the textbook backpropagation of that expression; they are
pinned independently by central finite differences (``fd_grad``).
Parameter layout: [W1 (H x F), b1 (H), W2 (K x H), b2 (K)].
"""
from __future__ import annotations

import numpy as np


def unpack(p, F, H, K):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + K * H].reshape(K, H); o += K * H
    b2 = p[o:o + K]
    return W1, b1, W2, b2


def forward(p, X, F, H, K):
    W1, b1, W2, b2 = unpack(p, F, H, K)
    X = np.asarray(X, dtype=np.float64)
    hid = np.maximum(X @ W1.T + b1, 0.0)
    return hid @ W2.T + b2, hid


def loss(p, X, T, M, F, H, K, scale):
    out, _ = forward(p, X, F, H, K)
    return scale * np.sum(M * (out - T) ** 2)


def grad(p, X, T, M, F, H, K, scale):
    """Returns (dout, grad) of the masked squared error."""
    W1, b1, W2, b2 = unpack(p, F, H, K)
    X = np.asarray(X, dtype=np.float64)
    out, hid = forward(p, X, F, H, K)
    dout = scale * 2.0 * (out - T) * M
    gW2 = dout.T @ hid
    gb2 = dout.sum(axis=0)
    dh = (dout @ W2) * (hid > 0)
    gW1 = dh.T @ X
    gb1 = dh.sum(axis=0)
    return dout, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2])


def fd_grad(p, X, T, M, F, H, K, scale, eps=1e-6):
    p = np.asarray(p, dtype=np.float64)
    g = np.zeros_like(p)
    for i in range(len(p)):
        a = p.copy(); a[i] += eps
        b = p.copy(); b[i] -= eps
        g[i] = (loss(a, X, T, M, F, H, K, scale) - loss(b, X, T, M, F, H, K, scale)) / (2 * eps)
    return g


def adam(p, g, m, v, lr, b1, b2, eps, step):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
