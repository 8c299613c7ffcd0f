"""Float64 restatement of the planner's 3-layer perceptron (TEST
INFRASTRUCTURE ONLY): forward pass, loss and backpropagated gradient with
the parameter layout of the library (W1 64x8, b1, W2 64x64, b2, W3 2x64, b3,
row-major).  Only tests/ use it, to check the GPU kernels; it is pinned by
central finite differences of its own loss.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 8, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_H * N_IN
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_OUT * N_H
NPARAM = OFF_B3 + N_OUT


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(N_H, N_IN), p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(N_H, N_H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(N_OUT, N_H), p[OFF_B3:NPARAM])


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    return h2 @ W3.T + b3, (X, z1, h1, z2, h2)


def loss(p, X, T, M):
    Y, _ = forward(p, X)
    return float(np.mean(np.sum(M * (Y - T) ** 2, axis=1)))


def gradient(p, X, T, M):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    Y, (X, z1, h1, z2, h2) = forward(p, X)
    B = X.shape[0]
    d3 = 2.0 / B * np.asarray(M) * (Y - np.asarray(T))        # B x 2
    gW3 = d3.T @ h2
    gb3 = d3.sum(0)
    d2 = (d3 @ W3) * (z2 > 0)                                   # B x 64
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (z1 > 0)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def finite_difference(p, X, T, M, idx, h=1e-6):
    p = np.array(p, dtype=np.float64)
    out = []
    for i in idx:
        a = p.copy(); a[i] += h
        b = p.copy(); b[i] -= h
        out.append((loss(a, X, T, M) - loss(b, X, T, M)) / (2 * h))
    return np.array(out)
