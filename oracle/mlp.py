"""FP64 restatement of the fusion-planner MLP -- TEST INFRASTRUCTURE ONLY.

Forward pass, mean-squared-error loss and its gradient for the packed
parameter layout of csrc/planner.hip (W1[16][64], b1, W2[64][64], b2, w3, b3),
written directly from the chain rule. tests/test_oracle_pins.py pins it by
central finite differences; the GPU kernels are checked against it.
"""

import numpy as np

F, H = 16, 64


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    o = 0
    W1 = w[o:o + F * H].reshape(F, H); o += F * H
    b1 = w[o:o + H]; o += H
    W2 = w[o:o + H * H].reshape(H, H); o += H * H
    b2 = w[o:o + H]; o += H
    w3 = w[o:o + H]; o += H
    b3 = w[o]
    return W1, b1, W2, b2, w3, b3


def forward(w, X):
    W1, b1, W2, b2, w3, b3 = unpack(w)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ W1 + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0.0)
    return h2 @ w3 + b3, (X, z1, h1, z2, h2)


def loss(w, X, y):
    out, _ = forward(w, X)
    return float(np.mean((out - np.asarray(y, dtype=np.float64)) ** 2))


def grad(w, X, y):
    W1, b1, W2, b2, w3, b3 = unpack(w)
    out, (X, z1, h1, z2, h2) = forward(w, X)
    n = len(out)
    dy = 2.0 * (out - np.asarray(y, dtype=np.float64)) / n
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    d2 = np.outer(dy, w3) * (z2 > 0)
    gW2 = h1.T @ d2
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ W2.T) * (z1 > 0)
    gW1 = X.T @ d1
    gb1 = d1.sum(axis=0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]]), float(np.mean((out - y) ** 2))
