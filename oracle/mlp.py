"""Float64 numpy restatement of the engine's MLP (pathfinding_gpu_amd/csrc/mlp.hip) -- TEST INFRASTRUCTURE ONLY.

The network has no reference counterpart; this oracle is pinned by central
finite differences (tests/test_mlp_oracle.py), not by reference vectors.
Parameter layout: W1[H][in], b1[H], W2[H][H], b2[H], W3[out][H], b3[out].
"""
import numpy as np


def split(p, n_in, h, n_out):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + h * n_in].reshape(h, n_in); o += h * n_in
    b1 = p[o:o + h]; o += h
    W2 = p[o:o + h * h].reshape(h, h); o += h * h
    b2 = p[o:o + h]; o += h
    W3 = p[o:o + n_out * h].reshape(n_out, h); o += n_out * h
    b3 = p[o:o + n_out]; o += n_out
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def num_params(n_in, h, n_out):
    return h * n_in + h + h * h + h + n_out * h + n_out


def forward(p, X, n_in, h, n_out):
    W1, b1, W2, b2, W3, b3 = split(p, n_in, h, n_out)
    X = np.asarray(X, np.float64)
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    return h2 @ W3.T + b3, (X, a1, h1, a2, h2)


def loss(p, X, T, M, n_in, h, n_out):
    Y, _ = forward(p, X, n_in, h, n_out)
    M = np.asarray(M, np.float64)
    return float((M * (Y - T) ** 2).sum() / max(M.sum(), 1e-12))


def loss_grad(p, X, T, M, n_in, h, n_out):
    W1, b1, W2, b2, W3, b3 = split(p, n_in, h, n_out)
    Y, (X, a1, h1, a2, h2) = forward(p, X, n_in, h, n_out)
    M = np.asarray(M, np.float64)
    ms = max(M.sum(), 1e-12)
    L = float((M * (Y - T) ** 2).sum() / ms)
    dY = 2.0 * M * (Y - T) / ms
    gW3 = dY.T @ h2
    gb3 = dY.sum(0)
    d2 = (dY @ W3) * (a2 > 0)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    return L, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def numeric_grad(p, X, T, M, n_in, h, n_out, idx, eps=1e-6):
    g = []
    for i in idx:
        q = np.array(p, np.float64)
        q[i] += eps
        lp = loss(q, X, T, M, n_in, h, n_out)
        q[i] -= 2 * eps
        lm = loss(q, X, T, M, n_in, h, n_out)
        g.append((lp - lm) / (2 * eps))
    return np.array(g)
