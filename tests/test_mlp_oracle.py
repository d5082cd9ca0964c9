"""Pins the MLP oracle: analytic gradient vs central finite differences, and forward vs a hand expansion."""
import numpy as np

from oracle import mlp


def _data(seed, n=20, n_in=6, h=9, n_out=4):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.5, mlp.num_params(n_in, h, n_out))
    X = rng.normal(0, 1, (n, n_in))
    T = rng.normal(0, 1, (n, n_out))
    M = (rng.random((n, n_out)) > 0.3).astype(np.float64)
    return p, X, T, M, n_in, h, n_out


def test_gradient_matches_finite_differences():
    for seed in range(4):
        p, X, T, M, a, h, b = _data(seed)
        L, g = mlp.loss_grad(p, X, T, M, a, h, b)
        assert L == mlp.loss(p, X, T, M, a, h, b)
        idx = np.arange(0, p.size, 3)
        num = mlp.numeric_grad(p, X, T, M, a, h, b, idx)
        np.testing.assert_allclose(g[idx], num, rtol=1e-5, atol=1e-8)


def test_forward_hand_expansion():
    p, X, T, M, a, h, b = _data(7, n=3)
    W1, b1, W2, b2, W3, b3 = mlp.split(p, a, h, b)
    Y, _ = mlp.forward(p, X, a, h, b)
    for s in range(3):
        h1 = [max(0.0, sum(W1[j, i] * X[s, i] for i in range(a)) + b1[j]) for j in range(h)]
        h2 = [max(0.0, sum(W2[j, i] * h1[i] for i in range(h)) + b2[j]) for j in range(h)]
        y = [sum(W3[j, i] * h2[i] for i in range(h)) + b3[j] for j in range(b)]
        np.testing.assert_allclose(Y[s], y, rtol=1e-12)


def test_mask_zero_removes_terms():
    p, X, T, M, a, h, b = _data(3)
    M2 = M.copy()
    M2[:, 0] = 0
    T2 = T.copy()
    T2[:, 0] += 100.0        # masked-out targets must not matter
    assert mlp.loss(p, X, T2, M2, a, h, b) == mlp.loss(p, X, T, M2, a, h, b)
