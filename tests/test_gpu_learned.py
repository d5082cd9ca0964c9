"""GPU: the learned-Delta network's HIP kernels against the float64 oracle, and the trained model's effect.

Tolerances (FP32 kernels vs float64 oracle): forward rtol 1e-5 / atol 1e-5,
loss rtol 1e-5, gradient rtol 1e-4 / atol 1e-6.
"""
import os

import numpy as np
import pytest

import pathfinding_gpu_amd as P
from pathfinding_gpu_amd import learned as L
from oracle import mlp as OM

pytestmark = pytest.mark.gpu


def _rand(seed, n=50, n_in=10, h=32, n_out=14):
    rng = np.random.default_rng(seed)
    p = (rng.normal(0, 0.3, OM.num_params(n_in, h, n_out))).astype(np.float32)
    X = rng.normal(0, 1, (n, n_in)).astype(np.float32)
    T = rng.normal(0, 1, (n, n_out)).astype(np.float32)
    M = (rng.random((n, n_out)) > 0.2).astype(np.float32)
    return p, X, T, M, n_in, h, n_out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_mlp_forward_loss_grad_vs_oracle(seed):
    p, X, T, M, a, h, b = _rand(seed)
    net = L.Mlp(a, h, b)
    try:
        net.set_params(p)
        Y = net.forward(X)
        Yo, _ = OM.forward(p.astype(np.float64), X, a, h, b)
        np.testing.assert_allclose(Y, Yo, rtol=1e-5, atol=1e-5)
        loss, g = net.loss_grad(X, T, M)
        lo, go = OM.loss_grad(p.astype(np.float64), X, T, M, a, h, b)
        np.testing.assert_allclose(loss, lo, rtol=1e-5)
        np.testing.assert_allclose(g, go, rtol=1e-4, atol=1e-6)
    finally:
        net.close()


def test_mlp_training_fits_a_function():
    rng = np.random.default_rng(4)
    X = rng.uniform(-1, 1, (200, 10)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2], np.abs(X[:, 3]) - X[:, 4]], 1).astype(np.float32)
    M = np.ones_like(T)
    net = L.Mlp(10, 32, 2, seed=3)
    try:
        l0, _ = net.loss_grad(X, T, M)
        l1 = net.train(X, T, M, epochs=1500, lr=3e-3)
        assert l1 < 0.15 * l0, (l0, l1)
        # the trained parameters reproduce the loss in the oracle
        lo = OM.loss(net.params().astype(np.float64), X, T, M, 10, 32, 2)
        np.testing.assert_allclose(l1, lo, rtol=1e-4)
    finally:
        net.close()


def _held_out_graphs():
    # none of these is in delta_sweep.family(): other sizes and seeds
    return [P.grid_road(450, 450, 991), P.grid_road(250, 250, 993, wlo=1.0, whi=3.0), P.rmat(15, 12, 997),
            P.rmat(17, 8, 999)]


@pytest.mark.skipif(not os.path.exists(L.MODEL_PATH), reason="no trained model file")
def test_learned_delta_is_bitexact_and_not_slower(orc):
    model = L.DeltaModel.load()
    tuner = L.Tuner(model)
    tot_def = tot_pick = 0.0
    for g in _held_out_graphs():
        src = g.max_degree_vertex()
        with P.GpuSolver(g) as sv:
            def tf(delta, mapping):
                o = P.GpuOptions(drive=P.Drive.Persistent, delta=delta, mapping=P.Mapping(mapping))
                sv.sssp(src, o, copy_back=False)
                return min(sv.sssp(src, o, copy_back=False).ms for _ in range(3))
            res = tuner.choose(sv, g, 1, tf)
            pick = model.choose(g, 1)
            o = P.GpuOptions(drive=P.Drive.Persistent, delta=pick["delta"], mapping=P.Mapping(pick["mapping"]))
            r = sv.sssp(src, o)
            d, _, pc, _, _ = orc.dijkstra(g.row_offsets, g.col_indices, g.weights, src)
            assert np.array_equal(r.dist, d) and np.array_equal(r.parent, pc), g.name
            t_def = res.get("default_ms", tf(res["auto_delta"], 2))
            t_sel = tf(res["delta"], res["mapping"])
            print(f"{g.name}: davidson {res['auto_delta']} {t_def:.3f} ms | mlp {pick['delta']}/{pick['mapping']} "
                  f"{res.get('predicted_ms', float('nan')):.3f} ms | selected {res['source']}")
            tot_def += t_def
            tot_pick += t_sel
    print(f"learned-Delta gain over Davidson (held-out graphs): {tot_def / tot_pick:.3f}x")
    assert tot_pick <= 1.10 * tot_def
