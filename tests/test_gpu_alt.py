"""ALT (landmark) heuristic for A*: heuristic values against landmark distances from
the oracle's Dijkstra, and A* results (target distance, canonical path) against the
oracle's canonical A* -- a consistent heuristic only changes what is explored."""
import numpy as np
import pytest

import pathfinding_gpu_amd as P
from pathfinding_gpu_amd import Drive as DR, GpuAlgo as GA, Heuristic as HK

pytestmark = pytest.mark.gpu
INF = 0xFFFFFFFF


def alt_bound(dl, t, symmetric):
    """max over landmarks of d_L(t) - d_L(v) (and d_L(v) - d_L(t) when symmetric), 0 floor;
    landmarks that miss v or t give no bound.  dl: [L][n] distances from each landmark."""
    a = dl[:, t][:, None].astype(np.int64)
    b = dl.astype(np.int64)
    ok = (dl[:, t][:, None] != INF) & (dl != INF)
    d = np.where(ok, a - b, 0)
    if symmetric:
        d = np.where(ok, np.abs(a - b), 0)
    return np.maximum(d.max(axis=0), 0).astype(np.uint32)


@pytest.fixture(scope="module")
def grid():
    return P.grid_road(60, 60, 17)


@pytest.fixture(scope="module")
def rmat():
    return P.rmat(11, 8, 5)


def test_landmarks_and_heuristic_values(orc, grid, rmat):
    for g, want_sym in ((grid, True), (rmat, False)):
        with P.GpuSolver(g, max_batch=4) as sv:
            ids, sym = sv.build_landmarks(8)
            assert sym == want_sym and len(set(ids)) == 8
            dl = np.stack([orc.dijkstra(g.row_offsets, g.col_indices, g.weights, int(L))[0] for L in ids])
            # farthest-first: each landmark maximizes the distance to the ones before it
            d0 = orc.dijkstra(g.row_offsets, g.col_indices, g.weights, 0)[0]
            fin = d0 != INF
            assert d0[ids[0]] == d0[fin].max()
            for t in (0, g.n // 2, g.n - 1):
                h = sv.debug_heuristic(HK.ALT, t)
                assert np.array_equal(h, alt_bound(dl, t, sym)), (g.name, t)
                # admissible: never above the true distance to t (symmetric graph: d(v,t) = d(t,v))
                if sym:
                    dt = orc.dijkstra(g.row_offsets, g.col_indices, g.weights, t)[0]
                    assert np.all(h[dt != INF] <= dt[dt != INF])


@pytest.mark.parametrize("drive", [DR.Host, DR.Persistent], ids=lambda a: a.name)
@pytest.mark.parametrize("algo", [GA.DeltaStep, GA.Frontier], ids=lambda a: a.name)
def test_alt_astar_matches_oracle(orc, grid, drive, algo):
    with P.GpuSolver(grid, max_batch=16) as sv:
        ids, sym = sv.build_landmarks(8)
        rng = np.random.default_rng(3)
        srcs = rng.integers(0, grid.n, 16)
        dsts = rng.integers(0, grid.n, 16)
        o = P.GpuOptions(algo=algo, drive=drive, heuristic=HK.ALT)
        res = sv.astar_batch(srcs, dsts, o)
        eu = sv.astar_batch(srcs, dsts, P.GpuOptions(algo=algo, drive=drive, heuristic=HK.FloorEuclid))
        for i in range(16):
            d, path, _, _ = orc.astar(grid.row_offsets, grid.col_indices, grid.weights, grid.x, grid.y,
                                      grid.h_scale, orc.H_FLOOR_EUCLID, int(srcs[i]), int(dsts[i]))
            assert res["dist"][i] == d and np.array_equal(res["paths"][i], path), i
        # the landmarks prune at least as much as floor(Euclid); with the bucketed (Delta) order, where
        # the target bound settles early, strictly more
        assert res["stats"].edges_scanned <= eu["stats"].edges_scanned
        if algo == GA.DeltaStep:
            assert res["stats"].edges_scanned < eu["stats"].edges_scanned


@pytest.mark.parametrize("async_light", [1, 2])
def test_alt_astar_asynchronous_and_tiled(orc, grid, async_light):
    """ALT with the asynchronous and the tiled light phase: same distances and paths."""
    with P.GpuSolver(grid, max_batch=16) as sv:
        sv.build_landmarks(8)
        rng = np.random.default_rng(5)
        srcs = rng.integers(0, grid.n, 16)
        dsts = rng.integers(0, grid.n, 16)
        res = sv.astar_batch(srcs, dsts, P.GpuOptions(drive=DR.Persistent, heuristic=HK.ALT, async_light=async_light))
        for i in range(16):
            d, path, _, _ = orc.astar(grid.row_offsets, grid.col_indices, grid.weights, grid.x, grid.y,
                                      grid.h_scale, orc.H_FLOOR_EUCLID, int(srcs[i]), int(dsts[i]))
            assert res["dist"][i] == d and np.array_equal(res["paths"][i], path), i


def test_alt_on_directed_graph_and_given_landmarks(orc, rmat):
    with P.GpuSolver(rmat, max_batch=4) as sv:
        ids, sym = sv.build_landmarks(4, ids=[1, 2, 3, 4])
        assert ids == [1, 2, 3, 4] and not sym
        rng = np.random.default_rng(9)
        for _ in range(6):
            s, t = (int(v) for v in rng.integers(0, rmat.n, 2))
            r = sv.astar(s, t, P.GpuOptions(heuristic=HK.ALT, drive=DR.Persistent))
            d, path, _, _ = orc.astar(rmat.row_offsets, rmat.col_indices, rmat.weights, None, None, 0.0,
                                      orc.H_NONE, s, t)
            assert r.target_dist == d and np.array_equal(r.path, path), (s, t)


def test_alt_requires_landmarks(grid):
    with P.GpuSolver(grid) as sv:
        with pytest.raises(RuntimeError):
            sv.astar(0, 5, P.GpuOptions(heuristic=HK.ALT))
        with pytest.raises(RuntimeError):
            sv.build_landmarks(6)
