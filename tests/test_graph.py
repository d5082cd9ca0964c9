"""Host graph layer (libpfgraph.so): generators, CSR builder, union, binary cache.

The product CSR builder is checked against the oracle's restatement of the
reference's buildCSR; the generators against the workload definitions they
implement.
"""
import math
import os

import numpy as np
import pytest

import pathfinding_gpu_amd as P
from pathfinding_gpu_amd import graph as G


def test_grid_road_definition():
    rows, cols, seed = 13, 17, 42
    n, src, dst, w, x, y = G.grid_road_edges(rows, cols, seed)
    assert n == rows * cols
    assert len(src) == 2 * (rows * (cols - 1) + cols * (rows - 1))
    # integer coordinates at 100*(c, r) + jitter in [-20, 20]
    c = np.arange(n) % cols
    r = np.arange(n) // cols
    jx, jy = x - 100 * c, y - 100 * r
    assert (x == np.floor(x)).all() and (y == np.floor(y)).all()
    assert jx.min() >= -20 and jx.max() <= 20 and jy.min() >= -20 and jy.max() <= 20
    assert len(np.unique(jx)) > 30                                  # jitter actually varies
    # 4-neighbour roads, both directions, shared weight
    assert (src[0::2] == dst[1::2]).all() and (dst[0::2] == src[1::2]).all() and (w[0::2] == w[1::2]).all()
    d = np.abs(src.astype(np.int64) - dst.astype(np.int64))
    assert set(np.unique(d)) == {1, cols}
    # w = ceil(len * U[1, 1.5))  =>  ceil(len) <= w <= ceil(1.5 len)
    ln = np.hypot(x[src] - x[dst], y[src] - y[dst])
    assert (w >= np.ceil(ln)).all() and (w <= np.ceil(1.5 * ln)).all()
    assert w.min() >= 1


def test_grid_config_sizes():
    for side, m in ((100, 39600), (1000, 3996000)):
        assert G.N.graphlib().pf_grid_road_arcs(side, side) == m


def test_generators_deterministic_and_seeded():
    a = G.grid_road_edges(20, 20, 3)
    b = G.grid_road_edges(20, 20, 3)
    c = G.grid_road_edges(20, 20, 4)
    assert all(np.array_equal(p, q) for p, q in zip(a[1:], b[1:]))
    assert not np.array_equal(a[3], c[3])
    r1 = G.rmat_edges(10, 4, 11)
    r2 = G.rmat_edges(10, 4, 11)
    assert all(np.array_equal(p, q) for p, q in zip(r1[1:], r2[1:]))


def test_rmat_definition():
    scale, ef = 12, 16
    n, src, dst, w = G.rmat_edges(scale, ef, 11)
    assert n == 4096 and len(src) == ef * n
    assert w.min() >= 1 and w.max() <= 255 and len(np.unique(w)) == 255
    # quadrant frequencies of the top bit: P(src bit) = c + d = 0.24, P(dst bit) = b + d = 0.24
    assert abs((src >> (scale - 1)).mean() - 0.24) < 0.01
    assert abs((dst >> (scale - 1)).mean() - 0.24) < 0.01
    assert abs(((src >> (scale - 1)) & (dst >> (scale - 1))).mean() - 0.05) < 0.005
    g = P.rmat(scale, ef, 11)
    assert g.m == int((src != dst).sum())           # self-loops dropped, duplicates kept
    deg = np.diff(g.row_offsets)
    assert g.max_degree_vertex() == int(np.argmax(deg)) == 0


@pytest.mark.parametrize("maker", ["grid", "rmat", "random"])
def test_build_csr_matches_oracle(orc, maker):
    if maker == "grid":
        n, s, d, w, x, y = G.grid_road_edges(30, 40, 9)
    elif maker == "rmat":
        n, s, d, w = G.rmat_edges(11, 8, 5)
        x = y = None
    else:
        rng = np.random.default_rng(1)
        n = 500
        s = rng.integers(0, n, 5000).astype(np.uint32)
        d = rng.integers(0, n, 5000).astype(np.uint32)
        w = rng.integers(0, 9, 5000).astype(np.uint32)
        x = y = None
    g = P.build_csr(n, s, d, w, x, y)
    row, col, wo = orc.build_csr(n, s, d, w)
    assert np.array_equal(g.row_offsets, row) and np.array_equal(g.col_indices, col) and np.array_equal(g.weights, wo)
    st = orc.csr_stats(row, col, wo, x, y)
    assert (g.min_w, g.max_w, g.max_deg) == (st["minW"], st["maxW"], st["maxDeg"])
    assert (g.avg_w, g.avg_deg, g.h_scale) == (st["avgW"], st["avgDeg"], st["hScale"])


def test_union_and_offsets(orc):
    gs = [P.grid_road(a, a, 10 + a) for a in (5, 9, 3, 12)]
    u, voff = P.union(gs)
    assert voff.tolist() == [0, 25, 106, 115, 259]
    assert u.n == 259 and u.m == sum(g.m for g in gs)
    for i, g in enumerate(gs):
        b, e = int(voff[i]), int(voff[i + 1])
        rows = u.row_offsets[b:e + 1] - u.row_offsets[b]
        assert np.array_equal(rows, g.row_offsets)
        k0 = int(u.row_offsets[b])
        assert np.array_equal(u.col_indices[k0:k0 + g.m] - b, g.col_indices)
        assert np.array_equal(u.weights[k0:k0 + g.m], g.weights)
    # multi-seed Dijkstra on the union = per-graph Dijkstra
    seeds = [int(voff[i]) for i in range(len(gs))]
    d, pc = orc.dijkstra_multi_seed(u.row_offsets, u.col_indices, u.weights, seeds)
    for i, g in enumerate(gs):
        di, _, pi, _, _ = orc.dijkstra(g.row_offsets, g.col_indices, g.weights, 0)
        assert np.array_equal(d[voff[i]:voff[i + 1]], di)


def test_binary_cache_roundtrip(tmp_path):
    g = P.grid_road(15, 11, 2)
    p = str(tmp_path / "g.bin")
    P.save_binary(g, p)
    h = P.load_binary(p)
    assert h.n == g.n and h.m == g.m
    for a in ("row_offsets", "col_indices", "weights", "x", "y"):
        assert np.array_equal(getattr(g, a), getattr(h, a))
    assert h.h_scale == g.h_scale and h.avg_w == g.avg_w
    with open(p, "rb") as f:
        assert f.read(8) == b"PFCSR001"
    r = P.rmat(8, 4, 1)
    P.save_binary(r, p)
    h = P.load_binary(p)
    assert h.x is None and np.array_equal(h.col_indices, r.col_indices)
    with open(p, "r+b") as f:
        f.write(b"XXXXXXXX")
    with pytest.raises(OSError):
        P.load_binary(p)


def test_reference_binary_layout(tmp_path, orc):
    """The cache layout matches the reference's saveBinary (src/graph.cpp:322-336) field by field."""
    g = P.grid_road(4, 3, 1)
    p = str(tmp_path / "g.bin")
    P.save_binary(g, p)
    raw = open(p, "rb").read()
    n, m, hasxy, nlen = np.frombuffer(raw[8:24], np.uint32)
    assert (n, m, hasxy) == (g.n, g.m, 1)
    off = 24 + nlen
    row = np.frombuffer(raw[off:off + 4 * (n + 1)], np.uint32)
    assert np.array_equal(row, g.row_offsets)
    off += 4 * (n + 1) + 8 * m
    x = np.frombuffer(raw[off:off + 8 * n], np.float64)
    assert np.array_equal(x, g.x)
