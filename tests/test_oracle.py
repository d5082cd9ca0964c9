"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each restated function is checked against something that would catch a
plausible mistake in it:
  * the reference's own CPU code, compiled in place into oracle/_ref (when
    present) and the fixtures it wrote into tests/golden/ (always present);
  * the CSR known answer of the reference's include/graph.h:8-11;
  * closed forms and brute force (Bellman-Ford, min-index tight predecessor,
    math.isqrt, numpy searchsorted / cumsum).
"""
import math

import numpy as np
import pytest

from conftest import random_graph


# ---------------------------------------------------------------- CSR
def test_csr_known_answer(orc):
    # include/graph.h:8-11: 0->1(5), 0->2(3), 1->2(1), 2->0(7); after the (w,dst) row sort
    # row 0 becomes [2(3), 1(5)]
    row, col, w = orc.build_csr(3, [0, 0, 1, 2], [1, 2, 2, 0], [5, 3, 1, 7])
    assert row.tolist() == [0, 2, 3, 4]
    assert col.tolist() == [2, 1, 2, 0]
    assert w.tolist() == [3, 5, 1, 7]


def test_csr_drops_self_loops_and_clamps_weights(orc):
    row, col, w = orc.build_csr(3, [0, 1, 1, 2, 2], [0, 2, 1, 0, 0], [4, 0, 9, 7, 2])
    assert row.tolist() == [0, 0, 1, 3]
    assert col.tolist() == [2, 0, 0]
    assert w.tolist() == [1, 2, 7]          # w=0 -> 1; duplicates kept, sorted by (w, dst)


def test_csr_rows_sorted_and_multiset_preserved(orc):
    rng = np.random.default_rng(5)
    n = 50
    src, dst, w = random_graph(rng, n, 600)
    row, col, wo = orc.build_csr(n, src, dst, w)
    keep = src != dst
    assert row[-1] == keep.sum()
    want = sorted(zip(src[keep].tolist(), np.maximum(w[keep], 1).tolist(), dst[keep].tolist()))
    got = [(u, int(wo[k]), int(col[k])) for u in range(n) for k in range(row[u], row[u + 1])]
    assert got == want


def test_csr_stats_closed_form(orc):
    rng = np.random.default_rng(9)
    n = 40
    src, dst, w = random_graph(rng, n, 300, zero_w=False)
    row, col, wo = orc.build_csr(n, src, dst, w)
    x = rng.integers(-500, 500, n).astype(np.float64)
    y = rng.integers(-500, 500, n).astype(np.float64)
    s = orc.csr_stats(row, col, wo, x, y)
    assert s["minW"] == wo.min() and s["maxW"] == wo.max()
    assert s["avgW"] == pytest.approx(wo.mean(), rel=1e-12)
    assert s["avgDeg"] == pytest.approx(len(col) / n, rel=1e-12)
    us = np.repeat(np.arange(n), np.diff(row))
    e = np.hypot(x[us] - x[col], y[us] - y[col])
    ok = e > 1e-12
    assert s["hScale"] == pytest.approx((wo[ok] / e[ok]).min(), rel=1e-12)


# ---------------------------------------------------------------- Dijkstra
def _tight_min_parent(row, col, w, dist, seeds):
    n = len(row) - 1
    par = np.full(n, 0xFFFFFFFF, np.uint32)
    for u in range(n):
        if dist[u] == 0xFFFFFFFF:
            continue
        for k in range(row[u], row[u + 1]):
            v = col[k]
            if int(dist[u]) + int(w[k]) == int(dist[v]) and u < par[v]:
                par[v] = u
    for s in seeds:
        par[s] = 0xFFFFFFFF
    return par


@pytest.mark.parametrize("seed", range(12))
def test_dijkstra_vs_bellman_ford_and_tie_rule(orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 40))
    src, dst, w = random_graph(rng, n, int(rng.integers(0, 4 * n)), wmax=int(rng.integers(1, 6)))
    row, col, wo = orc.build_csr(n, src, dst, w)
    s = int(rng.integers(0, n))
    d, pf, pc, settled, edges = orc.dijkstra(row, col, wo, s)
    assert np.array_equal(d, orc.bellman_ford(row, col, wo, s))
    assert np.array_equal(pc, _tight_min_parent(row, col, wo, d, [s]))
    assert np.array_equal(pc, orc.canonical_parents(row, col, wo, d))
    # the reference's first-reached parent is a tight predecessor (verify.cpp:74-87 layer)
    rep = orc.verify_sssp(row, col, wo, s, d, d, pf, None, samples=16)
    assert sum(rep["errors"].values()) == 0
    assert settled == int((d != 0xFFFFFFFF).sum())


def test_dijkstra_weights_near_inf(orc):
    big = 0xFFFFFFF0
    row, col, w = orc.build_csr(4, [0, 1, 0, 2], [1, 2, 2, 3], [big, 20, 5, big])
    d, pf, pc, _, _ = orc.dijkstra(row, col, w, 0)
    # 0->1 = big; 0->2 = 5 (1->2 would be big+20 >= INF); 2->3 = 5+big = 0xFFFFFFF5 < INF
    assert d.tolist() == [0, big, 5, 0xFFFFFFF5]
    assert pc.tolist() == [0xFFFFFFFF, 0, 0, 2]


def test_multi_seed_equals_union_of_sources(orc):
    rng = np.random.default_rng(3)
    parts, off = [], 0
    srcs = []
    for i in range(4):
        n = int(rng.integers(5, 30))
        s, d, w = random_graph(rng, n, 3 * n, zero_w=False)
        parts.append((s + off, d + off, w))
        srcs.append(off + int(rng.integers(0, n)))
        off += n
    src = np.concatenate([p[0] for p in parts])
    dst = np.concatenate([p[1] for p in parts])
    w = np.concatenate([p[2] for p in parts])
    row, col, wo = orc.build_csr(off, src, dst, w)
    d, pc = orc.dijkstra_multi_seed(row, col, wo, srcs)
    for s in srcs:
        ds, _, ps, _, _ = orc.dijkstra(row, col, wo, s)
        reach = ds != 0xFFFFFFFF
        assert np.array_equal(d[reach], ds[reach])
        assert np.array_equal(pc[reach], ps[reach])


def test_multi_source_matches_single(orc):
    rng = np.random.default_rng(4)
    n = 200
    s, dd, w = random_graph(rng, n, 1000, zero_w=False)
    row, col, wo = orc.build_csr(n, s, dd, w)
    srcs = rng.integers(0, n, 7)
    dist, edges, thr = orc.multi_source(row, col, wo, srcs, threads=3)
    tot = 0
    for i, q in enumerate(srcs):
        d, _, _, _, e = orc.dijkstra(row, col, wo, int(q))
        assert np.array_equal(dist[i], d)
        tot += e
    assert edges == tot and thr == 3


# ---------------------------------------------------------------- heuristic / A*
def test_isqrt_heuristic_closed_form(orc):
    rng = np.random.default_rng(11)
    x = rng.integers(-2_000_000, 2_000_000, 300).astype(np.float64)
    y = rng.integers(-2_000_000, 2_000_000, 300).astype(np.float64)
    for v in range(300):
        t = (v * 37) % 300
        dx, dy = int(x[v] - x[t]), int(y[v] - y[t])
        assert orc.heuristic(orc.H_FLOOR_EUCLID, x, y, 0.0, v, t) == math.isqrt(dx * dx + dy * dy)
    # perfect squares and their neighbours
    xs = np.array([0.0, 3.0, 3.0, 300000.0], np.float64)
    ys = np.array([0.0, 4.0, 5.0, 400000.0], np.float64)
    assert orc.heuristic(orc.H_FLOOR_EUCLID, xs, ys, 0.0, 1, 0) == 5
    assert orc.heuristic(orc.H_FLOOR_EUCLID, xs, ys, 0.0, 2, 0) == 5      # sqrt(34) = 5.83
    assert orc.heuristic(orc.H_FLOOR_EUCLID, xs, ys, 0.0, 3, 0) == 500000


def test_scaled_heuristic_formula(orc):
    rng = np.random.default_rng(12)
    x = rng.uniform(-1e5, 1e5, 200)
    y = rng.uniform(-1e5, 1e5, 200)
    for v in range(200):
        t = (v * 7 + 3) % 200
        e = math.sqrt((x[v] - x[t]) ** 2 + (y[v] - y[t]) ** 2) * 1.137 * (1.0 - 1e-9)
        want = 0 if not e > 0 else (4000000000 if e >= 4e9 else int(e))
        assert orc.heuristic(orc.H_SCALED, x, y, 1.137, v, t) == want


def _grid(orc, side, seed):
    import pathfinding_gpu_amd as P
    g = P.grid_road(side, side, seed)
    return g


@pytest.mark.parametrize("hkind", [0, 1, 2])
def test_astar_canonical_path_is_tree_path(orc, hkind):
    g = _grid(orc, 30, 5)
    row, col, w = g.row_offsets, g.col_indices, g.weights
    rng = np.random.default_rng(hkind)
    for _ in range(10):
        s, t = (int(v) for v in rng.integers(0, g.n, 2))
        d, _, pc, _, _ = orc.dijkstra(row, col, w, s)
        da, path, settled, edges = orc.astar(row, col, w, g.x, g.y, g.h_scale, hkind, s, t, canonical=True)
        assert da == d[t]
        assert np.array_equal(path, orc.reconstruct_path(pc, s, t))
        assert orc.validate_path(row, col, w, path, s, t, int(d[t]))
        dr, _, _, _ = orc.astar(row, col, w, g.x, g.y, g.h_scale, hkind, s, t, canonical=False)
        assert dr == d[t]


def test_floor_euclid_is_consistent_on_grid(orc):
    g = _grid(orc, 40, 7)
    row, col, w = g.row_offsets, g.col_indices, g.weights
    for t in (0, 555, g.n - 1):
        h = orc.heuristic_all(orc.H_FLOOR_EUCLID, g.x, g.y, 0.0, t)
        us = np.repeat(np.arange(g.n), np.diff(row))
        assert (h[us].astype(np.int64) <= w.astype(np.int64) + h[col]).all()   # consistent
        assert h[t] == 0


# ---------------------------------------------------------------- small pieces
def test_light_end_vs_searchsorted(orc):
    rng = np.random.default_rng(2)
    n = 60
    s, d, w = random_graph(rng, n, 500, wmax=50, zero_w=False)
    row, col, wo = orc.build_csr(n, s, d, w)
    for delta in (1, 7, 25, 50, 1000):
        le = orc.light_end(row, wo, delta)
        want = [row[u] + np.searchsorted(wo[row[u]:row[u + 1]], delta, side="right") for u in range(n)]
        assert le.tolist() == [int(v) for v in want]


def test_auto_delta_closed_form(orc):
    for aw, ad in [(127.46, 3.96), (128.0, 16.0), (1.0, 0.2), (0.01, 5.0), (1000.0, 1.0)]:
        assert orc.auto_delta(aw, ad) == max(1, int(math.floor(32.0 * aw / max(1.0, ad) + 0.5)))


def test_exclusive_scan_and_bounds(orc):
    rng = np.random.default_rng(1)
    a = rng.integers(0, 3, 1000).astype(np.uint32)
    ex, tot = orc.exclusive_scan(a)
    assert tot == int(a.sum())
    assert np.array_equal(ex, (np.cumsum(a) - a).astype(np.uint32))
    assert orc.next_bound(0, 100) == 100 and orc.next_bound(99, 100) == 100 and orc.next_bound(100, 100) == 200


def test_verify_detects_planted_mistakes(orc):
    g = _grid(orc, 20, 3)
    row, col, w = g.row_offsets, g.col_indices, g.weights
    d, pf, pc, _, _ = orc.dijkstra(row, col, w, 0)
    ok = orc.verify_sssp(row, col, w, 0, d, d, pc, pc)
    assert sum(ok["errors"].values()) == 0
    bad_d = d.copy()
    bad_d[57] += 1
    assert orc.verify_sssp(row, col, w, 0, d, bad_d, pc, pc)["errors"]["dist"] == 1
    bad_p = pc.copy()
    bad_p[57] = (bad_p[57] + 1) % g.n
    r = orc.verify_sssp(row, col, w, 0, d, d, bad_p, pc, samples=400)
    assert r["errors"]["tree"] == 1 and r["errors"]["parent"] == 1


def test_reachable_arcs(orc):
    row, col, w = orc.build_csr(5, [0, 1, 3, 3], [1, 2, 4, 0], [1, 1, 1, 1])
    d, *_ = orc.dijkstra(row, col, w, 0)
    assert orc.reachable_arcs(row, d) == (2, 3)
