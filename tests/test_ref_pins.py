"""Pins the oracle to the reference's own CPU code.

* golden: tests/golden/ref_vectors.npz, written by tests/golden/make_golden.py
  from oracle/_ref (the reference's src/cpu_sssp.cpp, graph.cpp, verify.cpp
  compiled in place).  Always runs.
* live: oracle/_ref itself on fresh random graphs.  Skipped where the
  reference sources were not available to build it.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import random_graph

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")


def _cases():
    z = np.load(GOLD, allow_pickle=False)
    return z, [str(c) for c in z["cases"]]


def test_golden_csr_and_stats(orc):
    z, cases = _cases()
    for c in cases:
        n = int(z[f"{c}/n"][0])
        row, col, w = orc.build_csr(n, z[f"{c}/src"], z[f"{c}/dst"], z[f"{c}/w"])
        assert np.array_equal(row, z[f"{c}/ref_row"]), c
        assert np.array_equal(col, z[f"{c}/ref_col"]), c
        assert np.array_equal(w, z[f"{c}/ref_wt"]), c
        x = z[f"{c}/x"] if f"{c}/x" in z else None
        y = z[f"{c}/y"] if f"{c}/y" in z else None
        s = orc.csr_stats(row, col, w, x, y)
        assert [s["minW"], s["maxW"], s["maxDeg"]] == z[f"{c}/ref_istats"].tolist(), c
        assert [s["avgW"], s["avgDeg"], s["hScale"]] == z[f"{c}/ref_dstats"].tolist(), c


def test_golden_dijkstra_exact(orc):
    z, cases = _cases()
    for c in cases:
        row, col, w = z[f"{c}/ref_row"], z[f"{c}/ref_col"], z[f"{c}/ref_wt"]
        for i, s in enumerate(z[f"{c}/sources"]):
            d, pf, pc, settled, edges = orc.dijkstra(row, col, w, int(s))
            assert np.array_equal(d, z[f"{c}/ref_dist"][i]), c
            assert np.array_equal(pf, z[f"{c}/ref_parent"][i]), c          # cpuDijkstra's own rule
            assert [settled, edges] == z[f"{c}/ref_dijkstra_stats"][i].tolist(), c
            # the min-index rule differs from the reference's only on ties, and is a valid tree
            rep = orc.verify_sssp(row, col, w, int(s), d, d, pc, None)
            assert sum(rep["errors"].values()) == 0
        md = z[f"{c}/ref_multi_dist"]
        dist, _, _ = orc.multi_source(row, col, w, z[f"{c}/sources"], threads=2)
        assert np.array_equal(dist, md), c


def test_golden_astar_exact(orc):
    z, cases = _cases()
    for c in cases:
        if f"{c}/pairs" not in z:
            continue
        row, col, w = z[f"{c}/ref_row"], z[f"{c}/ref_col"], z[f"{c}/ref_wt"]
        x, y = z[f"{c}/x"], z[f"{c}/y"]
        hs = float(z[f"{c}/ref_dstats"][2])
        j = 0
        for (s, t) in z[f"{c}/pairs"]:
            for useh in (1, 0):
                kind = orc.H_SCALED if useh else orc.H_NONE
                d, path, settled, edges = orc.astar(row, col, w, x, y, hs, kind, int(s), int(t), canonical=False)
                L = int(z[f"{c}/ref_astar_len"][j])
                assert d == int(z[f"{c}/ref_astar_dist"][j]), c
                assert np.array_equal(path, z[f"{c}/ref_astar_path"][j][:L]), c
                assert [settled, edges] == z[f"{c}/ref_astar_stats"][j].tolist(), c
                j += 1
        for i, (s, t) in enumerate(z[f"{c}/pairs"]):
            h = orc.heuristic_all(orc.H_SCALED, x, y, hs, int(t))
            assert np.array_equal(h, z[f"{c}/ref_h"][i]), c


# ------------------------------------------------------------------ live reference
@pytest.fixture(scope="module")
def ref(orc):
    r = orc.ref()
    if r is None:
        pytest.skip("oracle/_ref not built (reference sources absent on this machine)")
    return r


def _p(a, t=C.c_uint32):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


@pytest.mark.parametrize("seed", range(8))
def test_live_reference_dijkstra_astar(orc, ref, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(20, 300))
    src, dst, w = random_graph(rng, n, int(rng.integers(n, 6 * n)), wmax=int(rng.integers(1, 9)))
    x = rng.integers(0, 50, n).astype(np.float64)
    y = rng.integers(0, 50, n).astype(np.float64)
    g = ref.ref_build_csr(n, len(src), _p(src), _p(dst), _p(w), _p(x, C.c_double), _p(y, C.c_double))
    try:
        m = ref.ref_m(g)
        row, col, wo = orc.build_csr(n, src, dst, w)
        rr, rc, rw = np.empty(n + 1, np.uint32), np.empty(m, np.uint32), np.empty(m, np.uint32)
        st, ds = np.empty(3, np.uint32), np.empty(3, np.float64)
        ref.ref_csr(g, _p(rr), _p(rc), _p(rw), _p(st), _p(ds, C.c_double))
        assert np.array_equal(row, rr) and np.array_equal(col, rc) and np.array_equal(wo, rw)
        for s in rng.integers(0, n, 4):
            d, pf, pc, settled, edges = orc.dijkstra(row, col, wo, int(s))
            rd, rp, rs = np.empty(n, np.uint32), np.empty(n, np.uint32), np.empty(2, np.uint64)
            ref.ref_dijkstra(g, int(s), _p(rd), _p(rp), _p(rs, C.c_uint64))
            assert np.array_equal(d, rd) and np.array_equal(pf, rp)
            assert [settled, edges] == rs.tolist()
            t = int(rng.integers(0, n))
            path = np.empty(n + 1, np.uint32)
            ln = C.c_uint32()
            ss = np.empty(2, np.uint64)
            msv = C.c_double()
            rdist = ref.ref_astar(g, int(s), t, 1, _p(path), n + 1, C.byref(ln), _p(ss, C.c_uint64), C.byref(msv))
            od, opath, osettled, oedges = orc.astar(row, col, wo, x, y, float(ds[2]), orc.H_SCALED, int(s), t,
                                                    canonical=False)
            assert od == rdist and np.array_equal(opath, path[: ln.value])
            assert [osettled, oedges] == ss.tolist()
            # the reference's verifier accepts the oracle's min-index parents
            out = np.zeros(6, np.uint64)
            ref.ref_verify_sssp(g, int(s), _p(rd), _p(d), _p(pc), 32, 7, _p(out, C.c_uint64))
            assert out[:3].tolist() == [0, 0, 0]
    finally:
        ref.ref_free(g)


def test_live_reference_reconstruct_path(orc, ref):
    rng = np.random.default_rng(77)
    for _ in range(20):
        n = int(rng.integers(2, 50))
        par = rng.integers(0, n + 3, n).astype(np.uint32)
        par[rng.integers(0, n)] = 0xFFFFFFFF
        s, t = (int(v) for v in rng.integers(0, n, 2))
        want = np.empty(n + 1, np.uint32)
        L = ref.ref_reconstruct_path(_p(par), n, s, t, _p(want), n + 1)
        assert np.array_equal(orc.reconstruct_path(par, s, t), want[:L])
