"""Host graph layer: CSR container, benchmark workload generators, union, cache.

Thin wrappers over libpfgraph.so (include/pf_graph.h).  The CSR semantics are
those of the reference's buildCSR (src/graph.cpp:23-104).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _native as N


@dataclass
class CSRGraph:
    """CSR graph; field meaning follows the reference's CSRGraph (include/graph.h:24-48)."""

    n: int
    row_offsets: np.ndarray           # uint32[n+1]
    col_indices: np.ndarray           # uint32[m], rows sorted by (w, dst)
    weights: np.ndarray               # uint32[m]
    x: Optional[np.ndarray] = None    # float64[n]
    y: Optional[np.ndarray] = None
    name: str = ""
    min_w: int = 0
    max_w: int = 0
    max_deg: int = 0
    avg_w: float = 0.0
    avg_deg: float = 0.0
    h_scale: float = 0.0

    @property
    def m(self) -> int:
        return int(self.col_indices.shape[0])

    def has_coords(self) -> bool:
        return self.x is not None

    def degree(self, u: int) -> int:
        return int(self.row_offsets[u + 1] - self.row_offsets[u])

    def bytes_csr(self) -> int:
        return (self.n + 1) * 4 + self.m * 8

    def max_degree_vertex(self) -> int:
        return int(N.graphlib().pf_max_degree_vertex(self.n, N.ptr(self.row_offsets)))


def _stats(g: CSRGraph) -> CSRGraph:
    iout = np.zeros(3, np.uint32)
    dout = np.zeros(3, np.float64)
    N.graphlib().pf_csr_stats(g.n, N.ptr(g.row_offsets), N.ptr(g.col_indices), N.ptr(g.weights),
                              N.ptr(g.x, C.c_double), N.ptr(g.y, C.c_double), N.ptr(iout), N.ptr(dout, C.c_double))
    g.min_w, g.max_w, g.max_deg = (int(v) for v in iout)
    g.avg_w, g.avg_deg, g.h_scale = (float(v) for v in dout)
    return g


def build_csr(n: int, src: np.ndarray, dst: np.ndarray, w: np.ndarray, x=None, y=None, name: str = "") -> CSRGraph:
    """buildCSR (src/graph.cpp:23-104): drop self-loops, w>=1, rows sorted by (w, dst), statistics."""
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    w = np.ascontiguousarray(w, np.uint32)
    m_in = src.shape[0]
    row = np.zeros(n + 1, np.uint32)
    col = np.empty(max(m_in, 1), np.uint32)
    wo = np.empty(max(m_in, 1), np.uint32)
    m_out = C.c_uint32(0)
    rc = N.graphlib().pf_build_csr(n, m_in, N.ptr(src), N.ptr(dst), N.ptr(w), N.ptr(row), N.ptr(col), N.ptr(wo),
                                   C.byref(m_out))
    if rc != 0:
        raise ValueError(f"pf_build_csr failed ({rc})")
    m = m_out.value
    g = CSRGraph(n=n, row_offsets=row, col_indices=col[:m].copy(), weights=wo[:m].copy(),
                 x=None if x is None else np.ascontiguousarray(x, np.float64),
                 y=None if y is None else np.ascontiguousarray(y, np.float64), name=name)
    return _stats(g)


def grid_road_edges(rows: int, cols: int, seed: int, jitter: int = 20, wlo: float = 1.0, whi: float = 1.5):
    m = int(N.graphlib().pf_grid_road_arcs(rows, cols))
    n = rows * cols
    src = np.empty(m, np.uint32)
    dst = np.empty(m, np.uint32)
    w = np.empty(m, np.uint32)
    x = np.empty(n, np.float64)
    y = np.empty(n, np.float64)
    rc = N.graphlib().pf_gen_grid_road(rows, cols, seed, jitter, wlo, whi, N.ptr(src), N.ptr(dst), N.ptr(w),
                                       N.ptr(x, C.c_double), N.ptr(y, C.c_double))
    if rc != 0:
        raise ValueError("pf_gen_grid_road failed")
    return n, src, dst, w, x, y


def grid_road(rows: int, cols: int, seed: int, jitter: int = 20, wlo: float = 1.0, whi: float = 1.5) -> CSRGraph:
    """Grid road graph: 4-neighbour roads, integer jitter, w = ceil(len * U[wlo, whi))."""
    n, src, dst, w, x, y = grid_road_edges(rows, cols, seed, jitter, wlo, whi)
    return build_csr(n, src, dst, w, x, y, name=f"grid_{rows}x{cols}_s{seed}")


def rmat_edges(scale: int, ef: int, seed: int, a=0.57, b=0.19, c=0.19, wmax: int = 255):
    m = ef << scale
    src = np.empty(m, np.uint32)
    dst = np.empty(m, np.uint32)
    w = np.empty(m, np.uint32)
    rc = N.graphlib().pf_gen_rmat(scale, ef, a, b, c, wmax, seed, N.ptr(src), N.ptr(dst), N.ptr(w))
    if rc != 0:
        raise ValueError("pf_gen_rmat failed")
    return 1 << scale, src, dst, w


def rmat(scale: int, ef: int, seed: int, a=0.57, b=0.19, c=0.19, wmax: int = 255) -> CSRGraph:
    """Directed R-MAT, self-loops dropped, duplicates kept, w uniform in [1, wmax]."""
    n, src, dst, w = rmat_edges(scale, ef, seed, a, b, c, wmax)
    return build_csr(n, src, dst, w, name=f"rmat_s{scale}_ef{ef}_s{seed}")


def union(graphs: Sequence[CSRGraph]):
    """Disjoint union; returns (graph, vertex offsets[count+1])."""
    cnt = len(graphs)
    ns = np.array([g.n for g in graphs], np.uint32)
    ms = np.array([g.m for g in graphs], np.uint32)
    tn = int(ns.sum(dtype=np.uint64))
    tm = int(ms.sum(dtype=np.uint64))
    rows = (N.u32p * cnt)(*[N.ptr(g.row_offsets) for g in graphs])
    cols = (N.u32p * cnt)(*[N.ptr(g.col_indices) for g in graphs])
    ws = (N.u32p * cnt)(*[N.ptr(g.weights) for g in graphs])
    voff = np.zeros(cnt + 1, np.uint32)
    row = np.zeros(tn + 1, np.uint32)
    col = np.empty(max(tm, 1), np.uint32)
    w = np.empty(max(tm, 1), np.uint32)
    rc = N.graphlib().pf_csr_union(cnt, N.ptr(ns), N.ptr(ms), rows, cols, ws, N.ptr(voff), N.ptr(row), N.ptr(col),
                                   N.ptr(w))
    if rc != 0:
        raise ValueError("pf_csr_union failed (sizes exceed 2^32)")
    g = CSRGraph(n=tn, row_offsets=row, col_indices=col[:tm], weights=w[:tm], name=f"union_{cnt}")
    return _stats(g), voff


def save_binary(g: CSRGraph, path: str) -> None:
    rc = N.graphlib().pf_csr_save(path.encode(), g.name.encode(), g.n, g.m, N.ptr(g.row_offsets),
                                  N.ptr(g.col_indices), N.ptr(g.weights), N.ptr(g.x, C.c_double),
                                  N.ptr(g.y, C.c_double))
    if rc != 0:
        raise OSError(f"cannot write {path}")


def load_binary(path: str) -> CSRGraph:
    n, m, xy = C.c_uint32(), C.c_uint32(), C.c_uint32()
    if N.graphlib().pf_csr_peek(path.encode(), C.byref(n), C.byref(m), C.byref(xy)) != 0:
        raise OSError(f"not a PFCSR001 file: {path}")
    row = np.empty(n.value + 1, np.uint32)
    col = np.empty(max(m.value, 1), np.uint32)
    w = np.empty(max(m.value, 1), np.uint32)
    x = np.empty(n.value, np.float64) if xy.value else None
    y = np.empty(n.value, np.float64) if xy.value else None
    if N.graphlib().pf_csr_load(path.encode(), N.ptr(row), N.ptr(col), N.ptr(w), N.ptr(x, C.c_double),
                                N.ptr(y, C.c_double)) != 0:
        raise OSError(f"truncated PFCSR001 file: {path}")
    g = CSRGraph(n=n.value, row_offsets=row, col_indices=col[:m.value], weights=w[:m.value], x=x, y=y, name=path)
    return _stats(g)
