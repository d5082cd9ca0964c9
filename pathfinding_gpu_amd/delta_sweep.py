"""Delta sweep on synthetic graphs -> training data of the learned-Delta network.

Counterpart of the reference's runDeltaStudy (src/main.cpp:560-608), which times
GpuSolver::sssp over a range of Delta values on one graph.  Here the sweep runs
over a family of synthetic grid road graphs and R-MAT graphs (none of them is a
benchmark workload: the 1000x1000 grid with seed 7 and the scale-20 R-MAT with
seed 11 are held out), for single queries and batches, and records for every
graph the time of each candidate Delta (multiples of the Davidson default, hybrid
mapping) and of each mapping (at the Davidson Delta).  Targets are log2(time /
best time of that graph and batch size).

    python -m pathfinding_gpu_amd.delta_sweep OUT.npz [--quick]
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import graph as G
from .learned import DELTA_MULTS, FEATURES, MAPPINGS, N_OUT, davidson_delta, graph_features
from .solver import Drive, GpuOptions, GpuSolver


def family(quick: bool = False):
    """(kind, params) of the training graphs."""
    out = []
    sides = (120, 300) if quick else (150, 300, 500, 700, 900)
    for i, side in enumerate(sides):
        for j, (lo, hi) in enumerate(((1.0, 1.5), (1.0, 4.0))):
            out.append(("grid", dict(rows=side, cols=side, seed=100 + 10 * i + j, jitter=20, wlo=lo, whi=hi)))
    if not quick:
        out.append(("grid", dict(rows=400, cols=1000, seed=150, jitter=40, wlo=1.0, whi=1.5)))
        out.append(("grid", dict(rows=800, cols=800, seed=151, jitter=0, wlo=1.0, whi=2.0)))
    scales = (10, 12) if quick else (12, 14, 16, 17, 18)
    efs = (8,) if quick else (4, 8, 16, 32)
    for s in scales:
        for ef in efs:
            if s >= 17 and ef >= 32:
                continue
            out.append(("rmat", dict(scale=s, ef=ef, seed=200 + s * 10 + ef, wmax=255)))
    if not quick:
        for s, ef, wmax in ((14, 8, 15), (16, 16, 4095), (18, 8, 4095), (16, 8, 63)):
            out.append(("rmat", dict(scale=s, ef=ef, seed=300 + s, wmax=wmax)))
        for s in (14, 16, 18):   # near-uniform degree (Erdos-Renyi-like R-MAT)
            out.append(("rmat", dict(scale=s, ef=8, seed=400 + s, wmax=255, a=0.25, b=0.25, c=0.25)))
    return out


def build(kind: str, p: dict):
    if kind == "grid":
        return G.grid_road(p["rows"], p["cols"], p["seed"], p["jitter"], p["wlo"], p["whi"])
    return G.rmat(p["scale"], p["ef"], p["seed"], p.get("a", 0.57), p.get("b", 0.19), p.get("c", 0.19), p["wmax"])


def time_run(sv: GpuSolver, srcs, opt: GpuOptions, reps: int = 2) -> float:
    best = float("inf")
    for _ in range(reps):
        if len(srcs) == 1:
            r = sv.sssp(int(srcs[0]), opt, copy_back=False)
        else:
            r = sv.multi_source(srcs, opt, copy_back=False)
        best = min(best, r.ms)
    return best


def sweep(out_path: str, quick: bool = False, batches=(1, 32), log=print):
    X, Y, M, names = [], [], [], []
    rng = np.random.default_rng(5)
    for kind, p in family(quick):
        t0 = time.time()
        g = build(kind, p)
        kmax = max(batches)
        sv = GpuSolver(g, max_batch=kmax)
        auto = davidson_delta(g.avg_w, g.avg_deg)
        deg = np.diff(g.row_offsets)
        cand = np.flatnonzero(deg > 0)
        for k in batches:
            srcs = rng.choice(cand, size=k, replace=False) if k > 1 else [g.max_degree_vertex() if kind == "rmat" else 0]
            base = GpuOptions(drive=Drive.Persistent)
            time_run(sv, srcs, GpuOptions(drive=Drive.Persistent, delta=auto), reps=1)   # warm-up
            t = np.full(N_OUT, np.nan)
            for i, mlt in enumerate(DELTA_MULTS):
                d = max(1, int(round(auto * mlt)))
                t[i] = time_run(sv, srcs, GpuOptions(drive=Drive.Persistent, delta=d))
            for j, mp in enumerate(MAPPINGS):
                t[len(DELTA_MULTS) + j] = time_run(sv, srcs, GpuOptions(drive=Drive.Persistent, delta=auto, mapping=mp))
            best = np.nanmin(t)
            X.append(graph_features(g, k))
            Y.append(np.log2(t / best))
            M.append(np.isfinite(t).astype(np.float64))
            names.append(f"{g.name}|k{k}")
            log(f"[sweep] {g.name} k={k} auto={auto} best={best:.3f}ms "
                f"mults={' '.join(f'{v:.2f}' for v in np.log2(t / best)[:len(DELTA_MULTS)])} "
                f"maps={' '.join(f'{v:.2f}' for v in np.log2(t / best)[len(DELTA_MULTS):])}")
        sv.close()
        log(f"[sweep]   ({time.time() - t0:.1f}s)")
    np.savez(out_path, X=np.array(X), Y=np.nan_to_num(np.array(Y)), M=np.array(M), names=np.array(names),
             features=np.array(FEATURES), mults=DELTA_MULTS)
    return out_path


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--quick", action="store_true")
    a = ap.parse_args(argv)
    sweep(a.out, a.quick, log=lambda s: print(s, flush=True))


if __name__ == "__main__":
    main(sys.argv[1:])
