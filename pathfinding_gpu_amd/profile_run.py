"""Fixed workload for hardware-counter profiles (rocprofv3 --pmc ... -- python3 -m pathfinding_gpu_amd.profile_run).

Runs, in order: one streaming copy of 1 GiB (counter calibration: 1 GiB read, 1 GiB
written), then `--steps` batched runs of K=64 single-source Delta-stepping on the
1000x1000 grid road graph (seed 7; vertex 0 + 63 random sources, persistent
kernel, Delta = 1019, hybrid mapping) -- the benchmark's step -- and prints the
reach statistics that give the step kernel's algorithmic bytes.
"""
import argparse
import ctypes as C
import json

import numpy as np

from . import _native as N
from .graph import grid_road
from .solver import Drive, GpuOptions, GpuSolver, Mapping


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--mapping", type=int, default=2)
    a = ap.parse_args()
    ms = C.c_double()
    N.check(N.engine().pf_debug_copy(1 << 30, 1, C.byref(ms)))
    g = grid_road(1000, 1000, 7)
    rng = np.random.default_rng(2026)
    cand = np.flatnonzero(np.diff(g.row_offsets) > 0)
    srcs = np.concatenate([[0], rng.choice(cand[cand != 0], size=a.batch - 1, replace=False)]).astype(np.uint32)
    sv = GpuSolver(g, max_batch=a.batch)
    o = GpuOptions(drive=Drive.Persistent, delta=sv.auto_delta(), mapping=Mapping(a.mapping))
    rs = [sv.multi_source(srcs, o, copy_back=False) for _ in range(a.steps)]
    reached, arcs = sv.last_reach()
    print(json.dumps({"copy_ms_1GiB": ms.value, "copy_GBps": 2 * 2**30 / (ms.value * 1e-3) / 1e9,
                      "reached": reached, "arcs": arcs, "bytes_alg_per_launch": 16 * reached + 16 * arcs,
                      "loop_ms": [r.stats.loop_ms for r in rs], "edges_scanned": rs[-1].stats.edges_scanned,
                      "updates": rs[-1].stats.updates, "atomics": rs[-1].stats.atomics_issued}))
    sv.close()


if __name__ == "__main__":
    main()
