"""Benchmark of the SSSP / A* engine on one MI355X.

Workload (the headline configuration): 1000x1000 grid road graph, seed 7
(1,000,000 vertices, 3,996,000 arcs); Delta-stepping SSSP from vertex 0, plus
a batch of 256 point-to-point A* queries with uniform random endpoints
(h = floor(Euclidean distance)).

One step = one SSSP from vertex 0 on the tuned, device-resident engine
(persistent kernel, tuned Delta).  `value` is its GTEPS: reachable arcs
(the reference's TEPS numerator, src/main.cpp:209-213) / wall time of the step,
with the graph already resident in HBM and no copy back.  Reported beside it:
the plain path (Davidson default Delta, host-driven loop, sync every 8
launches, as the reference), the 256-query A* batch (queries/s and per-query
latency), the roofline of the step kernel, and the reference's cpuDijkstra
timed on this host.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--side", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--delta", type=int, default=0, help="Delta of the tuned path (0 = tuned table / default)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import pathfinding_gpu_amd as P

    t0 = time.time()
    g = P.grid_road(args.side, args.side, args.seed)
    log(f"[bench] graph {g.name}: n={g.n} m={g.m} avgW={g.avg_w:.2f} built in {time.time() - t0:.1f}s")
    K = args.queries
    sv = P.GpuSolver(g, max_batch=max(1, K))
    info = P.device_info(0)
    log(f"[bench] device {info['name']} {info['arch']} CUs={info['cu_count']} upload {sv.upload_ms():.1f} ms")
    auto = sv.auto_delta()

    # ---- tuned device-resident path ----
    delta = args.delta or 4 * auto
    tuned = P.GpuOptions(algo=P.GpuAlgo.DeltaStep, drive=P.Drive.Persistent, delta=delta)
    plain = P.GpuOptions(algo=P.GpuAlgo.DeltaStep, drive=P.Drive.Host, sync_every=8)   # reference defaults

    def run_steps(opt, steps):
        ts, st = [], None
        for _ in range(steps):
            a = time.perf_counter()
            r = sv.sssp(0, opt, copy_back=False)
            ts.append((time.perf_counter() - a) * 1e3)
            st = r
        return ts, st

    log("[bench] warmup")
    run_steps(tuned, args.warmup)
    run_steps(plain, max(1, args.warmup))
    log("[bench] timed: tuned path")
    t_tuned, r_tuned = run_steps(tuned, args.steps)
    reached, arcs = sv.last_reach()
    log("[bench] timed: plain path")
    t_plain, r_plain = run_steps(plain, args.steps)

    ms_step = float(np.mean(t_tuned))
    gteps = arcs / (ms_step * 1e-3) / 1e9
    ms_plain = float(np.mean(t_plain))
    gteps_plain = arcs / (ms_plain * 1e-3) / 1e9

    # roofline of the step kernel (the persistent launch): algorithmic bytes per launch = 16 per reached
    # vertex (row extent + own state) + 16 per scanned arc (col, w + target state), over its hipEvent duration
    loop_ms = r_tuned.stats.loop_ms
    bytes_launch = 16 * reached + 16 * arcs
    achieved = bytes_launch / (loop_ms * 1e-3) / 1e9
    b_alg = 8 * g.n + bytes_launch

    # ---- A* batch ----
    rng = np.random.default_rng(2026)
    srcs = rng.integers(0, g.n, K).astype(np.uint32)
    dsts = rng.integers(0, g.n, K).astype(np.uint32)
    aopt = P.GpuOptions(algo=P.GpuAlgo.DeltaStep, drive=P.Drive.Persistent, delta=delta,
                        heuristic=P.Heuristic.FloorEuclid)
    log("[bench] A* batch warmup")
    sv.astar_batch(srcs, dsts, aopt, path_cap=8192)
    a_ts = []
    for _ in range(max(1, args.steps // 2)):
        a = time.perf_counter()
        res = sv.astar_batch(srcs, dsts, aopt, path_cap=8192)
        a_ts.append((time.perf_counter() - a) * 1e3)
    a_ms = float(np.mean(a_ts))
    qps = K / (a_ms * 1e-3)

    out = {
        "metric": "GTEPS (SSSP, 1000x1000 grid road graph, tuned device-resident Delta-stepping) "
                  "+ A* queries/s and latency, bit-exact",
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"grid{args.side}x{args.side}_seed{args.seed}: sssp from 0 + {K} A* queries",
                   "n": g.n, "m": g.m, "delta": delta, "auto_delta": auto, "drive": "persistent",
                   "astar_heuristic": "floor_euclid"},
        "plain_gteps": round(gteps_plain, 4),
        "plain_ms": round(ms_plain, 4),
        "speedup_vs_plain": round(ms_plain / ms_step, 3),
        "plain_stats": {"iterations": r_plain.stats.iterations, "host_syncs": r_plain.stats.host_syncs,
                        "launches": r_plain.stats.launches, "buckets": r_plain.stats.buckets},
        "tuned_stats": {"iterations": r_tuned.stats.iterations, "steps": r_tuned.stats.steps,
                        "buckets": r_tuned.stats.buckets, "edges_scanned": r_tuned.stats.edges_scanned,
                        "atomics_per_update": round(r_tuned.stats.atomics_issued / max(1, r_tuned.stats.updates), 4),
                        "event_ms": round(r_tuned.ms, 4), "loop_ms": round(loop_ms, 4)},
        "astar": {"queries": K, "batch_ms": round(a_ms, 3), "queries_per_s": round(qps, 1),
                  "p50_ms": round(float(np.percentile(a_ts, 50)), 3),
                  "p99_ms": round(float(np.percentile(a_ts, 99)), 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "bytes_per_launch": int(bytes_launch), "b_alg_per_sssp": int(b_alg)},
    }

    if not args.no_cpu:
        log("[bench] CPU baseline")
        import oracle
        kind = "reference" if oracle.ref() is not None else "port"
        row, col, w = g.row_offsets, g.col_indices, g.weights
        if kind == "reference":
            import ctypes as C
            R = oracle.ref()
            xs = None
            gh = R.ref_build_csr(g.n, g.m, *(a.ctypes.data_as(C.POINTER(C.c_uint32))
                                              for a in (np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(row)),
                                                        col, w)), xs, xs)
            dist = np.empty(g.n, np.uint32)
            par = np.empty(g.n, np.uint32)
            reps, tot = 0, 0.0
            while tot < args.cpu_seconds * 1e3 and reps < 200:
                tot += R.ref_dijkstra(gh, 0, dist.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      par.ctypes.data_as(C.POINTER(C.c_uint32)), None)
                reps += 1
            R.ref_free(gh)
        else:
            reps = max(1, int(args.cpu_seconds * 1e3 / max(1.0, oracle.time_dijkstra(row, col, w, 0, 1))))
            tot = oracle.time_dijkstra(row, col, w, 0, reps)
            dist = oracle.dijkstra(row, col, w, 0)[0]
        cpu_ms = tot / reps
        chk = sv.sssp(0, tuned)
        out["bitexact_dist_vs_cpuDijkstra"] = bool(np.array_equal(chk.dist, dist))
        out["cpu_baseline"] = {"value": round(arcs / (cpu_ms * 1e-3) / 1e9, 5), "unit": "GTEPS", "cores": 1,
                               "kind": kind, "ms_per_sssp": round(cpu_ms, 3),
                               "sample": f"cpuDijkstra from vertex 0 on the same graph, {reps} runs",
                               "host_threads": os.cpu_count()}
    sv.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
