"""Benchmark of the SSSP / A* engine on one MI355X.

Workload (default, --workload grid1000): the 1000x1000 grid road graph, seed 7
(1,000,000 vertices, 3,996,000 arcs, integer jittered coordinates,
w = ceil(len * U[1, 1.5))), Delta-stepping from vertex 0, plus 256 point-to-point
A* queries with uniform random endpoints (h = floor(Euclidean distance)).
--workload rmat20 runs the scale-20 R-MAT (edge factor 16, seed 11, source = the
highest-degree vertex) the same way, without A*.

One step = one batched run of K = 64 single-source shortest-path queries
(vertex 0 and 63 random sources) in one persistent launch, with the Delta and
mapping the learned network picks for this graph and batch size (judged against
the Davidson default).  `value` = GTEPS of that step: reachable arcs summed over
the K queries (the reference's TEPS numerator, src/main.cpp:209-213) / wall time,
graph resident in HBM, no copy back.  Reported beside it:
  * plain: one query from vertex 0 with the Davidson default Delta (persistent
    kernel), and the reference-style host-driven loop (sync every 8 launches);
  * learned single-query: one query from vertex 0 with the learned Delta;
  * A*: the 256-query batch (queries/s, latency) and plain one-at-a-time A*;
  * roofline of the step kernel; the reference's cpuDijkstra (1 core) and
    cpuMultiSource (all cores) on this host.

    python bench.py [--steps K] [--warmup W] [--workload grid1000|rmat20]
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
TUNE_DB = os.path.join(ROOT, "pathfinding_gpu_amd", "data", "autotune.json")
TRAFFIC_PROFILES = {"grid1000": os.path.join(ROOT, "profiles", "stepkernel_grid1000_k64_hbm_traffic.md"),
                    "rmat20": os.path.join(ROOT, "profiles", "stepkernel_rmat20_k64_hbm_traffic.md")}


def measured_traffic(opts: dict, workload: str = "grid1000"):
    """HBM bytes per launch of the step kernel from the committed rocprofv3 counter summary,
    if it was measured with exactly these launch options (else None)."""
    try:
        for line in open(TRAFFIC_PROFILES[workload]):
            if line.startswith("traffic: "):
                rec = json.loads(line[len("traffic: "):])
                if rec.get("options") == opts:
                    return rec["fetch_bytes"] + rec["write_bytes"]
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(fn, reps):
    ts, r = [], None
    for _ in range(reps):
        a = time.perf_counter()
        r = fn()
        ts.append((time.perf_counter() - a) * 1e3)
    return ts, r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="grid1000", choices=["grid1000", "rmat20", "union512"])
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-parameter autotuner")
    ap.add_argument("--tune-db", default=TUNE_DB, help="autotuner results (JSON, per graph and batch size)")
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve", action="store_true", help="run the serving workload instead")
    ap.add_argument("--serve-rate", type=float, default=1500.0)
    ap.add_argument("--serve-requests", type=int, default=3000)
    ap.add_argument("--grid-slots", type=int, default=0,
                    help="A* slots of the grid engine (0 = as many as half the HBM holds, within the K*n limit)")
    ap.add_argument("--rmat-slots", type=int, default=24)
    ap.add_argument("--epoch", type=int, default=96)
    ap.add_argument("--landmarks", type=int, default=8, help="ALT landmarks for the serving A* (0 = floor(Euclid))")
    ap.add_argument("--grid-mapping", type=int, default=0,
                    help="mapping of the serving grid engine (0 thread, 1 wave, 2 hybrid; -1 = the network's pick)")
    ap.add_argument("--grid-delta-mult", type=float, default=16.0,
                    help="serving grid Delta with landmarks, in multiples of the Davidson default")
    ap.add_argument("--grid-blocks", type=int, default=256, help="workgroups of the serving grid engine (0 = automatic)")
    ap.add_argument("--rmat-mapping", type=int, default=2,
                    help="mapping of the serving R-MAT engine (-1 = the network's pick; hybrid measured fastest)")
    ap.add_argument("--rmat-blocks", type=int, default=1024, help="workgroups of the serving R-MAT engine (0 = automatic)")
    ap.add_argument("--concurrent", action="store_true",
                    help="drive the two serving engines from two host threads (their kernels overlap on the GPU)")
    args = ap.parse_args()
    if args.serve:
        return serve(args)
    if args.workload == "union512":
        return bench_union(args)
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import pathfinding_gpu_amd as P
    from pathfinding_gpu_amd import learned as L

    t0 = time.time()
    if args.workload == "grid1000":
        g = P.grid_road(1000, 1000, 7)
        src0 = 0
    else:
        g = P.rmat(20, 16, 11)
        src0 = g.max_degree_vertex()
    log(f"[bench] graph {g.name}: n={g.n} m={g.m} avgW={g.avg_w:.2f} avgDeg={g.avg_deg:.2f} ({time.time() - t0:.1f}s)")
    K = args.batch
    QA = args.queries if args.workload == "grid1000" else 0
    sv = P.GpuSolver(g, max_batch=max(K, QA, 1))
    info = P.device_info(0)
    log(f"[bench] device {info['name']} {info['arch']} CUs={info['cu_count']} upload {sv.upload_ms():.1f} ms")
    auto = sv.auto_delta()
    rng = np.random.default_rng(2026)
    deg = np.diff(g.row_offsets)
    cand = np.flatnonzero(deg > 0)
    srcs = np.concatenate([[src0], rng.choice(cand[cand != src0], size=K - 1, replace=False)]).astype(np.uint32)

    PERS = P.Drive.Persistent

    def opt(delta, mapping=2, **kw):
        return P.GpuOptions(drive=PERS, delta=int(delta), mapping=P.Mapping(mapping), **kw)

    # ---- learned choices (network on the GPU, judged once against the Davidson default) ----
    tuner = L.Tuner()

    def tf_single(d, mp):
        sv.sssp(src0, opt(d, mp), copy_back=False)
        return min(sv.sssp(src0, opt(d, mp), copy_back=False).ms for _ in range(2))

    def tf_batch(d, mp):
        sv.multi_source(srcs, opt(d, mp), copy_back=False)
        return min(sv.multi_source(srcs, opt(d, mp), copy_back=False).ms for _ in range(2))

    def tf_batch_o(o):
        sv.multi_source(srcs, o, copy_back=False)
        return min(sv.multi_source(srcs, o, copy_back=False).ms for _ in range(2))

    log("[bench] choosing Delta")
    pick1 = tuner.choose(sv, g, 1, tf_single)
    pickK = tuner.choose(sv, g, K, tf_batch)
    log(f"[bench] single: {pick1}")
    log(f"[bench] batch:  {pickK}")

    plain = opt(auto)
    host_ref = P.GpuOptions(drive=P.Drive.Host, sync_every=8)          # reference defaults
    learned1 = opt(pick1["delta"], pick1["mapping"])
    learnedK = opt(pickK["delta"], pickK["mapping"])
    tuned = None
    if not args.no_tune:
        from pathfinding_gpu_amd.autotune import autotune
        log("[bench] autotuning launch parameters of the batched run")

        def tfo(o):
            return sv.multi_source(srcs, o, copy_back=False).ms

        tuned = autotune(g, K, pickK["delta"], info["cu_count"], tfo, passes=2, reps=1, cache=args.tune_db)
        cand = tuned.options(info["cu_count"], delta=pickK["delta"])
        if tf_batch_o(cand) < tf_batch(pickK["delta"], pickK["mapping"]):
            learnedK = cand
        log(f"[bench] autotuned (batch): {tuned}")

        def tfo1(o):
            return sv.sssp(src0, o, copy_back=False).ms

        tuned1 = autotune(g, 1, pick1["delta"], info["cu_count"], tfo1, passes=2, reps=2, cache=args.tune_db)
        cand1 = tuned1.options(info["cu_count"], delta=pick1["delta"])
        if min(tfo1(cand1) for _ in range(2)) < tf_single(pick1["delta"], pick1["mapping"]):
            learned1 = cand1
        log(f"[bench] autotuned (single): {tuned1}")

    # ---- warm-up, then timed steps: the learned batched run ----
    for _ in range(args.warmup):
        sv.multi_source(srcs, learnedK, copy_back=False)
    log(f"[bench] timed: {args.steps} steps of K={K} batched SSSP")
    t_step, r_step = timed(lambda: sv.multi_source(srcs, learnedK, copy_back=False), args.steps)
    reached_k, arcs_k = sv.last_reach()
    ms_step = float(np.mean(t_step))
    gteps = arcs_k / (ms_step * 1e-3) / 1e9

    # ---- plain and learned single query ----
    reps = max(3, args.steps // 2)
    sv.sssp(src0, plain, copy_back=False)
    t_plain, r_plain = timed(lambda: sv.sssp(src0, plain, copy_back=False), reps)
    reached1, arcs1 = sv.last_reach()
    sv.sssp(src0, host_ref, copy_back=False)
    t_host, r_host = timed(lambda: sv.sssp(src0, host_ref, copy_back=False), reps)
    sv.sssp(src0, learned1, copy_back=False)
    t_l1, r_l1 = timed(lambda: sv.sssp(src0, learned1, copy_back=False), reps)
    ms_plain, ms_host, ms_l1 = float(np.mean(t_plain)), float(np.mean(t_host)), float(np.mean(t_l1))
    gteps_plain = arcs1 / (ms_plain * 1e-3) / 1e9

    # roofline of the dominant kernel: the persistent step kernel of the batched run (one launch per step).
    # Algorithmic bytes per launch: 16 per reached (k, v) (row extent + own state) + 16 per scanned arc
    # (col, w + target state) over all K queries; duration = its hipEvent span on the engine stream.
    bytes_launch = 16 * reached_k + 16 * arcs_k
    loop_ms = r_step.stats.loop_ms
    achieved = bytes_launch / (loop_ms * 1e-3) / 1e9
    timed_opts = {"delta": int(learnedK.delta), "mapping": int(learnedK.mapping), "block_size": int(learnedK.block_size),
                  "wave_cutoff": int(learnedK.wave_cutoff), "block_cutoff": int(learnedK.block_cutoff),
                  "grid_blocks": int(learnedK.grid_blocks), "async_light": int(learnedK.async_light),
                  "atomic": int(learnedK.atomic)}
    traffic = measured_traffic(timed_opts, args.workload) if K == 64 else None

    out = {
        "metric": "GTEPS, learned-Delta batched (K=64 SSSP per launch) vs plain default-Delta single query; "
                  "+ A* queries/s and latency; bit-exact",
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
        "config": {"workload": f"{g.name}: SSSP from {src0} (+{K - 1} random sources per batch)"
                               + (f" + {QA} A* queries" if QA else ""),
                   "n": g.n, "m": g.m, "batch": K, "auto_delta": auto,
                   "learned_single": {"delta": pick1["delta"], "mapping": pick1["mapping"], "source": pick1["source"]},
                   "learned_batch": {"delta": pickK["delta"], "mapping": pickK["mapping"], "source": pickK["source"]},
                   "drive": "persistent",
                   "autotuned": None if tuned is None else {"block_size": learnedK.block_size,
                                                            "wave_cutoff": learnedK.wave_cutoff,
                                                            "block_cutoff": learnedK.block_cutoff,
                                                            "grid_blocks": learnedK.grid_blocks,
                                                            "mapping": int(learnedK.mapping),
                                                            "trials": tuned.trials}},
        "plain_gteps": round(gteps_plain, 5),
        "plain_ms": round(ms_plain, 3),
        "speedup_vs_plain": round(gteps / gteps_plain, 3),
        "host_driven_reference_style_ms": round(ms_host, 3),
        "learned_single_ms": round(ms_l1, 3),
        "learned_single_options": {"delta": int(learned1.delta), "mapping": int(learned1.mapping),
                                   "block_size": int(learned1.block_size), "grid_blocks": int(learned1.grid_blocks),
                                   "async_light": int(learned1.async_light), "atomic": int(learned1.atomic)},
        "learned_single_gteps": round(arcs1 / (ms_l1 * 1e-3) / 1e9, 5),
        "learned_single_speedup": round(ms_plain / ms_l1, 3),
        "step_stats": {"steps": r_step.stats.steps, "buckets": r_step.stats.buckets,
                       "edges_scanned": r_step.stats.edges_scanned,
                       "atomics_per_update": round(r_step.stats.atomics_issued / max(1, r_step.stats.updates), 4),
                       "event_ms": round(r_step.ms, 3), "loop_ms": round(loop_ms, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None if traffic is None else int(traffic),
                     "bytes_per_launch": int(bytes_launch), "kernel": "stepKernel (persistent, K-source batch)",
                     "launch_options": timed_opts},
    }

    if QA:
        pa = np.stack([rng.integers(0, g.n, QA), rng.integers(0, g.n, QA)], 1).astype(np.uint32)
        pickA = L.DeltaModel.load().choose(g, QA) if tuner.model is not None else {"delta": auto, "mapping": 2}
        aopt = opt(pickA["delta"], pickA["mapping"], heuristic=P.Heuristic.FloorEuclid)
        if not args.no_tune:   # launch shape of the A* batch at the learned Delta (kept only if faster)
            from pathfinding_gpu_amd.autotune import autotune

            def tfe(o_):
                return sv.astar_batch(pa[:, 0], pa[:, 1], o_, path_cap=8192)["ms"]

            e_tuned = autotune(g, QA, int(pickA["delta"]), info["cu_count"], tfe, passes=2, reps=1,
                               cache=args.tune_db, kind="astar-euclid")
            cand = e_tuned.options(info["cu_count"], delta=int(pickA["delta"]), heuristic=P.Heuristic.FloorEuclid)
            tfe(cand)
            if min(tfe(cand) for _ in range(2)) < min(tfe(aopt) for _ in range(2)):
                aopt = cand
            log(f"[bench] autotuned (A* floor Euclid): {e_tuned}")
        sv.astar_batch(pa[:, 0], pa[:, 1], aopt, path_cap=8192)
        a_ts, _ = timed(lambda: sv.astar_batch(pa[:, 0], pa[:, 1], aopt, path_cap=8192), max(2, args.steps // 3))
        a_ms = float(np.mean(a_ts))
        popt = opt(auto, heuristic=P.Heuristic.FloorEuclid)
        nq = 8
        sv.astar(int(pa[0, 0]), int(pa[0, 1]), popt, path_cap=8192)
        p_ts, _ = timed(lambda: [sv.astar(int(pa[i, 0]), int(pa[i, 1]), popt, path_cap=8192) for i in range(nq)], 1)
        # one query at a time with the tiled light phase (4x the Davidson Delta, thread mapping, 256 x 128)
        topt = P.GpuOptions(drive=PERS, delta=int(auto * 4), mapping=P.Mapping.Thread, heuristic=P.Heuristic.FloorEuclid,
                            async_light=2, grid_blocks=256, block_size=128)
        sv.astar(int(pa[0, 0]), int(pa[0, 1]), topt, path_cap=8192)
        t_ts, _ = timed(lambda: [sv.astar(int(pa[i, 0]), int(pa[i, 1]), topt, path_cap=8192) for i in range(nq)], 1)
        lm_ids, _ = sv.build_landmarks(8)
        # Delta for the landmark heuristic: f = d + h is nearly constant along shortest paths, so
        # wider buckets pay off; judged once by timing among 1x, 4x, 8x and 16x the Davidson default
        alt_ms = {}
        for mult in (1, 4, 8, 16):
            for asy in (0, 1, 2):
                o_ = opt(auto * mult, pickA["mapping"], heuristic=P.Heuristic.ALT, async_light=asy)
                sv.astar_batch(pa[:, 0], pa[:, 1], o_, path_cap=8192)
                alt_ms[(mult, asy)] = sv.astar_batch(pa[:, 0], pa[:, 1], o_, path_cap=8192)["ms"]
        alt_mult, alt_async = min(alt_ms, key=alt_ms.get)
        lopt = opt(auto * alt_mult, pickA["mapping"], heuristic=P.Heuristic.ALT, async_light=alt_async)
        alt_tuned = None
        if not args.no_tune:   # launch shape of the ALT batch at that Delta (kept only if faster)
            from pathfinding_gpu_amd.autotune import autotune

            def tfa(o_):
                return sv.astar_batch(pa[:, 0], pa[:, 1], o_, path_cap=8192)["ms"]

            alt_tuned = autotune(g, QA, int(auto * alt_mult), info["cu_count"], tfa, passes=2, reps=1,
                                 cache=args.tune_db, kind="astar-alt")
            cand = alt_tuned.options(info["cu_count"], delta=int(auto * alt_mult), heuristic=P.Heuristic.ALT)
            tfa(cand)
            if min(tfa(cand) for _ in range(2)) < min(tfa(lopt) for _ in range(2)):
                lopt = cand
            log(f"[bench] autotuned (A* ALT): {alt_tuned}")
        ra = sv.astar_batch(pa[:, 0], pa[:, 1], lopt, path_cap=8192)
        re_ = sv.astar_batch(pa[:, 0], pa[:, 1], aopt, path_cap=8192)
        alt_same = bool(np.array_equal(ra["dist"], re_["dist"]) and
                        all(np.array_equal(x, y) for x, y in zip(ra["paths"], re_["paths"])))
        l_ts, _ = timed(lambda: sv.astar_batch(pa[:, 0], pa[:, 1], lopt, path_cap=8192), max(2, args.steps // 3))
        l_ms = float(np.mean(l_ts))
        out["astar_alt"] = {"queries": QA, "landmarks": len(lm_ids), "delta": int(auto * alt_mult),
                            "async_light": int(lopt.async_light), "mapping": int(lopt.mapping),
                            "grid_blocks": lopt.grid_blocks, "block_size": lopt.block_size,
                            "judged_ms": {f"{m}x{('', '/async', '/tiled')[a]}": round(v, 2) for (m, a), v in alt_ms.items()},
                            "batch_ms": round(l_ms, 3),
                            "queries_per_s": round(QA / (l_ms * 1e-3), 1),
                            "edges_scanned": ra["stats"].edges_scanned,
                            "edges_scanned_floor_euclid": re_["stats"].edges_scanned,
                            "same_dist_and_paths_as_floor_euclid": alt_same}
        out["astar"] = {"queries": QA, "batch_ms": round(a_ms, 3), "queries_per_s": round(QA / (a_ms * 1e-3), 1),
                        "latency_ms_p50": round(float(np.percentile(a_ts, 50)), 3),
                        "latency_ms_p99": round(float(np.percentile(a_ts, 99)), 3),
                        "delta": pickA["delta"], "mapping": int(aopt.mapping), "grid_blocks": aopt.grid_blocks,
                        "block_size": aopt.block_size, "async_light": int(aopt.async_light),
                        "plain_one_at_a_time_queries_per_s": round(nq / (p_ts[0] * 1e-3), 1),
                        "tiled_one_at_a_time_queries_per_s": round(nq / (t_ts[0] * 1e-3), 1),
                        "note": "all queries of a batch complete together: latency = batch time"}

    if not args.no_cpu:
        log("[bench] CPU baseline")
        import ctypes as C

        import oracle
        R = oracle.ref()
        kind = "reference" if R is not None else "port"
        row, col, w = g.row_offsets, g.col_indices, g.weights
        u32p = C.POINTER(C.c_uint32)
        if kind == "reference":
            srcv = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(row))
            gh = R.ref_build_csr(g.n, g.m, srcv.ctypes.data_as(u32p), col.ctypes.data_as(u32p), w.ctypes.data_as(u32p),
                                 None, None)
            dist = np.empty(g.n, np.uint32)
            par = np.empty(g.n, np.uint32)
            reps_c, tot = 0, 0.0
            while tot < args.cpu_seconds * 1e3 * 0.5 and reps_c < 100:
                tot += R.ref_dijkstra(gh, int(src0), dist.ctypes.data_as(u32p), par.ctypes.data_as(u32p), None)
                reps_c += 1
            kk = min(K, 64)
            md = np.empty(kk * g.n, np.uint32)
            e = C.c_uint64()
            thr = C.c_int()
            ms_omp = R.ref_multi_source(gh, srcs[:kk].ctypes.data_as(u32p), kk, 0, md.ctypes.data_as(u32p), C.byref(e),
                                        C.byref(thr))
            R.ref_free(gh)
            threads = thr.value
        else:
            reps_c = max(1, int(args.cpu_seconds * 500 / max(1.0, oracle.time_dijkstra(row, col, w, int(src0), 1))))
            tot = oracle.time_dijkstra(row, col, w, int(src0), reps_c)
            dist = oracle.dijkstra(row, col, w, int(src0))[0]
            kk = min(K, 64)
            a = time.perf_counter()
            md, _, threads = oracle.multi_source(row, col, w, srcs[:kk], 0)
            ms_omp = (time.perf_counter() - a) * 1e3
            md = md.ravel()
        cpu_ms = tot / reps_c
        gpu1 = sv.sssp(int(src0), learned1)
        gpuK = sv.multi_source(srcs[:kk] if kk == K else srcs, learnedK)
        out["bitexact_vs_reference_cpu"] = {
            "single_dist": bool(np.array_equal(gpu1.dist, dist)),
            "batch_dist": bool(np.array_equal(gpuK.dist[:kk].ravel(), md)),
        }
        out["cpu_baseline"] = {"value": round(arcs1 / (cpu_ms * 1e-3) / 1e9, 5), "unit": "GTEPS", "cores": 1,
                               "kind": kind, "ms_per_sssp": round(cpu_ms, 3),
                               "sample": f"cpuDijkstra from vertex {src0} on the same graph, {reps_c} runs"}
        out["cpu_multisource_omp"] = {"value": round(arcs1 * kk / (ms_omp * 1e-3) / 1e9, 5), "unit": "GTEPS",
                                      "cores": threads, "kind": kind, "ms": round(ms_omp, 2),
                                      "sample": f"cpuMultiSource over the first {kk} batch sources"}
        out["host"] = {"cpu_threads": os.cpu_count()}
    sv.close()
    print(json.dumps(out))


def bench_union(args):
    """512 grid road graphs with sides uniform in [64, 512] (seeds 1000+i), one random source each, run as
    one query over their disjoint union (one seed per graph) in a single launch: the persistent kernel, and
    the captured hipGraph of 64 steps replayed until done."""
    import pathfinding_gpu_amd as P
    from pathfinding_gpu_amd import learned as L
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(4)
    sides = rng.integers(64, 513, 512)
    t0 = time.time()
    gs = [P.grid_road(int(a), int(a), 1000 + i) for i, a in enumerate(sides)]
    u, voff = P.union(gs)
    local = np.array([int(rng.integers(0, gs[i].n)) for i in range(len(gs))], np.uint32)
    seeds = (voff[:-1] + local).astype(np.uint32)
    log(f"[union] {len(gs)} graphs, n={u.n} m={u.m} built in {time.time() - t0:.1f}s")
    sv = P.GpuSolver(u)
    pick = L.DeltaModel.load().choose(gs[int(np.argmax(sides))], 1)
    auto = sv.auto_delta()
    res, opts = {}, {}
    for name, o in (("persistent_learned", P.GpuOptions(drive=P.Drive.Persistent, delta=pick["delta"],
                                                         mapping=P.Mapping(pick["mapping"]))),
                    ("graph_learned", P.GpuOptions(drive=P.Drive.Graph, sync_every=64, delta=pick["delta"],
                                                   mapping=P.Mapping(pick["mapping"]))),
                    ("persistent_davidson", P.GpuOptions(drive=P.Drive.Persistent, delta=auto))):
        sv.multi_seed(seeds, o, copy_back=False)
        ts, r = timed(lambda: sv.multi_seed(seeds, o, copy_back=False), max(2, args.steps // 2))
        res[name] = (float(np.mean(ts)), r)
        opts[name] = o
    if not args.no_tune:   # launch shape (mapping, light phase, workgroups) at the learned Delta
        from pathfinding_gpu_amd.autotune import autotune
        cus = P.device_info(0)["cu_count"]
        tuned = autotune(u, 1, int(pick["delta"]), cus, lambda o_: sv.multi_seed(seeds, o_, copy_back=False).ms,
                         passes=2, reps=1, cache=args.tune_db, kind="union512")
        o = tuned.options(cus, delta=int(pick["delta"]))
        sv.multi_seed(seeds, o, copy_back=False)
        ts, r = timed(lambda: sv.multi_seed(seeds, o, copy_back=False), max(2, args.steps // 2))
        res["persistent_tuned"] = (float(np.mean(ts)), r)
        opts["persistent_tuned"] = o
        # the same launch shape driven as one captured hipGraph of 64 steps, replayed until done
        og = tuned.options(cus, delta=int(pick["delta"]))
        og.drive, og.sync_every = P.Drive.Graph, 64
        sv.multi_seed(seeds, og, copy_back=False)
        ts, r = timed(lambda: sv.multi_seed(seeds, og, copy_back=False), max(2, args.steps // 2))
        res["graph_tuned"] = (float(np.mean(ts)), r)
        opts["graph_tuned"] = og
        log(f"[union] autotuned: {tuned}")
    reached, arcs = sv.last_reach()
    best = min(res, key=lambda k: res[k][0])
    ms = res[best][0]
    out = {"metric": "GTEPS, 512 uneven grid road graphs in one launch (disjoint union, one seed per graph)",
           "value": round(arcs / (ms * 1e-3) / 1e9, 4), "unit": "GTEPS", "n_gpus": 1, "ms_per_step": round(ms, 3),
           "steps": max(2, args.steps // 2), "higher_is_better": True, "dtype": "u32", "data": "synthetic",
           "config": {"workload": "512 grids, side U[64,512], seeds 1000+i", "n": u.n, "m": u.m, "best": best,
                      "delta": pick["delta"], "mapping": pick["mapping"], "auto_delta": auto},
           "variants_ms": {k: round(v[0], 3) for k, v in res.items()},
           "graph_replays": res["graph_learned"][1].stats.launches,
           "graph_tuned_replays": res["graph_tuned"][1].stats.launches if "graph_tuned" in res else None}
    if not args.no_cpu:
        import oracle
        r = sv.multi_seed(seeds, opts[best])       # the timed variant's results, per instance

        def check(i):
            g = gs[i]
            d, _, pc, _, _ = oracle.dijkstra(g.row_offsets, g.col_indices, g.weights, int(local[i]))
            sl = slice(int(voff[i]), int(voff[i + 1]))
            par = r.parent[sl].astype(np.int64)
            par[par != 0xFFFFFFFF] -= int(voff[i])
            return bool(np.array_equal(r.dist[sl], d) and np.array_equal(par, pc.astype(np.int64)))

        a = time.perf_counter()
        with ThreadPoolExecutor(16) as ex:
            ok = list(ex.map(check, range(len(gs))))
        out["bitexact_instances"] = f"{sum(ok)}/{len(ok)}"
        out["checked_variant"] = best
        out["cpu_oracle_check_s"] = round(time.perf_counter() - a, 2)
    sv.close()
    print(json.dumps(out))


def serve(args):
    """Serving workload: a 2000x2000 grid road graph (seed 3) and an R-MAT scale-22 edge-factor-16 graph
    (weights U[1,255], seed 22) resident in HBM; Poisson arrivals, 90% A* on the grid, 10% full SSSP on
    the R-MAT, continuous batching into the persistent multi-slot kernel."""
    import pathfinding_gpu_amd as P
    from pathfinding_gpu_amd import learned as L
    from pathfinding_gpu_amd.serving import Server, SlotEngine, poisson_requests

    t0 = time.time()
    grid = P.grid_road(2000, 2000, 3)
    rmat = P.rmat(22, 16, 22)
    log(f"[serve] graphs built in {time.time() - t0:.1f}s: grid n={grid.n} m={grid.m}, rmat n={rmat.n} m={rmat.m}")
    model = L.DeltaModel.load()
    if args.grid_slots <= 0:   # memory-based admission: the slots half the HBM holds
        from pathfinding_gpu_amd.serving import slots_for_memory
        args.grid_slots = slots_for_memory(grid.n, int(P.device_info(0)["global_mem"]), fraction=0.5)
    pg = model.choose(grid, args.grid_slots)
    pr = model.choose(rmat, args.rmat_slots)
    if args.rmat_mapping >= 0:
        pr = {**pr, "mapping": args.rmat_mapping}
    heur = P.Heuristic.ALT if args.landmarks else P.Heuristic.FloorEuclid
    if args.landmarks:   # wider buckets for the landmark bound (f nearly constant along shortest paths)
        pg = {**pg, "delta": int(args.grid_delta_mult * L.davidson_delta(grid.avg_w, grid.avg_deg))}
    if args.grid_mapping >= 0:   # the A* batch sweeps on the 1M grid: thread mapping, 256 workgroups fastest
        pg = {**pg, "mapping": args.grid_mapping}
    eng = {
        "grid": SlotEngine(grid, args.grid_slots, P.GpuOptions(drive=P.Drive.Persistent, delta=pg["delta"],
                                                               mapping=P.Mapping(pg["mapping"]), heuristic=heur,
                                                               grid_blocks=args.grid_blocks),
                           astar=True, landmarks=args.landmarks),
        "rmat": SlotEngine(rmat, args.rmat_slots, P.GpuOptions(drive=P.Drive.Persistent, delta=pr["delta"],
                                                               mapping=P.Mapping(pr["mapping"]),
                                                               grid_blocks=args.rmat_blocks), astar=False,
                           keep_sssp=False),
    }
    reqs = poisson_requests(args.serve_rate, args.serve_requests, {"grid": 0.9, "rmat": 0.1},
                            {"grid": grid.n, "rmat": rmat.n}, {"grid": True, "rmat": False}, seed=5)
    warm = poisson_requests(args.serve_rate, 64, {"grid": 0.9, "rmat": 0.1}, {"grid": grid.n, "rmat": rmat.n},
                            {"grid": True, "rmat": False}, seed=6)
    srv = Server(eng, epoch_steps=args.epoch, concurrent=args.concurrent)
    srv.run(warm, time_limit_s=120, log=log)
    rep = srv.run(reqs, time_limit_s=600, log=log)
    astar_done = [r for r in reqs if r.graph == "grid"]
    out = {
        "metric": "serving: A* + SSSP queries/s and p50/p99 latency (continuous batching), bit-exact",
        "value": round(rep.qps, 1), "unit": "queries/s", "n_gpus": 1, "higher_is_better": True,
        "offered_rate_qps": args.serve_rate, "requests": rep.requests, "wall_s": round(rep.wall_s, 3),
        "latency_ms": {k: round(v, 2) for k, v in rep.latency_ms.items()},
        "per_graph": rep.per_graph,
        "astar_qps": round(len(astar_done) / rep.wall_s, 1),
        "config": {"workload": "grid 2000x2000 seed 3 (90% A*) + rmat s22 ef16 seed 22 (10% SSSP), "
                               "Poisson arrivals", "grid_slots": args.grid_slots, "rmat_slots": args.rmat_slots,
                   "epoch_steps": args.epoch, "grid_delta": pg["delta"], "grid_mapping": pg["mapping"],
                   "grid_blocks": args.grid_blocks, "concurrent_engines": args.concurrent,
                   "astar_heuristic": heur.name, "landmarks": args.landmarks,
                   "rmat_delta": pr["delta"], "rmat_mapping": pr["mapping"], "rmat_blocks": args.rmat_blocks},
        "dtype": "u32", "data": "synthetic",
    }
    for e in eng.values():
        e.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
