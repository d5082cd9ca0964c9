/*
 * oracle.h -- CPU restatement of the reference's shortest-path semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP engine, its
 * C-ABI, the C++ GpuSolver, the Python package) links, loads or calls this
 * code.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it, and only as the checker / the timed CPU baseline.
 *
 * Every function cites the reference file:line it restates.  All arithmetic
 * is integer (u32 distances and weights, u64 keys) except the scaled
 * heuristic, which repeats the reference's double-precision formula.
 *
 * Parent rule.  The reference's cpuDijkstra keeps the first predecessor that
 * reaches the final distance (src/cpu_sssp.cpp:48-51), i.e. the tight
 * predecessor with the smallest (dist, index) pair.  The GPU engine packs
 * (dist<<32 | parent) into one word and takes a 64-bit minimum, so on equal
 * distance the smaller parent index wins (src/gpu_kernels.cuh:9-13).  The
 * oracle reports both: parent_first (reference rule) and parent_canon
 * (min-index rule, the one the engine is held to bit for bit).
 */
#ifndef PF_ORACLE_H
#define PF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_INF 0xFFFFFFFFu
#define OR_NO_PARENT 0xFFFFFFFFu

/* Heuristic kinds (match the engine's enum values). */
#define OR_H_NONE 0
#define OR_H_SCALED 1       /* include/heuristic.h:19-25 (double, hScale)   */
#define OR_H_FLOOR_EUCLID 2 /* h = floor(||p(v)-p(t)||) on integer coords */

typedef struct {
    uint32_t minW, maxW, maxDeg;
    double avgW, avgDeg, hScale;
} or_stats;

/* buildCSR (src/graph.cpp:23-104): drop self-loops, w>=1, counting sort by
 * source (stable), then sort each row by (w, dst).  row[n+1], col/w[m_in]
 * capacity.  Returns m (arcs kept). */
uint32_t or_build_csr(uint32_t n, size_t m_in, const uint32_t* src, const uint32_t* dst,
                      const uint32_t* w, uint32_t* row, uint32_t* col, uint32_t* wout);

/* Statistics of buildCSR (src/graph.cpp:76-101).  x/y may be NULL. */
void or_csr_stats(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                  const double* x, const double* y, or_stats* out);

/* autoDelta (src/gpu_sssp.cu:202-206). */
uint32_t or_auto_delta(double avgW, double avgDeg);

/* lightEndKernel semantics (src/gpu_kernels.cuh:361-371): upper_bound of
 * w <= delta inside each weight-sorted row. */
void or_light_end(uint32_t n, const uint32_t* row, const uint32_t* w, uint32_t delta,
                  uint32_t* light_end);

/* Heuristic values.  kind OR_H_SCALED: heuristicValue (include/heuristic.h:19-25)
 * on double coords.  kind OR_H_FLOOR_EUCLID: exact floor(sqrt(dx^2+dy^2)) on the
 * integer-valued coords. */
uint32_t or_heuristic(int kind, const double* x, const double* y, double hScale,
                      uint32_t v, uint32_t t);

/* cpuDijkstra (src/cpu_sssp.cpp:28-57), binary heap keyed (dist<<32|v) with
 * lazy deletion.  Any output pointer may be NULL.  stats_out[0]=settled,
 * stats_out[1]=edgesScanned. */
void or_dijkstra(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                 uint32_t src, uint32_t* dist, uint32_t* parent_first, uint32_t* parent_canon,
                 uint64_t* stats_out);

/* Multi-seed Dijkstra: every seed starts at distance 0 with NO_PARENT (used for
 * the disjoint-union batch of many graphs).  Same outputs as or_dijkstra. */
void or_dijkstra_multi_seed(uint32_t n, const uint32_t* row, const uint32_t* col,
                            const uint32_t* w, const uint32_t* seeds, uint32_t nseeds,
                            uint32_t* dist, uint32_t* parent_canon);

/* cpuAStar (src/cpu_sssp.cpp:64-101).  canonical=0: stop when the target is
 * popped, exactly as the reference (dist and stats then match cpuAStar).
 * canonical=1: keep settling every vertex with f <= dist[t] under the
 * min-index parent rule, so the returned path is the canonical path of the
 * shortest-path tree.  path may be NULL; *path_len receives its length (0 if
 * unreachable or it does not fit path_cap).  stats_out[0]=settled,
 * stats_out[1]=edgesScanned.  Returns dist[t]. */
uint32_t or_astar(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                  const double* x, const double* y, double hScale, int hkind, uint32_t src,
                  uint32_t dst, int canonical, uint32_t* path, uint32_t path_cap,
                  uint32_t* path_len, uint64_t* stats_out);

/* cpuMultiSource (src/cpu_sssp.cpp:103-122): OpenMP dynamic,1 over sources.
 * dist_out is k*n (may be NULL).  threads<=0: all.  Returns threads used;
 * edges_out receives the summed edgesScanned. */
int or_multi_source(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                    const uint32_t* srcs, uint32_t k, int threads, uint32_t* dist_out,
                    uint64_t* edges_out);

/* reconstructPath (src/cpu_sssp.cpp:124-135).  Returns the length written
 * (0 = broken chain / cycle / does not fit). */
uint32_t or_reconstruct_path(uint32_t n, const uint32_t* parent, uint32_t src, uint32_t dst,
                             uint32_t* path, uint32_t path_cap);

/* Closed form of the min-index parent rule:
 * parent(v) = min{u : (u,v) in E, dist[u] + w(u,v) == dist[v]} for reached
 * v != seed, NO_PARENT otherwise.  O(m). */
void or_canonical_parents(uint32_t n, const uint32_t* row, const uint32_t* col,
                          const uint32_t* w, const uint32_t* dist, uint32_t* parent);

/* Brute-force Bellman-Ford (n-1 synchronous rounds) for tiny graphs. */
void or_bellman_ford(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                     uint32_t src, uint32_t* dist);

/* verifySSSP (src/verify.cpp:54-99) layers (1) dist equality, (2) tree
 * validity, (3) sampled path validity; plus (4) exact parent equality with
 * parent_canon when it is non-NULL.  errs[0..3] receive the error counts of
 * the four layers, checks[0..3] the numbers of checks. */
void or_verify_sssp(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                    uint32_t src, const uint32_t* ref_dist, const uint32_t* dist,
                    const uint32_t* parent, const uint32_t* parent_canon, uint32_t sample_paths,
                    uint64_t seed, uint64_t* checks, uint64_t* errs);

/* validatePath (src/verify.cpp:25-52): arcs exist (min-weight parallel arc)
 * and the cost equals expected.  Returns 1 if valid. */
int or_validate_path(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                     const uint32_t* path, uint32_t len, uint32_t src, uint32_t dst,
                     uint32_t expected);

/* Reachable arcs: sum of out-degrees of reached vertices (src/main.cpp:209-213),
 * and the reached-vertex count. */
uint64_t or_reachable_arcs(uint32_t n, const uint32_t* row, const uint32_t* dist,
                           uint64_t* reached_out);

/* Exclusive prefix sum of u32 flags (the arithmetic of cub::DeviceScan::
 * ExclusiveSum at src/gpu_sssp.cu:312).  Returns the total. */
uint64_t or_exclusive_scan(const uint32_t* in, uint32_t* out, size_t count);

/* Bucket bound arithmetic (src/gpu_sssp.cu:370,441): (key/delta+1)*delta. */
uint64_t or_next_bound(uint64_t key, uint32_t delta);

/* Timed single-thread Dijkstra repeated reps times (CPU baseline leg);
 * returns total milliseconds. */
double or_time_dijkstra(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                        uint32_t src, int reps);

#ifdef __cplusplus
}
#endif
#endif
