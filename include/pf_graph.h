/*
 * pf_graph.h -- C-ABI of the host graph layer: workload generators, CSR
 * construction, disjoint union, binary cache.  Caller-allocated buffers only.
 *
 * Reference counterparts: src/graph.cpp buildCSR :23-104 (same semantics:
 * self-loops dropped, w >= 1, stable by source, each row sorted by (w, dst),
 * statistics and hScale), saveBinary/loadBinary :320-367 (same on-disk
 * layout, magic "PFCSR001").  The generators are NOT the reference's
 * genGridRoad/genRMAT (graph.cpp:113-199): they produce the workloads of
 * this engine's benchmark configurations (grid road graph with integer
 * jitter and ceil(len*U[lo,hi]) weights; directed R-MAT with duplicates and
 * U[1,wmax] weights), from a counter-based SplitMix64 stream so that every
 * arc can be generated independently and in parallel.
 */
#ifndef PF_GRAPH_H
#define PF_GRAPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of directed arcs of a rows x cols grid road graph: 2 per road. */
uint64_t pf_grid_road_arcs(uint32_t rows, uint32_t cols);

/* Grid road graph.  Vertex id = r*cols + c at (100c + jx, 100r + jy), jx, jy
 * uniform integers in [-jitter, jitter].  Roads join 4-neighbours; each road
 * gets one weight ceil(len * U[wlo, whi)) shared by its two arcs.  Arcs are
 * emitted road by road (horizontal roads row-major, then vertical roads),
 * each as (a->b, b->a).  src/dst/w: pf_grid_road_arcs entries; x/y: n. */
int pf_gen_grid_road(uint32_t rows, uint32_t cols, uint64_t seed, int32_t jitter, double wlo, double whi,
                     uint32_t* src, uint32_t* dst, uint32_t* w, double* x, double* y);

/* Directed R-MAT: ef * 2^scale arcs, quadrant probabilities (a, b, c, 1-a-b-c)
 * chosen per level, most significant bit first; w = 1 + (draw % wmax).  No
 * vertex permutation; self-loops are left for the CSR builder to drop,
 * duplicates are kept. */
int pf_gen_rmat(uint32_t scale, uint32_t ef, double a, double b, double c, uint32_t wmax, uint64_t seed,
                uint32_t* src, uint32_t* dst, uint32_t* w);

/* buildCSR semantics (src/graph.cpp:23-74).  row: n+1, col/wout: m_in capacity.
 * *m_out receives the arcs kept. */
int pf_build_csr(uint32_t n, uint64_t m_in, const uint32_t* src, const uint32_t* dst, const uint32_t* w,
                 uint32_t* row, uint32_t* col, uint32_t* wout, uint32_t* m_out);

/* Statistics (src/graph.cpp:76-101): iout = {minW, maxW, maxDeg},
 * dout = {avgW, avgDeg, hScale}.  x/y may be NULL. */
int pf_csr_stats(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w, const double* x,
                 const double* y, uint32_t* iout, double* dout);

/* Disjoint union of `count` CSR graphs: vertex v of graph i becomes
 * voff[i] + v.  voff: count+1 entries (output).  row: (sum n)+1, col/w: sum m. */
int pf_csr_union(uint32_t count, const uint32_t* ns, const uint32_t* ms, const uint32_t* const* rows,
                 const uint32_t* const* cols, const uint32_t* const* ws, uint32_t* voff, uint32_t* row,
                 uint32_t* col, uint32_t* w);

/* Binary cache in the reference's PFCSR001 layout (src/graph.cpp:320-336). */
int pf_csr_save(const char* path, const char* name, uint32_t n, uint32_t m, const uint32_t* row,
                const uint32_t* col, const uint32_t* w, const double* x, const double* y);
/* Header only: n, m, has_xy. */
int pf_csr_peek(const char* path, uint32_t* n, uint32_t* m, uint32_t* has_xy);
/* Body into caller buffers sized from pf_csr_peek (x/y may be NULL). */
int pf_csr_load(const char* path, uint32_t* row, uint32_t* col, uint32_t* w, double* x, double* y);

/* Vertex of maximum out-degree (smallest id on ties). */
uint32_t pf_max_degree_vertex(uint32_t n, const uint32_t* row);

#ifdef __cplusplus
}
#endif
#endif
