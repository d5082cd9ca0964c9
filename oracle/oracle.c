/*
 * oracle.c -- CPU restatement of the reference's shortest-path semantics.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product path.  See oracle.h for
 * the parent rules.  Reference citations are relative to the reference
 * repository root.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Binary min-heaps (textbook; the reference uses std::priority_queue with    */
/* std::greater, src/cpu_sssp.cpp:33,73).                                     */
/* ------------------------------------------------------------------------- */
typedef struct {
    uint64_t* a;
    size_t n, cap;
} heap64;

static void h64_push(heap64* h, uint64_t x) {
    if (h->n == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 1024;
        h->a = (uint64_t*)realloc(h->a, h->cap * sizeof(uint64_t));
    }
    size_t i = h->n++;
    while (i > 0) {
        size_t p = (i - 1) / 2;
        if (h->a[p] <= x) break;
        h->a[i] = h->a[p];
        i = p;
    }
    h->a[i] = x;
}

static uint64_t h64_pop(heap64* h) {
    uint64_t top = h->a[0];
    uint64_t x = h->a[--h->n];
    size_t i = 0;
    for (;;) {
        size_t l = 2 * i + 1;
        if (l >= h->n) break;
        size_t c = (l + 1 < h->n && h->a[l + 1] < h->a[l]) ? l + 1 : l;
        if (h->a[c] >= x) break;
        h->a[i] = h->a[c];
        i = c;
    }
    if (h->n) h->a[i] = x;
    return top;
}

typedef struct {
    uint64_t f;
    uint32_t v;
} fitem;

typedef struct {
    fitem* a;
    size_t n, cap;
} heapf;

static int fless(fitem a, fitem b) { return a.f < b.f || (a.f == b.f && a.v < b.v); }

static void hf_push(heapf* h, fitem x) {
    if (h->n == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 1024;
        h->a = (fitem*)realloc(h->a, h->cap * sizeof(fitem));
    }
    size_t i = h->n++;
    while (i > 0) {
        size_t p = (i - 1) / 2;
        if (!fless(x, h->a[p])) break;
        h->a[i] = h->a[p];
        i = p;
    }
    h->a[i] = x;
}

static fitem hf_pop(heapf* h) {
    fitem top = h->a[0];
    fitem x = h->a[--h->n];
    size_t i = 0;
    for (;;) {
        size_t l = 2 * i + 1;
        if (l >= h->n) break;
        size_t c = (l + 1 < h->n && fless(h->a[l + 1], h->a[l])) ? l + 1 : l;
        if (!fless(h->a[c], x)) break;
        h->a[i] = h->a[c];
        i = c;
    }
    if (h->n) h->a[i] = x;
    return top;
}

/* ------------------------------------------------------------------------- */
/* CSR                                                                       */
/* ------------------------------------------------------------------------- */
static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* src/graph.cpp:23-74 */
uint32_t or_build_csr(uint32_t n, size_t m_in, const uint32_t* src, const uint32_t* dst,
                      const uint32_t* w, uint32_t* row, uint32_t* col, uint32_t* wout) {
    uint32_t* deg = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
    size_t keep = 0;
    for (size_t i = 0; i < m_in; ++i)
        if (src[i] != dst[i]) { deg[src[i]]++; keep++; }           /* graph.cpp:30-33 */
    row[0] = 0;
    for (uint32_t u = 0; u < n; ++u) row[u + 1] = row[u] + deg[u]; /* graph.cpp:37-38 */
    uint32_t* cur = deg;
    for (uint32_t u = 0; u < n; ++u) cur[u] = row[u];
    for (size_t i = 0; i < m_in; ++i) {                            /* graph.cpp:44-51 */
        if (src[i] == dst[i]) continue;
        uint32_t p = cur[src[i]]++;
        col[p] = dst[i];
        wout[p] = w[i] < 1u ? 1u : w[i];
    }
    free(deg);
    uint64_t* tmp = NULL;
    size_t tcap = 0;
    for (uint32_t u = 0; u < n; ++u) {                             /* graph.cpp:57-74 */
        uint32_t b = row[u], e = row[u + 1];
        if (e - b < 2) continue;
        if (e - b > tcap) {
            tcap = e - b;
            tmp = (uint64_t*)realloc(tmp, tcap * sizeof(uint64_t));
        }
        for (uint32_t k = b; k < e; ++k) tmp[k - b] = ((uint64_t)wout[k] << 32) | col[k];
        qsort(tmp, e - b, sizeof(uint64_t), cmp_u64);
        for (uint32_t k = b; k < e; ++k) {
            wout[k] = (uint32_t)(tmp[k - b] >> 32);
            col[k] = (uint32_t)tmp[k - b];
        }
    }
    free(tmp);
    return (uint32_t)keep;
}

/* src/graph.cpp:76-101 */
void or_csr_stats(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                  const double* x, const double* y, or_stats* out) {
    uint32_t m = row[n];
    out->minW = m ? 0xFFFFFFFFu : 0;
    out->maxW = 0;
    double sumW = 0.0;
    for (uint32_t k = 0; k < m; ++k) {
        if (w[k] < out->minW) out->minW = w[k];
        if (w[k] > out->maxW) out->maxW = w[k];
        sumW += w[k];
    }
    out->avgW = m ? sumW / m : 0.0;
    out->maxDeg = 0;
    for (uint32_t u = 0; u < n; ++u)
        if (row[u + 1] - row[u] > out->maxDeg) out->maxDeg = row[u + 1] - row[u];
    out->avgDeg = n ? (double)m / n : 0.0;
    out->hScale = 0.0;
    if (x && y) {
        double scale = INFINITY;
        for (uint32_t u = 0; u < n; ++u)
            for (uint32_t k = row[u]; k < row[u + 1]; ++k) {
                double dx = x[u] - x[col[k]], dy = y[u] - y[col[k]];
                double e = sqrt(dx * dx + dy * dy);
                if (e > 1e-12 && w[k] / e < scale) scale = w[k] / e;
            }
        out->hScale = isfinite(scale) ? scale : 0.0;
    }
}

/* src/gpu_sssp.cu:202-206 */
uint32_t or_auto_delta(double avgW, double avgDeg) {
    double d = 32.0 * avgW / (avgDeg > 1.0 ? avgDeg : 1.0);
    long long r = llround(d);
    return (uint32_t)(r < 1 ? 1 : r);
}

/* src/gpu_kernels.cuh:361-371 */
void or_light_end(uint32_t n, const uint32_t* row, const uint32_t* w, uint32_t delta,
                  uint32_t* light_end) {
    for (uint32_t u = 0; u < n; ++u) {
        uint32_t k = row[u];
        while (k < row[u + 1] && w[k] <= delta) ++k;   /* linear scan: independent of the binary search */
        light_end[u] = k;
    }
}

/* ------------------------------------------------------------------------- */
/* Heuristics                                                                */
/* ------------------------------------------------------------------------- */
static uint32_t isqrt_floor(uint64_t a) {
    /* bit-by-bit integer square root: largest r with r*r <= a */
    uint64_t r = 0, bit = (uint64_t)1 << 62;
    while (bit > a) bit >>= 2;
    while (bit) {
        if (a >= r + bit) {
            a -= r + bit;
            r = (r >> 1) + bit;
        } else {
            r >>= 1;
        }
        bit >>= 2;
    }
    return (uint32_t)r;
}

/* include/heuristic.h:19-25 for OR_H_SCALED */
uint32_t or_heuristic(int kind, const double* x, const double* y, double hScale, uint32_t v,
                      uint32_t t) {
    if (kind == OR_H_SCALED) {
        volatile double dx = x[v] - x[t], dy = y[v] - y[t];
        volatile double s = dx * dx;
        s = s + dy * dy;
        double e = sqrt(s) * hScale * (1.0 - 1e-9);
        if (!(e > 0.0)) return 0u;
        if (e >= 4.0e9) return 4000000000u;
        return (uint32_t)e;
    }
    if (kind == OR_H_FLOOR_EUCLID) {
        int64_t dx = (int64_t)x[v] - (int64_t)x[t], dy = (int64_t)y[v] - (int64_t)y[t];
        return isqrt_floor((uint64_t)(dx * dx) + (uint64_t)(dy * dy));
    }
    return 0u;
}

/* ------------------------------------------------------------------------- */
/* Dijkstra (src/cpu_sssp.cpp:28-57)                                          */
/* ------------------------------------------------------------------------- */
static void dijkstra_core(uint32_t n, const uint32_t* row, const uint32_t* col,
                          const uint32_t* w, const uint32_t* seeds, uint32_t nseeds,
                          uint32_t* dist, uint32_t* pfirst, uint32_t* pcanon, uint64_t* st) {
    uint32_t* d = dist ? dist : (uint32_t*)malloc((size_t)n * 4);
    for (uint32_t v = 0; v < n; ++v) d[v] = OR_INF;
    if (pfirst)
        for (uint32_t v = 0; v < n; ++v) pfirst[v] = OR_NO_PARENT;
    if (pcanon)
        for (uint32_t v = 0; v < n; ++v) pcanon[v] = OR_NO_PARENT;
    heap64 h = {0};
    for (uint32_t i = 0; i < nseeds; ++i) {
        d[seeds[i]] = 0;
        h64_push(&h, (uint64_t)seeds[i]);
    }
    uint64_t settled = 0, edges = 0;
    while (h.n) {
        uint64_t top = h64_pop(&h);
        uint32_t du = (uint32_t)(top >> 32), u = (uint32_t)top;
        if (du > d[u]) continue;                     /* cpu_sssp.cpp:42 lazy deletion */
        settled++;
        for (uint32_t k = row[u]; k < row[u + 1]; ++k) {
            uint32_t v = col[k];
            uint64_t nd = (uint64_t)du + w[k];
            edges++;
            if (nd < d[v]) {                         /* cpu_sssp.cpp:48-51 */
                d[v] = (uint32_t)nd;
                if (pfirst) pfirst[v] = u;
                if (pcanon) pcanon[v] = u;
                h64_push(&h, (nd << 32) | v);
            } else if (nd == d[v] && pcanon && u < pcanon[v]) {
                pcanon[v] = u;                       /* min-index tie rule */
            }
        }
    }
    free(h.a);
    if (!dist) free(d);
    if (st) {
        st[0] = settled;
        st[1] = edges;
    }
}

void or_dijkstra(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                 uint32_t src, uint32_t* dist, uint32_t* parent_first, uint32_t* parent_canon,
                 uint64_t* stats_out) {
    dijkstra_core(n, row, col, w, &src, 1, dist, parent_first, parent_canon, stats_out);
}

void or_dijkstra_multi_seed(uint32_t n, const uint32_t* row, const uint32_t* col,
                            const uint32_t* w, const uint32_t* seeds, uint32_t nseeds,
                            uint32_t* dist, uint32_t* parent_canon) {
    dijkstra_core(n, row, col, w, seeds, nseeds, dist, NULL, parent_canon, NULL);
}

/* ------------------------------------------------------------------------- */
/* A* (src/cpu_sssp.cpp:64-101)                                               */
/* ------------------------------------------------------------------------- */
uint32_t or_astar(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                  const double* x, const double* y, double hScale, int hkind, uint32_t src,
                  uint32_t dst, int canonical, uint32_t* path, uint32_t path_cap,
                  uint32_t* path_len, uint64_t* stats_out) {
    int useH = hkind != OR_H_NONE && x && y && (hkind != OR_H_SCALED || hScale > 0.0);
    uint32_t* dist = (uint32_t*)malloc((size_t)n * 4);
    uint32_t* par = (uint32_t*)malloc((size_t)n * 4);
    for (uint32_t v = 0; v < n; ++v) { dist[v] = OR_INF; par[v] = OR_NO_PARENT; }
#define HV(v) (useH ? (uint64_t)or_heuristic(hkind, x, y, hScale, (v), dst) : 0u)
    heapf h = {0};
    dist[src] = 0;
    fitem s0 = {HV(src), src};
    hf_push(&h, s0);
    uint64_t settled = 0, edges = 0;
    while (h.n) {
        fitem it = hf_pop(&h);
        uint32_t u = it.v;
        uint64_t du = dist[u];
        if (it.f > du + HV(u)) continue;             /* cpu_sssp.cpp:83 stale entry */
        if (canonical) {
            if (dist[dst] != OR_INF && it.f > dist[dst]) break;   /* every f <= d(t) settled */
        }
        settled++;
        if (!canonical && u == dst) break;           /* cpu_sssp.cpp:85 */
        for (uint32_t k = row[u]; k < row[u + 1]; ++k) {
            uint32_t v = col[k];
            uint64_t nd = du + w[k];
            edges++;
            if (nd < dist[v]) {                      /* cpu_sssp.cpp:90-94 */
                dist[v] = (uint32_t)nd;
                par[v] = u;
                fitem e = {nd + HV(v), v};
                hf_push(&h, e);
            } else if (canonical && nd == dist[v] && u < par[v]) {
                par[v] = u;
            }
        }
    }
#undef HV
    uint32_t d = dist[dst];
    uint32_t len = 0;
    if (d != OR_INF && path) len = or_reconstruct_path(n, par, src, dst, path, path_cap);
    if (path_len) *path_len = len;
    if (stats_out) { stats_out[0] = settled; stats_out[1] = edges; }
    free(h.a);
    free(dist);
    free(par);
    return d;
}

/* src/cpu_sssp.cpp:103-122 */
int or_multi_source(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                    const uint32_t* srcs, uint32_t k, int threads, uint32_t* dist_out,
                    uint64_t* edges_out) {
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#else
    threads = 1;
#endif
    uint64_t edges = 0;
    int K = (int)k;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) reduction(+ : edges)
#endif
    for (int i = 0; i < K; ++i) {
        uint64_t st[2];
        uint32_t* d = dist_out ? dist_out + (size_t)i * n : NULL;
        dijkstra_core(n, row, col, w, &srcs[i], 1, d, NULL, NULL, st);
        edges += st[1];
    }
    if (edges_out) *edges_out = edges;
    return threads;
}

/* src/cpu_sssp.cpp:124-135 */
uint32_t or_reconstruct_path(uint32_t n, const uint32_t* parent, uint32_t src, uint32_t dst,
                             uint32_t* path, uint32_t path_cap) {
    if (dst >= n) return 0;
    uint32_t v = dst, len = 0;
    for (size_t steps = 0; steps <= n; ++steps) {
        if (len >= path_cap) return 0;
        path[len++] = v;
        if (v == src) {
            for (uint32_t i = 0; i < len / 2; ++i) {
                uint32_t t = path[i];
                path[i] = path[len - 1 - i];
                path[len - 1 - i] = t;
            }
            return len;
        }
        v = parent[v];
        if (v == OR_NO_PARENT || v >= n) break;
    }
    return 0;
}

void or_canonical_parents(uint32_t n, const uint32_t* row, const uint32_t* col,
                          const uint32_t* w, const uint32_t* dist, uint32_t* parent) {
    for (uint32_t v = 0; v < n; ++v) parent[v] = OR_NO_PARENT;
    for (uint32_t u = 0; u < n; ++u) {
        if (dist[u] == OR_INF) continue;
        for (uint32_t k = row[u]; k < row[u + 1]; ++k) {
            uint32_t v = col[k];
            if ((uint64_t)dist[u] + w[k] == dist[v] && u < parent[v]) parent[v] = u;
        }
    }
}

void or_bellman_ford(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                     uint32_t src, uint32_t* dist) {
    for (uint32_t v = 0; v < n; ++v) dist[v] = OR_INF;
    dist[src] = 0;
    for (uint32_t round = 0; round + 1 < n || round == 0; ++round) {
        int changed = 0;
        for (uint32_t u = 0; u < n; ++u) {
            if (dist[u] == OR_INF) continue;
            for (uint32_t k = row[u]; k < row[u + 1]; ++k) {
                uint64_t nd = (uint64_t)dist[u] + w[k];
                if (nd < dist[col[k]]) { dist[col[k]] = (uint32_t)nd; changed = 1; }
            }
        }
        if (!changed) break;
    }
}

/* ------------------------------------------------------------------------- */
/* Verification (src/verify.cpp)                                             */
/* ------------------------------------------------------------------------- */
static uint32_t min_arc_weight(const uint32_t* row, const uint32_t* col, const uint32_t* w,
                               uint32_t a, uint32_t b) {
    uint32_t best = OR_INF;                          /* verify.cpp:18-23 */
    for (uint32_t k = row[a]; k < row[a + 1]; ++k)
        if (col[k] == b && w[k] < best) best = w[k];
    return best;
}

int or_validate_path(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                     const uint32_t* path, uint32_t len, uint32_t src, uint32_t dst,
                     uint32_t expected) {
    (void)n;
    if (expected == OR_INF) return len == 0;         /* verify.cpp:27-30 */
    if (len == 0 || path[0] != src || path[len - 1] != dst) return 0;
    uint64_t sum = 0;
    for (uint32_t i = 0; i + 1 < len; ++i) {
        uint32_t a = min_arc_weight(row, col, w, path[i], path[i + 1]);
        if (a == OR_INF) return 0;
        sum += a;
    }
    return sum == expected;
}

/* SplitMix64 stream for the sampled-path layer (the reference samples targets
 * with its own Rng, include/common.h:44-56; any fixed sampler serves). */
static uint64_t sm64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void or_verify_sssp(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                    uint32_t src, const uint32_t* ref_dist, const uint32_t* dist,
                    const uint32_t* parent, const uint32_t* parent_canon, uint32_t sample_paths,
                    uint64_t seed, uint64_t* checks, uint64_t* errs) {
    for (int i = 0; i < 4; ++i) checks[i] = errs[i] = 0;
    for (uint32_t v = 0; v < n; ++v) {               /* verify.cpp:63-72 */
        checks[0]++;
        if (dist[v] != ref_dist[v]) errs[0]++;
    }
    for (uint32_t v = 0; v < n; ++v) {               /* verify.cpp:74-87 */
        if (v == src || dist[v] == OR_INF) continue;
        checks[1]++;
        uint32_t p = parent[v];
        int ok = p < n && dist[p] != OR_INF;
        if (ok) {
            ok = 0;
            for (uint32_t k = row[p]; k < row[p + 1]; ++k)
                if (col[k] == v && (uint64_t)dist[p] + w[k] == dist[v]) { ok = 1; break; }
        }
        if (!ok) errs[1]++;
    }
    uint32_t* path = (uint32_t*)malloc(((size_t)n + 1) * 4);
    uint64_t s = seed;                               /* verify.cpp:89-97 */
    for (uint32_t i = 0, tries = 0; i < sample_paths && tries < sample_paths * 50u; ++tries) {
        uint32_t t = (uint32_t)(sm64(&s) % n);
        if (ref_dist[t] == OR_INF) continue;
        ++i;
        checks[2]++;
        uint32_t len = or_reconstruct_path(n, parent, src, t, path, n + 1);
        if (!or_validate_path(n, row, col, w, path, len, src, t, ref_dist[t])) errs[2]++;
    }
    free(path);
    if (parent_canon) {
        for (uint32_t v = 0; v < n; ++v) {
            checks[3]++;
            if (parent[v] != parent_canon[v]) errs[3]++;
        }
    }
}

uint64_t or_reachable_arcs(uint32_t n, const uint32_t* row, const uint32_t* dist,
                           uint64_t* reached_out) {
    uint64_t s = 0, r = 0;
    for (uint32_t v = 0; v < n; ++v)
        if (dist[v] != OR_INF) { s += row[v + 1] - row[v]; r++; }
    if (reached_out) *reached_out = r;
    return s;
}

uint64_t or_exclusive_scan(const uint32_t* in, uint32_t* out, size_t count) {
    uint64_t acc = 0;
    for (size_t i = 0; i < count; ++i) {
        out[i] = (uint32_t)acc;
        acc += in[i];
    }
    return acc;
}

uint64_t or_next_bound(uint64_t key, uint32_t delta) { return (key / delta + 1) * (uint64_t)delta; }

double or_time_dijkstra(uint32_t n, const uint32_t* row, const uint32_t* col, const uint32_t* w,
                        uint32_t src, int reps) {
    uint32_t* d = (uint32_t*)malloc((size_t)n * 4);
    uint32_t* p = (uint32_t*)malloc((size_t)n * 4);
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    for (int r = 0; r < reps; ++r) dijkstra_core(n, row, col, w, &src, 1, d, p, NULL, NULL);
    clock_gettime(CLOCK_MONOTONIC, &b);
    free(d);
    free(p);
    return (b.tv_sec - a.tv_sec) * 1e3 + (b.tv_nsec - a.tv_nsec) * 1e-6;
}
