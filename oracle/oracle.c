/*
 * ORACLE -- test infrastructure, not product code.  See oracle.h.
 *
 * Compiled with -ffp-contract=off so FP64 expressions are evaluated exactly
 * as written (one rounding per operation).
 *
 * Planted faults (orc_set_fault), used only by the oracle's own tests:
 *   1  BFS relax proposes source value + 2 instead of + 1
 *   2  SSSP relax of on-demand parts drops the edge weight
 *   3  CC relax propagates the larger label instead of the smaller
 *   4  reference PageRank divides by the in-degree of the target instead of
 *      the out-degree of the source
 *   5  normalised PageRank drops the dangling-mass term
 *   6  on-demand part relax indexes the gathered buffer with the global
 *      prefix instead of the part-relative one
 */
#include "oracle.h"

#include <limits.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

static int g_fault = 0;
void orc_set_fault(int fault) { g_fault = fault; }

/* Bounded CPU-baseline samples: the restated traversal stops after this many
 * iterations (0 = run to the end); its values are then partial. */
static int g_iter_limit = 0;
void orc_set_iteration_limit(int n) { g_iter_limit = n > 0 ? n : 0; }
/* Timing samples: once this many edges are relaxed (or pulled), the
 * remaining vertices of the pass are skipped and the run ends (0 = off). */
static uint64_t g_edge_limit = 0;
void orc_set_edge_limit(uint64_t n) { g_edge_limit = n; }
static inline int over_edge_limit(uint64_t done) { return g_edge_limit && done >= g_edge_limit; }
void orc_set_num_threads(int n) { if (n > 0) omp_set_num_threads(n); }

enum { ALG_BFS = 0, ALG_SSSP = 1, ALG_CC = 2 };

static inline uint32_t edst(const orc_graph* g, uint64_t i) { return g->dst[i * (uint64_t)g->stride]; }
static inline uint32_t ewt(const orc_graph* g, uint64_t i) {
    return g->w ? g->w[i * (uint64_t)g->stride] : 1u;
}
static inline uint64_t degree(const orc_graph* g, uint64_t v) { return g->off[v + 1] - g->off[v]; }

/* check-then-atomicMin as in gpu_kernels.cuh:390-392 */
static inline void relax_min(uint32_t* val, uint8_t* active, uint32_t v, uint32_t cand) {
    uint32_t cur = __atomic_load_n(&val[v], __ATOMIC_RELAXED);
    if (cand < cur) {
        while (cand < cur &&
               !__atomic_compare_exchange_n(&val[v], &cur, cand, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
        __atomic_store_n(&active[v], 1, __ATOMIC_RELAXED);
    }
}

static inline uint32_t propose(int algo, uint32_t sval, uint32_t w) {
    switch (algo) {
        case ALG_BFS: return sval + (g_fault == 1 ? 2u : 1u);   /* gpu_kernels.cuh:387 */
        case ALG_SSSP: return sval + w;                         /* gpu_kernels.cuh:496 */
        default: return sval;                                   /* gpu_kernels.cuh:453 */
    }
}

static void relax_cc_fault(uint32_t* val, uint8_t* active, uint32_t v, uint32_t cand) {
    /* fault 3: larger label wins */
    uint32_t cur = __atomic_load_n(&val[v], __ATOMIC_RELAXED);
    if (cand > cur) {
        while (cand > cur &&
               !__atomic_compare_exchange_n(&val[v], &cur, cand, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
        __atomic_store_n(&active[v], 1, __ATOMIC_RELAXED);
    }
}

static int nthreads_or_default(int n) { return n > 0 ? n : omp_get_max_threads(); }

/* Reference static flag: off[v] < S && off[v] + deg - 1 < S, in unsigned
 * 64-bit arithmetic (GraphMeta.cuh:357-365). */
static inline int ref_is_static(const orc_graph* g, uint64_t v, uint64_t S) {
    uint64_t o = g->off[v], d = degree(g, v);
    return o < S && (o + d - 1) < S;
}

/* Vertex-granular split restated from GraphMeta.cuh:516-578, with a guard
 * where the reference loops forever.  Falls back to edge-granular parts in
 * that case (the reference would hang; results are unaffected). */
static int split_parts_for_relax(const uint64_t* prefix, const uint32_t* deg, uint32_t n, uint64_t total,
                                 uint64_t part, uint32_t** starts, uint32_t** counts, int* vertex_granular) {
    int cap = 16;
    uint32_t* s = (uint32_t*)malloc(sizeof(uint32_t) * cap);
    uint32_t* c = (uint32_t*)malloc(sizeof(uint32_t) * cap);
    uint64_t* e = (uint64_t*)malloc(sizeof(uint64_t) * cap);
    int np;
    for (;;) {
        np = orc_split_reference(prefix, deg, n, total, part, s, c, e, cap);
        if (np != -2) break;
        cap *= 4;
        s = (uint32_t*)realloc(s, sizeof(uint32_t) * cap);
        c = (uint32_t*)realloc(c, sizeof(uint32_t) * cap);
        e = (uint64_t*)realloc(e, sizeof(uint64_t) * cap);
    }
    free(e);
    *vertex_granular = np >= 0;
    if (np < 0) { /* hub larger than a part: one part holds everything */
        s[0] = 0;
        c[0] = n;
        np = 1;
    }
    *starts = s;
    *counts = c;
    return np;
}

static int traverse_ref(const orc_graph* g, int algo, uint32_t src, uint64_t S, uint64_t part_edges,
                        int nthreads, uint32_t* val, orc_stats* st) {
    const uint64_t V = g->V;
    if (V == 0 || V >= UINT32_MAX - 1) return -1;
    if (algo != ALG_CC && src >= V) return -1;
    if (algo == ALG_SSSP && !g->w) return -1;
    const int nt = nthreads_or_default(nthreads);
    uint8_t* active = (uint8_t*)calloc(V, 1);
    uint8_t* isstat = (uint8_t*)malloc(V);
    uint32_t* slist = (uint32_t*)malloc(sizeof(uint32_t) * V);
    uint32_t* olist = (uint32_t*)malloc(sizeof(uint32_t) * V);
    uint64_t* oprefix = (uint64_t*)malloc(sizeof(uint64_t) * (V + 1));
    uint32_t* odeg = (uint32_t*)malloc(sizeof(uint32_t) * V);
    if (!active || !isstat || !slist || !olist || !oprefix || !odeg) return -2;
    uint32_t* gather = NULL;
    uint32_t* gatherw = NULL;
    uint64_t gather_cap = 0;
    orc_stats local;
    memset(&local, 0, sizeof(local));

    for (uint64_t v = 0; v < V; v++) isstat[v] = (uint8_t)ref_is_static(g, v, S);
    /* refreshLabelAndValue (GraphMeta.cuh:412-467) */
    if (algo == ALG_CC) {
        for (uint64_t v = 0; v < V; v++) { val[v] = (uint32_t)v; active[v] = 1; }
    } else {
        for (uint64_t v = 0; v < V; v++) val[v] = (uint32_t)(V + 1);
        val[src] = 1;
        active[src] = 1;
    }
    uint64_t count = 0;
    for (uint64_t v = 0; v < V; v++) count += active[v];
    local.node_sum = count;

    while (count && !(g_iter_limit > 0 && local.iterations >= (uint32_t)g_iter_limit) &&
           !over_edge_limit(local.edges_relaxed)) {
        local.iterations++;
        /* setStaticAndOverloadLabelBool + scan + compaction (CalculateOpt.cu:63-91):
         * lists in increasing vertex order. */
        uint64_t ns = 0, no = 0;
        for (uint64_t v = 0; v < V; v++) {
            if (!active[v]) continue;
            if (isstat[v]) slist[ns++] = (uint32_t)v; else olist[no++] = (uint32_t)v;
        }
        /* host prefix of overload degrees (CalculateOpt.cu:99-109) */
        uint64_t ototal = 0;
        for (uint64_t i = 0; i < no; i++) {
            oprefix[i] = ototal;
            odeg[i] = (uint32_t)degree(g, olist[i]);
            ototal += odeg[i];
        }
        oprefix[no] = ototal;
        /* setLabelDefaultOpt on both lists (CalculateOpt.cu:126-135) */
        for (uint64_t i = 0; i < ns; i++) active[slist[i]] = 0;
        for (uint64_t i = 0; i < no; i++) active[olist[i]] = 0;

        /* static relax (bfs_kernelStatic / sssp_kernel / cc_kernelStaticSwap) */
        uint64_t relaxed = 0;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 64) reduction(+ : relaxed)
        for (uint64_t i = 0; i < ns; i++) {
            if (over_edge_limit(local.edges_relaxed + relaxed)) continue;
            uint32_t u = slist[i];
            uint32_t sval = __atomic_load_n(&val[u], __ATOMIC_RELAXED);
            uint64_t b = g->off[u], e = g->off[u + 1];
            relaxed += e - b;
            for (uint64_t k = b; k < e; k++) {
                uint32_t v = edst(g, k);
                uint32_t cand = propose(algo, sval, ewt(g, k));
                if (algo == ALG_CC && g_fault == 3) relax_cc_fault(val, active, v, cand);
                else relax_min(val, active, v, cand);
            }
        }
        local.edges_relaxed += relaxed;

        if (no && !over_edge_limit(local.edges_relaxed)) {
            /* fillEdgeArrByMultiThread (GraphMeta.cuh:469-514) */
            if (ototal > gather_cap) {
                free(gather);
                free(gatherw);
                gather_cap = ototal;
                gather = (uint32_t*)malloc(sizeof(uint32_t) * gather_cap);
                gatherw = g->w ? (uint32_t*)malloc(sizeof(uint32_t) * gather_cap) : NULL;
                if (!gather || (g->w && !gatherw)) return -2;
            }
#pragma omp parallel for num_threads(nt) schedule(dynamic, 64)
            for (uint64_t i = 0; i < no; i++) {
                uint64_t from = g->off[olist[i]];
                for (uint32_t j = 0; j < odeg[i]; j++) {
                    gather[oprefix[i] + j] = edst(g, from + j);
                    if (gatherw) gatherw[oprefix[i] + j] = ewt(g, from + j);
                }
            }
            local.overload_edges += ototal;
            /* caculatePartInfoForEdgeList + per-part relax (CalculateOpt.cu:161-178) */
            uint32_t *pstart, *pcount;
            int vg;
            int np = split_parts_for_relax(oprefix, odeg, (uint32_t)no, ototal,
                                           part_edges ? part_edges : ototal, &pstart, &pcount, &vg);
            local.parts += (uint32_t)np;
            for (int p = 0; p < np; p++) {
                uint64_t first = pstart[p], cnt = pcount[p];
                uint64_t base = oprefix[first];
                relaxed = 0;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 64) reduction(+ : relaxed)
                for (uint64_t i = first; i < first + cnt; i++) {
                    if (over_edge_limit(local.edges_relaxed + relaxed)) continue;
                    uint32_t u = olist[i];
                    uint32_t sval = __atomic_load_n(&val[u], __ATOMIC_RELAXED);
                    /* part-relative offset, as in gpu_kernels.cuh:358-359 */
                    uint64_t at = (g_fault == 6) ? oprefix[i] : oprefix[i] - base;
                    relaxed += odeg[i];
                    for (uint32_t j = 0; j < odeg[i]; j++) {
                        uint64_t idx = base + at + j;
                        if (idx >= ototal) idx = ototal - 1;
                        uint32_t v = gather[idx];
                        uint32_t w = gatherw ? gatherw[idx] : 1u;
                        if (algo == ALG_SSSP && g_fault == 2) w = 0;
                        uint32_t cand = propose(algo, sval, w);
                        if (algo == ALG_CC && g_fault == 3) relax_cc_fault(val, active, v, cand);
                        else relax_min(val, active, v, cand);
                    }
                }
                local.edges_relaxed += relaxed;
            }
            free(pstart);
            free(pcount);
        }
        count = 0;
#pragma omp parallel for num_threads(nt) reduction(+ : count)
        for (uint64_t v = 0; v < V; v++) count += active[v];
        local.node_sum += count;
    }
    free(active); free(isstat); free(slist); free(olist); free(oprefix); free(odeg);
    free(gather); free(gatherw);
    if (st) *st = local;
    return 0;
}

int orc_bfs_ref(const orc_graph* g, uint32_t src, uint64_t static_edges, uint64_t part_edges, int nthreads,
                uint32_t* val, orc_stats* st) {
    return traverse_ref(g, ALG_BFS, src, static_edges, part_edges, nthreads, val, st);
}
int orc_sssp_ref(const orc_graph* g, uint32_t src, uint64_t static_edges, uint64_t part_edges, int nthreads,
                 uint32_t* val, orc_stats* st) {
    return traverse_ref(g, ALG_SSSP, src, static_edges, part_edges, nthreads, val, st);
}
int orc_cc_ref(const orc_graph* g, uint64_t static_edges, uint64_t part_edges, int nthreads, uint32_t* val,
               orc_stats* st) {
    return traverse_ref(g, ALG_CC, 0, static_edges, part_edges, nthreads, val, st);
}

/* ---- PageRank ---------------------------------------------------------- */
int orc_pr_ref(const orc_graph* g, const uint32_t* outdeg, double damping, double threshold, int max_iter,
               double* val, orc_stats* st) {
    const uint64_t V = g->V;
    if (V == 0) return -1;
    /* The reference hard-codes the literals 0.15 and 0.85
     * (gpu_kernels.cuh:538); 1 - 0.85 differs from 0.15 in the last bit. */
    const double base = (damping == 0.85) ? 0.15 : 1.0 - damping;
    double* sum = (double*)calloc(V, sizeof(double));
    uint8_t* active = (uint8_t*)malloc(V);
    if (!sum || !active) return -2;
    for (uint64_t v = 0; v < V; v++) { val[v] = 1.0 / (double)V; active[v] = 1; }
    uint64_t count = V;
    int it = 0;
    orc_stats local;
    memset(&local, 0, sizeof(local));
    local.node_sum = count;
    while (count && (max_iter <= 0 || it < max_iter)) {
        it++;
        /* prSumKernel_static / _dynamic: sequential sum in CSR order */
#pragma omp parallel for schedule(dynamic, 256)
        for (uint64_t v = 0; v < V; v++) {
            if (!active[v]) continue;
            double s = 0;
            for (uint64_t k = g->off[v]; k < g->off[v + 1]; k++) {
                uint32_t u = edst(g, k);
                double d = (g_fault == 4) ? (double)(g->off[v + 1] - g->off[v]) : (double)outdeg[u];
                s += val[u] / d;
            }
            sum[v] = s;
        }
        /* prKernel_Opt (gpu_kernels.cuh:533-552) */
        count = 0;
#pragma omp parallel for reduction(+ : count)
        for (uint64_t v = 0; v < V; v++) {
            if (!active[v]) continue;
            double t = base + damping * sum[v];
            double diff = t > val[v] ? (t - val[v]) : (val[v] - t);
            if (diff > threshold) {
                val[v] = t;
                count++;
            } else {
                active[v] = 0;
            }
            sum[v] = 0;
        }
        local.node_sum += count;
        local.edges_relaxed += 0;
    }
    local.iterations = (uint32_t)it;
    if (st) *st = local;
    free(sum);
    free(active);
    return it;
}

static uint64_t g_last_pulled = 0;
uint64_t orc_last_pulled(void) { return g_last_pulled; }

int orc_pr_norm(const orc_graph* g, const uint32_t* outdeg, double damping, int iters, double* rank,
                double* residual) {
    g_last_pulled = 0;
    const uint64_t V = g->V;
    if (V == 0) return -1;
    double* nr = (double*)malloc(sizeof(double) * V);
    double* contrib = (double*)malloc(sizeof(double) * V);
    if (!nr || !contrib) return -2;
    for (uint64_t v = 0; v < V; v++) rank[v] = 1.0 / (double)V;
    for (int it = 0; it < iters; it++) {
        double dangling = 0;
#pragma omp parallel for reduction(+ : dangling)
        for (uint64_t v = 0; v < V; v++) {
            contrib[v] = outdeg[v] ? rank[v] / (double)outdeg[v] : 0.0;
            if (!outdeg[v]) dangling += rank[v];
        }
        if (g_fault == 5) dangling = 0;
        const double base = (1.0 - damping) / (double)V + damping * dangling / (double)V;
        double res = 0;
        uint64_t pulled = 0;
#pragma omp parallel for reduction(+ : res, pulled)
        for (uint64_t v = 0; v < V; v++) {
            if (over_edge_limit(pulled)) continue;
            double s = 0;
            for (uint64_t k = g->off[v]; k < g->off[v + 1]; k++) s += contrib[edst(g, k)];
            pulled += g->off[v + 1] - g->off[v];
            nr[v] = base + damping * s;
            res += fabs(nr[v] - rank[v]);
        }
        memcpy(rank, nr, sizeof(double) * V);
        if (residual) residual[it] = res;
        g_last_pulled += pulled;
        if (over_edge_limit(pulled)) break;
    }
    free(nr);
    free(contrib);
    return 0;
}

/* ---- sampled edges of generated graphs --------------------------------- */
static uint64_t sm_finalise(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint32_t prob_threshold(double x) {
    double t = x * 4294967296.0;
    return t >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)t;
}

int orc_rmat_pairs(uint64_t seed, int scale, double a, double b, double c, uint64_t M, uint32_t wmin,
                   uint32_t wmax, uint64_t k, uint32_t* u, uint32_t* v, uint32_t* w) {
    if (scale < 1 || scale > 31 || wmax < wmin) return -1;
    const uint64_t key = sm_finalise(seed ^ 0x5A17EC0DEull), wkey = sm_finalise(seed ^ 0x3E16A7ull);
    const uint32_t ta = prob_threshold(a), tab = prob_threshold(a + b), tabc = prob_threshold(a + b + c);
    const uint64_t span = (uint64_t)wmax - wmin + 1;
#pragma omp parallel for schedule(static)
    for (uint64_t j = 0; j < k; j++) {
        const uint64_t i = (uint64_t)(((unsigned __int128)j * M) / k);
        uint32_t x = 0, y = 0;
        uint64_t word = 0;
        for (int l = 0; l < scale; l++) {
            if ((l & 1) == 0) word = sm_finalise(key + ((i << 5) + (uint64_t)(l >> 1)) * 0x9E3779B97F4A7C15ull);
            const uint32_t r = (uint32_t)(word >> ((l & 1) * 32));
            x = (x << 1) | (uint32_t)(r >= tab);
            y = (y << 1) | (uint32_t)((r >= ta && r < tab) || r >= tabc);
        }
        const uint32_t lo = x < y ? x : y, hi = x < y ? y : x;
        const uint64_t h = sm_finalise(wkey ^ (((uint64_t)lo << 32) | hi));
        u[j] = x;
        v[j] = y;
        w[j] = wmin + (uint32_t)(((unsigned __int128)h * span) >> 64);
    }
    return 0;
}

/* fillEdgeArrByMultiThread (GraphMeta.cuh:469-514): the adjacency of each
 * listed vertex copied, in list order, into one contiguous buffer; the list
 * is cut into `nthreads` chunks of ceil(n / nthreads) vertices (one thread
 * when n < 50), each copied by its own thread.  Records are copied whole
 * (stride * 4 bytes: {dst, w} when the graph is AoS).  Returns the edges
 * copied; *copy_ms is the wall time of the threaded copy alone (the
 * reference's fillProcess), the degree prefix excluded. */
long long orc_gather_ref(const orc_graph* g, const uint32_t* list, uint64_t n, int nthreads, uint32_t* out,
                         double* copy_ms) {
    uint64_t* pre = (uint64_t*)malloc(sizeof(uint64_t) * (n + 1));
    if (!pre) return -2;
    pre[0] = 0;
    for (uint64_t i = 0; i < n; i++) pre[i + 1] = pre[i] + degree(g, list[i]);
    const int nt = n < 50 ? 1 : (nthreads > 0 ? nthreads : 20);
    const uint64_t chunk = (n + (uint64_t)nt - 1) / (uint64_t)nt;
    const int words = g->stride;
    const double t0 = omp_get_wtime();
#pragma omp parallel num_threads(nt)
    {
        const uint64_t t = (uint64_t)omp_get_thread_num();
        const uint64_t lo = t * chunk < n ? t * chunk : n, hi = lo + chunk < n ? lo + chunk : n;
        for (uint64_t i = lo; i < hi; i++) {
            const uint64_t from = g->off[list[i]], d = pre[i + 1] - pre[i];
            memcpy(out + pre[i] * (uint64_t)words, g->dst + from * (uint64_t)words, 4 * d * (uint64_t)words);
        }
    }
    if (copy_ms) *copy_ms = (omp_get_wtime() - t0) * 1e3;
    const long long total = (long long)pre[n];
    free(pre);
    return total;
}

/* ---- textbook algorithms ----------------------------------------------- */
int orc_bfs_queue(const orc_graph* g, uint32_t src, uint32_t* val) {
    const uint64_t V = g->V;
    if (src >= V) return -1;
    uint32_t* q = (uint32_t*)malloc(sizeof(uint32_t) * V);
    if (!q) return -2;
    for (uint64_t v = 0; v < V; v++) val[v] = (uint32_t)(V + 1);
    uint64_t head = 0, tail = 0;
    val[src] = 1;
    q[tail++] = src;
    while (head < tail) {
        uint32_t u = q[head++];
        for (uint64_t k = g->off[u]; k < g->off[u + 1]; k++) {
            uint32_t v = edst(g, k);
            if (val[v] == V + 1) { val[v] = val[u] + 1; q[tail++] = v; }
        }
    }
    free(q);
    return 0;
}

typedef struct { uint64_t d; uint32_t v; } heap_item;
static void heap_push(heap_item* h, uint64_t* n, heap_item it) {
    uint64_t i = (*n)++;
    while (i > 0) {
        uint64_t p = (i - 1) / 2;
        if (h[p].d <= it.d) break;
        h[i] = h[p];
        i = p;
    }
    h[i] = it;
}
static heap_item heap_pop(heap_item* h, uint64_t* n) {
    heap_item top = h[0], last = h[--(*n)];
    uint64_t i = 0;
    for (;;) {
        uint64_t c = 2 * i + 1;
        if (c >= *n) break;
        if (c + 1 < *n && h[c + 1].d < h[c].d) c++;
        if (last.d <= h[c].d) break;
        h[i] = h[c];
        i = c;
    }
    if (*n) h[i] = last;
    return top;
}

int orc_dijkstra(const orc_graph* g, uint32_t src, uint32_t* val) {
    const uint64_t V = g->V, E = g->E;
    if (src >= V || !g->w) return -1;
    uint64_t* d = (uint64_t*)malloc(sizeof(uint64_t) * V);
    heap_item* h = (heap_item*)malloc(sizeof(heap_item) * (E + 1));
    if (!d || !h) return -2;
    for (uint64_t v = 0; v < V; v++) d[v] = UINT64_MAX;
    uint64_t n = 0;
    d[src] = 0;
    heap_item s0 = {0, src};
    heap_push(h, &n, s0);
    while (n) {
        heap_item it = heap_pop(h, &n);
        if (it.d != d[it.v]) continue;
        for (uint64_t k = g->off[it.v]; k < g->off[it.v + 1]; k++) {
            uint32_t v = edst(g, k);
            uint64_t nd = it.d + ewt(g, k);
            if (nd < d[v]) {
                d[v] = nd;
                heap_item x = {nd, v};
                heap_push(h, &n, x);
            }
        }
    }
    for (uint64_t v = 0; v < V; v++)
        val[v] = (d[v] == UINT64_MAX || d[v] + 1 >= V + 1) ? (uint32_t)(V + 1) : (uint32_t)(d[v] + 1);
    free(d);
    free(h);
    return 0;
}

int orc_cc_minreach(const orc_graph* g, uint32_t* val) {
    const uint64_t V = g->V;
    uint32_t* q = (uint32_t*)malloc(sizeof(uint32_t) * (V ? V : 1));
    if (!q) return -2;
    for (uint64_t v = 0; v < V; v++) val[v] = UINT32_MAX;
    for (uint64_t s = 0; s < V; s++) {
        if (val[s] != UINT32_MAX) continue;
        uint64_t head = 0, tail = 0;
        val[s] = (uint32_t)s;
        q[tail++] = (uint32_t)s;
        while (head < tail) {
            uint32_t u = q[head++];
            for (uint64_t k = g->off[u]; k < g->off[u + 1]; k++) {
                uint32_t v = edst(g, k);
                if (val[v] == UINT32_MAX) { val[v] = (uint32_t)s; q[tail++] = v; }
            }
        }
    }
    free(q);
    return 0;
}

static uint32_t uf_find(uint32_t* p, uint32_t x) {
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}
int orc_cc_unionfind(const orc_graph* g, uint32_t* val) {
    const uint64_t V = g->V;
    uint32_t* p = (uint32_t*)malloc(sizeof(uint32_t) * (V ? V : 1));
    if (!p) return -2;
    for (uint64_t v = 0; v < V; v++) p[v] = (uint32_t)v;
    for (uint64_t u = 0; u < V; u++)
        for (uint64_t k = g->off[u]; k < g->off[u + 1]; k++) {
            uint32_t a = uf_find(p, (uint32_t)u), b = uf_find(p, edst(g, k));
            if (a < b) p[b] = a; else if (b < a) p[a] = b;
        }
    for (uint64_t v = 0; v < V; v++) val[v] = uf_find(p, (uint32_t)v);
    free(p);
    return 0;
}

/* ---- planning ---------------------------------------------------------- */
int orc_plan_reference(uint64_t V, uint64_t E, const uint64_t* off, uint64_t edge_size, uint64_t free_bytes,
                       int param_size, double advise_rate, orc_plan* out, uint8_t* is_static) {
    if (!out || !off || edge_size == 0) return -1;
    /* reduceMem = paramSize * sizeof(uint) * V (GraphMeta.cuh:319) */
    uint64_t reduce = (uint64_t)param_size * 4u * V;
    uint64_t total = free_bytes > reduce ? (free_bytes - reduce) / edge_size : 0;   /* :322 */
    /* float adviseK = (10 - E/total)/9, clamped (GraphMeta.cuh:326-339) */
    float k = total ? (float)((10.0 - (double)E / (double)total) / 9.0) : 0.5f;
    if (k < 0) k = 0.5f;
    if (k > 1) k = 1.0f;
    if (advise_rate > 0) k = (float)advise_rate;
    uint64_t maxp = (uint64_t)((double)k * (double)total);                           /* :341 */
    if (maxp > E) maxp = E;                                                          /* :342-344 */
    if (maxp > UINT_MAX) maxp = UINT_MAX;                                            /* :349-352 */
    maxp -= maxp % 4096;                                                             /* :353-354 */
    uint64_t nstatic = 0, edges_in_static = 0;
    for (uint64_t v = 0; v < V; v++) {
        uint64_t o = off[v], d = off[v + 1] - off[v];
        int s = o < maxp && (o + d - 1) < maxp;                                      /* :358 */
        if (is_static) is_static[v] = (uint8_t)s;
        if (s) { nstatic++; edges_in_static += d; }
    }
    out->total_gpu_size = total;
    out->advise_k = k;
    out->max_partition_size = maxp;
    out->static_vertices = nstatic;
    out->part_overload_size = total > maxp ? total - maxp : 0;                       /* :369 */
    out->overload_size = E - edges_in_static;                                        /* :370 */
    return 0;
}

int orc_split_reference(const uint64_t* prefix, const uint32_t* deg, uint32_t n, uint64_t total,
                        uint64_t part_size, uint32_t* starts, uint32_t* counts, uint64_t* edges, int cap) {
    int np = 0;
    if (n == 0) return 0;
    if (part_size < total) {
        uint32_t left = 0, right = n - 1;
        while (prefix[right] + deg[right] - prefix[left] > part_size) {
            uint32_t start = left, end = right, mid = left;
            while (start <= end) {
                mid = (start + end) / 2;
                uint64_t head = prefix[mid] - prefix[left];
                uint64_t tail = prefix[mid] + deg[mid] - prefix[left];
                if (head <= part_size && tail > part_size) break;
                else if (tail <= part_size) start = mid + 1;
                else if (head > part_size) { if (mid == 0) break; end = mid - 1; }
            }
            if (mid == left) return -1; /* no progress: the reference never terminates */
            if (np >= cap) return -2;
            starts[np] = left;
            counts[np] = mid - left;
            edges[np] = prefix[mid] - prefix[left];
            np++;
            left = mid;
        }
        if (np >= cap) return -2;
        starts[np] = left;
        counts[np] = right - left + 1;
        edges[np] = prefix[right] + deg[right] - prefix[left];
        np++;
    } else {
        if (cap < 1) return -2;
        starts[0] = 0;
        counts[0] = n;
        edges[0] = total;
        np = 1;
    }
    return np;
}
