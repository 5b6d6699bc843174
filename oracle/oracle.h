/*
 * ORACLE -- test infrastructure, not product code.
 *
 * A CPU restatement in plain C of the algorithms on Ascetic's hot path
 * (NKU-EmbeddedSystem/Ascetic: CalculateOpt.cu, GraphMeta.cuh,
 * gpu_kernels.cuh), plus textbook algorithms used to pin it.  Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark load this
 * library, and only as a checker / baseline; the product path never does.
 *
 * Every graph here is an out-CSR with V+1 offsets (a sentinel off[V] = E is
 * always present in this API; the reference format has none and derives the
 * last degree as E - off[V-1], GraphMeta.cuh:226).  Weighted graphs pass the
 * weight array with a stride so that both the AoS {dst,w} records of the
 * reference (globals.cuh:9-12) and separate arrays work.
 *
 * Value conventions (GraphMeta.cuh:374-467): BFS/SSSP start the source at 1
 * and leave unreached vertices at V+1; CC starts every vertex at its own id.
 */
#ifndef ASC_ORACLE_H
#define ASC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t V, E;
    const uint64_t* off;     /* V+1 entries                                */
    const uint32_t* dst;     /* element i at dst[i*stride]                  */
    const uint32_t* w;       /* weights, element i at w[i*stride]; may be 0 */
    int stride;              /* 1 for plain arrays, 2 for AoS {dst,w}        */
} orc_graph;

typedef struct {
    uint32_t iterations;
    uint64_t node_sum;       /* initial active count + actives per iteration */
    uint64_t edges_relaxed;  /* edges visited by relax loops                */
    uint64_t overload_edges; /* edges gathered for on-demand parts          */
    uint32_t parts;
} orc_stats;

/* --- restatement of the reference's iteration (CalculateOpt.cu) ----------
 * static_edges: length of the device-resident edge prefix; a vertex is
 * static iff its whole adjacency lies in [0, static_edges)
 * (GraphMeta.cuh:357-365).  Active vertices outside it have their edges
 * gathered into a buffer (GraphMeta.cuh:469-514) and relaxed from there in
 * parts of at most part_edges edges (GraphMeta.cuh:516-578, CalculateOpt.cu:
 * 161-178).  nthreads <= 0 uses all cores.  Results do not depend on
 * static_edges / part_edges / nthreads. */
int orc_bfs_ref(const orc_graph* g, uint32_t src, uint64_t static_edges,
                uint64_t part_edges, int nthreads, uint32_t* val, orc_stats* st);
int orc_sssp_ref(const orc_graph* g, uint32_t src, uint64_t static_edges,
                 uint64_t part_edges, int nthreads, uint32_t* val, orc_stats* st);
int orc_cc_ref(const orc_graph* g, uint64_t static_edges, uint64_t part_edges,
               int nthreads, uint32_t* val, orc_stats* st);

/* PageRank over an in-CSR (g->dst holds in-neighbours) with out-degrees.
 * Reference mode (gpu_kernels.cuh:533-594, CalculateOpt.cu:547-721): FP64,
 * init 1/V, per iteration sum over in-edges of val[u]/outdeg[u] in CSR
 * order for active vertices, t = 0.15 + 0.85*sum, keep active and store t
 * iff |t - val| > threshold, otherwise deactivate without storing.
 * Multiplications and additions are not fused (compile with
 * -ffp-contract=off). Returns the number of iterations. */
int orc_pr_ref(const orc_graph* g, const uint32_t* outdeg, double damping,
               double threshold, int max_iter, double* val, orc_stats* st);
/* Normalised PageRank, FP64 power iteration: r0 = 1/V,
 * r' = (1-d)/V + d*(sum_in r[u]/outdeg[u] + dangling/V), `iters` steps;
 * residual[i] = sum |r_{i+1} - r_i| (optional). */
int orc_pr_norm(const orc_graph* g, const uint32_t* outdeg, double damping,
                int iters, double* rank, double* residual);

/* --- textbook algorithms (independent of the reference's structure) ----- */
int orc_bfs_queue(const orc_graph* g, uint32_t src, uint32_t* val);
/* Dijkstra with a binary heap; returns min(1 + d(v), V + 1). */
int orc_dijkstra(const orc_graph* g, uint32_t src, uint32_t* val);
/* label(v) = min{u : u reaches v}: BFS from u = 0,1,... over unlabelled. */
int orc_cc_minreach(const orc_graph* g, uint32_t* val);
/* Union-find (link larger root under smaller): component minimum.  Equal to
 * orc_cc_minreach on symmetric graphs. */
int orc_cc_unionfind(const orc_graph* g, uint32_t* val);

/* --- static / on-demand planning (GraphMeta.cuh:310-372) ---------------
 * free_bytes replaces cudaMemGetInfo's available memory.  Returns the
 * static prefix length in edges (max_partition_size), the number of static
 * vertices and sizes as the reference computes them, including its 4096-edge
 * floor and UINT_MAX cap. */
typedef struct {
    uint64_t total_gpu_size;     /* edges that fit the free memory          */
    double advise_k;
    uint64_t max_partition_size; /* static prefix, edges                    */
    uint64_t static_vertices;    /* count of vertices flagged isInStatic     */
    uint64_t part_overload_size; /* part buffer, edges                      */
    uint64_t overload_size;      /* E - sum of static degrees                */
} orc_plan;
int orc_plan_reference(uint64_t V, uint64_t E, const uint64_t* off, uint64_t edge_size,
                       uint64_t free_bytes, int param_size, double advise_rate,
                       orc_plan* out, uint8_t* is_static /* V entries or NULL */);

/* Reference part split (GraphMeta.cuh:516-578), vertex-granular.  prefix[i]
 * = start of active vertex i in the gathered list, deg[i] its degree.
 * Writes (start index, node count, edge count) triples; returns the number
 * of parts, or -1 where the reference would never terminate (a single
 * vertex larger than part_size). */
int orc_split_reference(const uint64_t* prefix, const uint32_t* deg, uint32_t n,
                        uint64_t total, uint64_t part_size, uint32_t* starts,
                        uint32_t* counts, uint64_t* edges, int cap);

/* CPU-baseline helpers.  orc_set_iteration_limit(n): the restated
 * traversals (orc_*_ref) stop after n iterations (0 = no limit), for bounded
 * timing samples; orc_set_num_threads(n): OpenMP threads of the functions
 * without an nthreads argument (orc_pr_norm).  orc_gather_ref restates
 * fillEdgeArrByMultiThread (GraphMeta.cuh:469-514): the adjacency records of
 * list[0..n) copied into `out` by nthreads threads (20 in the reference; one
 * when n < 50); returns the edges copied and the copy's wall time. */
void orc_set_iteration_limit(int n);
void orc_set_num_threads(int n);
/* orc_set_edge_limit(n): the restated traversals and orc_pr_norm skip the
 * rest of a pass, and stop, once n edges are relaxed or pulled (0 = no
 * limit); with one thread this bounds a timing sample to about n edges of
 * the full-size graph.  orc_last_pulled(): edges the last orc_pr_norm
 * call pulled. */
void orc_set_edge_limit(uint64_t n);
uint64_t orc_last_pulled(void);
long long orc_gather_ref(const orc_graph* g, const uint32_t* list, uint64_t n, int nthreads, uint32_t* out,
                         double* copy_ms);

/* Sampled edges of a synthetic R-MAT graph too large for a host copy: the
 * library's sampler (a SplitMix64 finaliser over (seed, sample index), two
 * 32-bit quadrant draws per 64-bit word, weights per unordered pair) restated
 * independently, so that a test can draw edges of the generated graph and
 * check properties of a traversal's values on them.  Writes the pair of
 * sample i_j = floor(j * M / k) for j < k (u == v for a dropped self-loop)
 * and its weight in [wmin, wmax]. */
int orc_rmat_pairs(uint64_t seed, int scale, double a, double b, double c, uint64_t M, uint32_t wmin,
                   uint32_t wmax, uint64_t k, uint32_t* u, uint32_t* v, uint32_t* w);

/* Fault injection for the oracle's own tests: 0 = none; any other value
 * plants one deliberate mistake (listed in oracle.c) so a test can show the
 * checks catch it. */
void orc_set_fault(int fault);

#ifdef __cplusplus
}
#endif
#endif
