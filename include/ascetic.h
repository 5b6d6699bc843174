/*
 * ascetic.h -- C ABI of the MI355X (gfx950) graph-traversal library.
 *
 * The library runs BFS, SSSP, connected components (min-label propagation)
 * and PageRank over binary CSR graphs on one AMD Instinct MI355X, with the
 * edge array split into a device-resident static prefix and an on-demand
 * remainder that is streamed from pinned host memory each iteration.
 *
 * Every entry point returns 0 on success or a negative asc_status code; no
 * entry point terminates the process.  Output buffers are caller-allocated
 * host arrays of length V (number of vertices).  All sizes are 64-bit.
 *
 * Reference interfaces replaced (NKU-EmbeddedSystem/Ascetic):
 *   bfs_opt / cc_opt / sssp_opt / pr_opt ........ CalculateOpt.cuh:11-14
 *   GraphMeta<E>::readDataFromFile .............. GraphMeta.cuh:151-171
 *   GraphMeta<E>::getMaxPartitionSize ........... GraphMeta.cuh:310-372
 *   GraphMeta<E>::caculatePartInfoForEdgeList ... GraphMeta.cuh:516-578
 *   ArgumentParser / main (the ptGraph CLI) ..... ArgumentParser.cu:3-48, main.cu:4-22
 */
#ifndef ASCETIC_H
#define ASCETIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASC_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------- */
typedef enum {
    ASC_OK = 0,
    ASC_ERR_INVALID = -1,   /* bad argument (null pointer, source >= V, ...) */
    ASC_ERR_IO = -2,        /* file could not be opened / read / written     */
    ASC_ERR_FORMAT = -3,    /* file size does not match any known layout     */
    ASC_ERR_NOMEM = -4,     /* host or device allocation failed              */
    ASC_ERR_DEVICE = -5,    /* a HIP call failed                             */
    ASC_ERR_NOGPU = -6,     /* no HIP device available                       */
    ASC_ERR_RANGE = -7,     /* size beyond a documented limit                */
    ASC_ERR_UNSUPPORTED = -8,/* graph kind does not support this algorithm   */
    ASC_ERR_BUSY = -9,      /* server queue full / no free query lane        */
    ASC_ERR_STATE = -10     /* call not valid in the object's current state  */
} asc_status;

const char* asc_strerror(int code);
int asc_abi_version(void);

/* ---- on-disk / in-memory formats --------------------------------------- */
/* Binary layouts, little-endian (reader: GraphMeta.cuh:151-171,
 * writer: converter/converter.cpp:178-181, 296-300, 413-416):
 *   BCSR     u64 V, u64 E, u64 off[V], u32 dst[E]
 *   BWCSR    u64 V, u64 E, u64 off[V], {u32 dst, u32 w}[E]
 *   BCSC     u64 V, u64 E, u32 outdeg[V], u64 inoff[V], u32 src[E]
 *   LEGACY32 u32 V, u32 E, u32 off[V], u32 dst[E]          (bfs.cu:294-304)
 *   LEGACY32W u32 V, u32 E, u32 off[V], {u32 dst, u32 w}[E] (sssp.cu:176-186)
 * Offsets carry no off[V] sentinel; the last degree is E - off[V-1]
 * (GraphMeta.cuh:226).  ASC_FMT_AUTO tries, in the enum's order, every layout
 * whose size matches the header and keeps the first that validates (BWCSR
 * and BCSC have equal sizes when V == E). */
typedef enum {
    ASC_FMT_AUTO = 0,
    ASC_FMT_BCSR = 1,
    ASC_FMT_BWCSR = 2,
    ASC_FMT_BCSC = 3,
    ASC_FMT_LEGACY32 = 4,
    ASC_FMT_LEGACY32W = 5
} asc_format;

/* Graph kinds. */
typedef enum {
    ASC_KIND_CSR = 0,       /* out-CSR, unweighted (bfs, cc)                 */
    ASC_KIND_WCSR = 1,      /* out-CSR, AoS {dst,w} (sssp; bfs/cc too)        */
    ASC_KIND_CSC = 2        /* in-CSR + out-degree (pr)                      */
} asc_kind;

/* Value conventions of the integer outputs (bfs / sssp).
 * REFERENCE: source = 1, unreached = V+1 (GraphMeta.cuh:432-446).
 * ZERO_BASED: source = 0, unreached = 0xFFFFFFFF. */
typedef enum {
    ASC_VALUES_REFERENCE = 0,
    ASC_VALUES_ZERO_BASED = 1
} asc_value_convention;

/* ---- context ----------------------------------------------------------- */
typedef enum {
    ASC_STREAM_GATHER = 0,  /* host threads gather active on-demand edges into
                               pinned staging; double-buffered hipMemcpyAsync */
    ASC_STREAM_ZEROCOPY = 1,/* kernels read on-demand edges from mapped pinned
                               host memory directly                         */
    ASC_STREAM_HYBRID = 2   /* dense chunks copied whole by DMA (as in
                               GATHER); the remaining active edges read by
                               the kernels from mapped pinned memory        */
} asc_stream_mode;

/* BFS direction (asc_config.bfs_direction). */
typedef enum {
    ASC_BFS_AUTO = 0,        /* direction-optimizing (top-down / bottom-up
                                switching) when every edge is resident and
                                the adjacency is symmetric; else top-down    */
    ASC_BFS_TOP_DOWN = 1     /* always top-down frontier expansion          */
} asc_bfs_direction;

/* CC method (asc_config.cc_method).  Both give every vertex the smallest
 * vertex id that reaches it; on a symmetric graph that is the smallest id of
 * its connected component. */
typedef enum {
    ASC_CC_AUTO = 0,         /* union-find (sampled hooking + compression)
                                when every edge is resident and the adjacency
                                is symmetric; else label propagation        */
    ASC_CC_LABEL_PROP = 1    /* always min-label propagation                */
} asc_cc_method;

typedef struct {
    int device;                  /* HIP device ordinal                        */
    double advise_rate;          /* --adviseK: fraction of the device edge
                                    capacity used as static region; <= 0 means
                                    the reference's automatic rule
                                    (GraphMeta.cuh:326-339)                   */
    uint64_t edge_budget_bytes;  /* > 0: static region limited to this many
                                    bytes of edge records (overrides
                                    advise_rate); 0: derived from free memory */
    uint64_t part_bytes;         /* bytes of each on-demand part buffer (four
                                    on the device, two pinned staging buffers
                                    on the host); 0 = 64 MiB                  */
    uint64_t device_mem_limit;   /* > 0: pretend the device has this much free
                                    memory when sizing, and refuse (ASC_ERR_
                                    NOMEM) a graph whose planned device copies
                                    exceed it (testing); 0 = query           */
    int gather_threads;          /* host gather threads; 0 = min(16, cores)   */
    int stream_mode;             /* asc_stream_mode                           */
    int value_convention;        /* asc_value_convention                      */
    int heavy_threshold;         /* degree at/above which a frontier vertex is
                                    expanded grid-cooperatively; 0 = tuned    */
    int pr_group;                /* lanes per vertex in the PageRank pull
                                    (4,8,16,32,64); 0 = tuned                 */
    int sssp_delta;              /* SSSP bucket width (near-far); 0 = tuned,
                                    <0 = plain frontier Bellman-Ford          */
    int verbose;                 /* print per-phase timings like the reference */
    int profile;                 /* 1: time every relax kernel with HIP events
                                    on its launch stream (asc_stats.kern)    */
    int stream_pack;             /* 1 (default): stream weighted on-demand
                                    edges as one u32 word, dst | (w - wmin) <<
                                    bits(V-1), when both fit, else as a u32 id
                                    plus a byte w - wmin when the weights span
                                    at most 256 values (5 bytes); 2: that
                                    split form whenever the weights allow it;
                                    0: 8-byte records                        */
    int bfs_direction;           /* asc_bfs_direction                         */
    int cc_method;               /* asc_cc_method                             */
    double dense_threshold;      /* gather streaming: an on-demand chunk (a
                                    vertex range of at most one part) whose
                                    active edges reach this fraction of its
                                    edges is copied whole by DMA from the
                                    pinned tail instead of gathered; 0 = chosen
                                    per round (at least half active, then by a
                                    link/gather balance), > 1 = never        */
    int stream_overlap;          /* SSSP/CC: n >= 1 = expand the resident
                                    frontier after every n-th streamed part or
                                    chunk, while the next one is in flight
                                    (default 1); 0 = after the whole round  */
    int bfs_persistent;          /* direction-optimizing BFS: 1 = the whole
                                    traversal in one persistent cooperative
                                    kernel that decides each level's direction
                                    and the end itself (same levels); 0
                                    (default) = one host round trip per level,
                                    faster on MI355X because its kernels run
                                    at full occupancy and a grid barrier costs
                                    about as much as a round trip            */
    int release_host_edges;      /* 1: once the on-demand edges are kept in
                                    their packed pinned form, free the graph's
                                    host copy of the {dst, w} records (it is
                                    then no longer needed: the device holds
                                    the static prefix); the graph can then not
                                    be re-planned.  0 (default): keep it     */
    int part_buffers;            /* device buffers for streamed parts and
                                    chunks, each part_bytes (0 = 4; 2..16):
                                    more let more of the on-demand region be
                                    copied while the resident part is relaxed
                                    at the start of a run                    */
    int pad_cfg;
    uint64_t max_pinned_bytes;   /* largest page-locked on-demand tail (packed
                                    records) a graph takes; a larger one, or
                                    one the system refuses to lock (about 70
                                    GB on a 300 GB host), lives in pageable
                                    memory, and its dense chunks go through
                                    pinned staging (gather streaming only).
                                    0 = no limit                             */
} asc_config;

typedef struct asc_ctx asc_ctx;
typedef struct asc_graph asc_graph;

void asc_config_init(asc_config* cfg);
int asc_ctx_create(const asc_config* cfg, asc_ctx** out);
int asc_ctx_destroy(asc_ctx* ctx);
/* Number of visible HIP devices (0 when none); never fails. */
int asc_device_count(void);
/* Blocks until all work the library queued on the context's device ended. */
int asc_ctx_synchronize(asc_ctx* ctx);
/* Host-link probe: copies `bytes` between a pinned host buffer and the
 * device `reps` times in each direction (HIP events on the copy stream) and
 * reports GB/s; the roofline for streamed on-demand edges. */
int asc_measure_link(asc_ctx* ctx, uint64_t bytes, int reps, double* h2d_gbps, double* d2h_gbps);
/* Calibration kernel for hardware counters (no counterpart in the
 * reference): one launch of a plain device copy (kernel `k_copy_probe`) that
 * reads and writes exactly `bytes` each; *ms is its duration. */
int asc_copy_probe(asc_ctx* ctx, uint64_t bytes, double* ms);

/* ---- graphs ------------------------------------------------------------ */
/* Host view of a CSR/CSC graph.  Exactly one of off64/off32 is set; it holds
 * V entries (no sentinel) or V+1 entries (has_sentinel = 1).  For ASC_KIND_CSR
 * and ASC_KIND_CSC `edges` points to u32 ids; for ASC_KIND_WCSR to E pairs
 * {u32 dst, u32 w}.  For ASC_KIND_CSC out_degree holds V u32 values. */
typedef struct {
    uint64_t num_vertices;
    uint64_t num_edges;
    int kind;                    /* asc_kind                                 */
    int has_sentinel;
    const uint64_t* off64;
    const uint32_t* off32;
    const uint32_t* edges;
    const uint32_t* out_degree;
} asc_csr_view;

/* Reads a graph file (GraphMeta.cuh:151-171).  The file is validated against
 * its header by size; offsets must be non-decreasing and <= E, ids < V. */
int asc_graph_load(asc_ctx* ctx, const char* path, int format, asc_graph** out);
/* Builds a graph from caller arrays (they are copied; the caller keeps them). */
int asc_graph_from_csr(asc_ctx* ctx, const asc_csr_view* view, asc_graph** out);
/* The same graph without copying the caller's records: the edge array, the
 * out-degrees and u64 offsets with their sentinel are borrowed (narrower
 * offsets are widened into a copy), and the caller keeps them alive and
 * unchanged until asc_graph_destroy.  Host memory the library holds is then
 * the offsets at most, plus the packed on-demand copy when a budget streams
 * part of the graph.  Replaces the reference's host arrays filled by
 * readDataFromFile (GraphMeta.cuh:151-171) for callers that already hold the
 * graph. */
int asc_graph_borrow_csr(asc_ctx* ctx, const asc_csr_view* view, asc_graph** out);
int asc_graph_destroy(asc_graph* g);

typedef struct {
    uint64_t num_vertices;
    uint64_t num_edges;
    int kind;
    int offset_bytes;            /* 4 or 8: width of offsets on the device    */
    uint64_t static_edges;       /* edges in the device-resident prefix       */
    uint64_t static_vertices;    /* vertices [0, static_vertices) are static  */
    uint64_t ondemand_edges;     /* edges left to stream                      */
    uint64_t part_edges;         /* edges per on-demand part buffer           */
    uint64_t device_bytes;       /* device memory the graph holds             */
    uint32_t max_degree;
    uint32_t max_degree_vertex;
    int symmetric;               /* 1: every u->v has a v->u of the same
                                    multiplicity; 0: not; -1: not checked
                                    (checked on upload when fully resident)  */
    int tail_pinned;             /* 1: the on-demand records are page-locked
                                    (DMA and zero-copy reads straight from
                                    them); 0: pageable or no tail            */
    uint64_t host_bytes;         /* host memory the graph holds: offsets, the
                                    edge records unless released, pinned
                                    on-demand copies, out-degrees            */
} asc_graph_info;
int asc_graph_get_info(const asc_graph* g, asc_graph_info* info);

/* Writes a graph file in one of the binary layouts. */
int asc_write_graph(const char* path, int format, const asc_csr_view* view);

/* Re-plans the static / on-demand split of a loaded graph with a new byte
 * budget (0 = automatic rule) or advise rate; uploads the new static prefix.
 * The budget and rate apply to this graph only (graphs created later in the
 * context keep the context's).  A plan whose device copies cannot fit is
 * refused before anything changes (the graph keeps its current split); if a
 * HIP allocation or copy fails after the old copies were freed, every later
 * call on the graph returns ASC_ERR_STATE until a re-plan succeeds.  Refused
 * with ASC_ERR_STATE while a server runs over the graph (so is
 * asc_graph_destroy). */
int asc_graph_replan(asc_graph* g, uint64_t edge_budget_bytes, double advise_rate);

/* ---- per-run statistics (phase names follow CalculateOpt.cu:30-36) ------ */
typedef struct {
    uint32_t launches;
    uint32_t pad;
    double ms;                   /* summed launch durations (HIP events)      */
    uint64_t edges;              /* edge records relaxed                      */
    uint64_t vertices;           /* frontier / list entries expanded          */
} asc_kernel_stats;

typedef struct {
    double total_ms;             /* first iteration start .. last end         */
    double pre_ms;               /* host-side frontier bookkeeping            */
    double static_ms;            /* resident relax kernels (HIP events)       */
    double overload_ms;          /* on-demand phase wall time                 */
    double overload_move_ms;     /* H2D copies of on-demand parts (events)    */
    double fill_ms;              /* host gather of on-demand edges            */
    uint32_t iterations;
    uint32_t parts;              /* on-demand parts shipped                   */
    uint64_t node_sum;           /* sum over iterations of active vertices    */
    uint64_t edges_traversed;    /* edges relaxed (static + on-demand)        */
    uint64_t static_edges_traversed;
    uint64_t streamed_edges;     /* on-demand edges shipped over the link     */
    uint64_t streamed_bytes;
    uint64_t dense_edges;        /* of streamed_edges: copied as whole chunks
                                    by DMA (asc_config.dense_threshold)       */
    uint64_t vertices_processed; /* frontier entries expanded                 */
    uint64_t reached;            /* vertices with a final value != unreached  */
    uint64_t reached_degree_sum; /* sum of out-degrees of reached vertices    */
    double kernel_ms;            /* sum of relax-kernel durations (events)    */
    double pr_residual;          /* PageRank: last iteration's L1 residual    */
    int host_cores;
    int path;                    /* 0 frontier expansion, 1 direction-
                                    optimizing BFS, 2 union-find CC          */
    /* Per kernel kind; durations filled when asc_config.profile = 1, edge
     * and vertex counts always:
     * [0] resident frontier expansion, [1] grid-cooperative heavy vertices,
     * [2] on-demand parts read from a device part buffer, [3] PageRank
     * pull, [4] BFS bottom-up steps, [5] union-find link passes, [6] the
     * persistent BFS kernel (whole traversal; its edges and vertices are
     * counted under [0], [1], [4]), [7] on-demand parts read straight from
     * mapped host memory (zero-copy and hybrid streaming, small rounds). */
    asc_kernel_stats kern[8];
    uint64_t reused_edges;       /* on-demand edges read again from a chunk an
                                    earlier round of the run left in a part
                                    buffer (no transfer; not in streamed_edges) */
} asc_stats;

/* ---- algorithms (CalculateOpt.cu:14-721) ------------------------------- */
int asc_bfs(asc_graph* g, uint32_t source, uint32_t* level_out, asc_stats* st);
int asc_sssp(asc_graph* g, uint32_t source, uint32_t* dist_out, asc_stats* st);
int asc_cc(asc_graph* g, uint32_t* label_out, asc_stats* st);

/* BFS that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory at call time (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same levels as asc_bfs (value convention of the context), written into
 * d_level_out, a DEVICE buffer of V u32 values owned by the caller; the
 * frontier queues, visited bitmap and counters live in one statically sized
 * device array per GPU (about 17 MB, set up by asc_ctx_create), so the call
 * makes no hipMalloc / hipFree.  Limits: the graph must be fully resident
 * (ASC_ERR_UNSUPPORTED otherwise) with V <= 2^21 (ASC_ERR_RANGE above).
 * `stream`: a hipStream_t to run on, or NULL for the context's stream; the
 * call returns after its work on that stream ended.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int asc_bfs_noalloc(asc_graph* g, uint32_t source, uint32_t* d_level_out, void* stream);

typedef enum {
    ASC_PR_NORMALISED = 0,       /* r = (1-d)/V + d*(sum r/outdeg + dangling/V),
                                    fixed iterations, FP32 ranks, FP64 residual */
    ASC_PR_REFERENCE = 1         /* 0.15 + 0.85*sum, active set with threshold,
                                    FP64 (gpu_kernels.cuh:533-594)             */
} asc_pr_mode;

typedef struct {
    int mode;                    /* asc_pr_mode                               */
    double damping;              /* 0.85                                      */
    int iterations;              /* normalised mode: fixed count (20)         */
    double threshold;            /* reference mode: 0.001                     */
    int max_iterations;          /* reference mode safety cap (0 = 10000)     */
    int method;                  /* normalised mode: asc_pr_method            */
    /* ASC_PR_TILED tile shape (0 = tuned): sources per segment 2^src_bits,
     * destinations per block 2^dst_bits (<= 14), src_bits + dst_bits <= 32 */
    int tile_src_bits;
    int tile_dst_bits;
} asc_pr_opts;

/* Normalised PageRank schedule (asc_pr_opts.method). */
typedef enum {
    ASC_PR_AUTO = 0,     /* ASC_PR_TILED when every edge is resident and
                            V >= 2^20 (contributions larger than an XCD's
                            L2), else the pull                              */
    ASC_PR_PULL = 1,     /* G-lane pull over the in-CSR                      */
    ASC_PR_BLOCKED = 2,  /* propagation-blocked (scatter to slots by
                            destination bin, LDS sums per bin); needs every
                            edge resident, else ASC_ERR_UNSUPPORTED         */
    ASC_PR_TILED = 3     /* source-segmented: in-edges regrouped once into
                            (destination block, source segment) tiles, a
                            workgroup sums one block in LDS while walking
                            the segments, whose contribution slices stay in
                            L2; needs every edge resident and V <= 2^31,
                            else ASC_ERR_UNSUPPORTED                        */
} asc_pr_method;
void asc_pr_opts_init(asc_pr_opts* o);

/* rank_f32: V floats (normalised mode); rank_f64: V doubles (reference mode);
 * residual_out (optional): per-iteration L1 residuals, `iterations` entries. */
int asc_pr(asc_graph* g, const asc_pr_opts* opts, float* rank_f32, double* rank_f64,
           double* residual_out, asc_stats* st);

/* ---- static / on-demand planning (GraphMeta.cuh:310-372, 516-578) ------- */
typedef struct {
    uint64_t total_device_edges; /* (free - param_size*4*V) / edge_size        */
    double advise_k;             /* rate actually used                        */
    uint64_t static_edges;       /* static prefix length in edges             */
    uint64_t static_vertices;    /* vertices whose adjacency is all static    */
    uint64_t ondemand_edges;     /* E - sum of static degrees                 */
    uint64_t part_edges;         /* on-demand part capacity in edges          */
} asc_plan;

/* Pure host function: plans the split for a graph with the given offsets.
 * free_bytes: device memory available; param_size: 13 (bfs, cc), 15 (sssp),
 * 17 (pr) as in CalculateOpt.cu:20,214,389,552. */
int asc_plan_partition(uint64_t V, uint64_t E, const uint64_t* off64, uint64_t edge_size,
                       uint64_t free_bytes, int param_size, double advise_rate,
                       uint64_t edge_budget_bytes, uint64_t part_bytes, asc_plan* out);

/* Edge-granular part split of an on-demand list: given prefix[0..n] (u64
 * exclusive starts of each active vertex's adjacency in the gathered list,
 * prefix[n] = total) and a part capacity in edges, writes the edge start of
 * each part into part_starts (n_parts+1 entries, last = total).  Unlike the
 * reference's vertex-granular split, a vertex larger than a part spans
 * several parts.  Returns the number of parts via *n_parts. */
int asc_split_parts(const uint64_t* prefix, uint64_t n, uint64_t part_edges,
                    uint64_t* part_starts, uint64_t cap, uint64_t* n_parts);

/* ---- synthetic graphs -------------------------------------------------- */
typedef struct {
    int type;                    /* 0 = R-MAT, 1 = Erdos-Renyi                */
    int scale;                   /* R-MAT: V = 2^scale                        */
    uint64_t num_vertices;       /* ER: V                                     */
    uint64_t num_samples;        /* ER: undirected pairs drawn; R-MAT: 0 means
                                    edge_factor * V                           */
    int edge_factor;             /* R-MAT                                     */
    double a, b, c;              /* R-MAT quadrant probabilities (d = 1-a-b-c) */
    uint64_t seed;
    int symmetrize;              /* add (v,u) for every (u,v)                 */
    int weighted;                /* AoS {dst,w}; w uniform in [wmin,wmax],
                                    w(u,v) = w(v,u)                           */
    uint32_t wmin, wmax;
    int offset_bytes;            /* 4 or 8: offsets the graph keeps on device */
    int csc;                     /* also produce out-degree (for PageRank)    */
    int threads;                 /* 0 = all cores                             */
} asc_gen_params;
void asc_gen_params_init(asc_gen_params* p);

/* Host CSR owned by the library (free with asc_host_csr_free). */
typedef struct {
    uint64_t num_vertices;
    uint64_t num_edges;
    int kind;
    uint64_t* off;               /* V+1 entries                               */
    uint32_t* edges;             /* E ids, or 2E words for weighted AoS       */
    uint32_t* out_degree;        /* V entries when kind == ASC_KIND_CSC       */
} asc_host_csr;

/* Generates a graph: self-loops and duplicate edges removed, adjacency sorted. */
int asc_generate(const asc_gen_params* p, asc_host_csr* out);
/* The same graph as asc_generate, bit for bit, built on the context's GPU
 * (raw degrees, per vertex range the raw pairs as sort keys, radix sort,
 * unique, records) and returned in host memory owned by the caller (free
 * with asc_host_csr_free).  The in-CSR (csc) only of symmetrized graphs
 * (ASC_ERR_UNSUPPORTED otherwise).  No counterpart in the reference, which
 * reads its graphs from files; it makes the large synthetic inputs of the
 * tests and benchmarks in seconds. */
int asc_generate_device(asc_ctx* ctx, const asc_gen_params* p, asc_host_csr* out);

/* ---- edge-list conversion (converter/converter.cpp:22-420) -------------- */
typedef struct {
    uint64_t num_vertices;       /* 0: one more than the largest id           */
    int symmetrize;              /* add v->u for every u->v                   */
    int dedup;                   /* keep one copy of each (u, v), the lightest */
    int keep_self_loops;
    int one_based;               /* ids start at 1 (set by a Matrix Market
                                    header by itself)                         */
    int weighted;                /* {dst, w} records: the input's third column,
                                    else w(u,v) = w(v,u) drawn in [wmin, wmax] */
    int csc;                     /* in-CSR of sources + out-degrees (PageRank) */
    uint32_t wmin, wmax;
    uint64_t seed;
    int threads;                 /* 0 = all cores                             */
    int pad;
} asc_edge_list_opts;
void asc_edge_list_opts_init(asc_edge_list_opts* o);
/* Reads a text edge list ("u v [w]" per line, separated by blanks, tabs or
 * commas; lines starting with '#' or '%' skipped), a Matrix Market file, or
 * binary {u32 src, u32 dst} pairs (a path ending in ".toB") into a host CSR
 * with sorted adjacency; asc_write_graph stores it as .bcsr / .bwcsr /
 * .bcsc.  Replaces the reference's two-step converter (text -> .toB ->
 * binary CSR), without its defects (V from sources only, shifted offsets
 * of trailing empty vertices, zero weights). */
int asc_read_edge_list(const char* path, const asc_edge_list_opts* o, asc_host_csr* out);
/* Reads a graph file into host memory only (no device needed); same
 * validation as asc_graph_load (GraphMeta.cuh:151-171). */
int asc_read_graph(const char* path, int format, asc_host_csr* out);
void asc_host_csr_free(asc_host_csr* h);
/* View over a host CSR (for asc_graph_from_csr / asc_write_graph). */
void asc_host_csr_view(const asc_host_csr* h, asc_csr_view* v);
/* Builds a graph directly from a generated CSR, taking ownership of its arrays
 * (no copy; *h is emptied). */
int asc_graph_adopt(asc_ctx* ctx, asc_host_csr* h, int offset_bytes, asc_graph** out);
/* Generates a weighted graph (the CSR asc_generate builds for the same
 * parameters) straight into a device graph: the static prefix the context's
 * budget allows is written on the GPU, the on-demand edges go to pinned host
 * memory in their packed 4-byte or split 5-byte form, and the host never
 * holds the 8-byte {dst, w} records; at R-MAT scale 30 with edge factor 16
 * (34 G directed edges) that is 8 GB of offsets and ~110 GB of records in
 * host memory instead of 275 GB.  Two sweeps over vertex ranges on the GPU
 * (degrees, then records); p->offset_bytes = 8 keeps 8-byte device offsets
 * (forced when E >= 2^32).  The graph cannot be re-planned (its host records
 * do not exist).  ASC_ERR_UNSUPPORTED for unweighted or in-CSR parameters,
 * or with stream_pack 0.  The reference reads its large graphs from files
 * (GraphMeta.cuh:151-171); this replaces writing and reading one. */
int asc_graph_generate(asc_ctx* ctx, const asc_gen_params* p, asc_graph** out);

/* ---- batched query serving --------------------------------------------- */
/* No counterpart in the reference (which runs one traversal per process).
 * A server keeps up to `lanes` (<= 64) BFS/SSSP queries in flight over one
 * resident graph: each iteration of a multi-source kernel advances every
 * admitted query by one frontier step; queries that finish free their lane
 * and waiting queries are admitted at the next iteration boundary, subject
 * to the device memory their state needs. */
typedef enum { ASC_QUERY_BFS = 0, ASC_QUERY_SSSP = 1 } asc_query_kind;

typedef struct {
    int lanes;                   /* max queries in flight per kind (1..64);
                                    0 = the graph's tuned width (64 untuned) */
    uint64_t state_budget_bytes; /* device bytes for per-query state (V bytes
                                    per BFS lane, 2 V per SSSP lane); 0 = half
                                    the free memory                           */
    int queue_capacity;          /* max submitted-but-unfinished queries      */
    int bfs_direction;           /* asc_bfs_direction: AUTO lets a BFS lane
                                    whose frontier is large take bottom-up
                                    steps on a symmetric graph               */
    int step_graphs;             /* 1 (default): each engine step replays the
                                    step's launch sequence recorded once into
                                    a hipGraph; 0: launches every kernel      */
    int pad;
} asc_server_opts;
void asc_server_opts_init(asc_server_opts* o);

typedef struct asc_server asc_server;

typedef struct {
    double submit_to_admit_ms;
    double submit_to_done_ms;
    uint32_t iterations;         /* engine steps the query was in flight      */
    uint32_t lane;
} asc_qstats;

typedef struct {
    uint64_t submitted, completed;
    uint64_t steps;              /* engine steps run                          */
    int lanes_bfs, lanes_sssp;   /* lanes granted by the memory budget        */
    double busy_ms;              /* wall time with at least one query in flight */
    uint64_t bottom_up_steps;    /* BFS engine steps with bottom-up lanes      */
    uint64_t fallbacks;          /* queries whose BFS levels left 8 bits or SSSP
                                    distances 16 bits, answered by the
                                    single-query engine                      */
    uint64_t graph_steps;        /* engine steps replayed from a hipGraph     */
} asc_server_stats;

/* Starts a background thread that admits and advances queries. */
int asc_server_create(asc_graph* g, const asc_server_opts* opts, asc_server** out);
int asc_server_destroy(asc_server* s);
/* Thread-safe.  Queues a BFS or SSSP query from `source`.  When `out` is not
 * NULL it must hold V values; it is filled (reference convention: source 1,
 * unreached V+1) before the query completes.  *ticket identifies the query.
 * Returns ASC_ERR_BUSY when queue_capacity queries are outstanding. */
int asc_submit(asc_server* s, int kind, uint32_t source, uint32_t* out, uint64_t* ticket);
/* Thread-safe.  Blocks until the query completed and releases the ticket.
 * One thread waits per ticket: a second asc_wait on a ticket that another
 * thread is waiting for returns ASC_ERR_INVALID, as does a released ticket. */
int asc_wait(asc_server* s, uint64_t ticket, asc_qstats* qs);
/* Blocks until every submitted query completed. */
int asc_server_drain(asc_server* s);
int asc_server_get_stats(asc_server* s, asc_server_stats* st);

/* ---- autotuning -------------------------------------------------------- */
typedef struct {
    int heavy_threshold;         /* degree from which a frontier vertex is
                                    expanded grid-cooperatively              */
    int blocks_per_cu;           /* workgroups per CU of traversal launches   */
    int pr_group;                /* lanes per vertex of the PageRank pull     */
    int sssp_delta;              /* SSSP band width (-1 = no far pile)        */
    int serve_lanes;             /* serving batch width                       */
    int pr_tile_src_bits;        /* PageRank tiled pull: log2 of the source
                                    segment (ASC_PR_TILED)                   */
    double score_ms;             /* summed time of the sampled workload with
                                    the chosen values                         */
    double untuned_ms;           /* the same workload with the defaults       */
    uint64_t fingerprint;        /* graph the values were tuned on            */
} asc_tuning;
/* Sweeps the tunables on sampled sources of g (coordinate descent, one
 * parameter at a time), stores the best values in the graph and in *out.
 * `budget_ms` bounds the sweep's wall time (0 = 30 s). */
int asc_autotune(asc_graph* g, double budget_ms, asc_tuning* out);
/* Tuning tables: one line per graph fingerprint, "key value" pairs.  Save
 * replaces the line of this graph; load applies the line matching the
 * graph's fingerprint or returns ASC_ERR_STATE when there is none. */
int asc_tuning_save(const char* path, const asc_graph* g, const asc_tuning* t);
int asc_tuning_load(const char* path, asc_graph* g, asc_tuning* t);
/* Applies explicit values to a graph (0 keeps a field unchanged). */
int asc_tuning_apply(asc_graph* g, const asc_tuning* t);

#ifdef __cplusplus
}
#endif
#endif /* ASCETIC_H */
