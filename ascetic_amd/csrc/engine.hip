// Context and graph lifecycle, device upload of the static region, and the
// algorithm entry points of the C ABI.
//
// Reference: GraphMeta::initGraphHost / initGraphDevice (GraphMeta.cuh:
// 208-297) and the drivers bfs_opt / cc_opt / sssp_opt / pr_opt
// (CalculateOpt.cuh:11-14).  The reference allocates ~12 V-sized device
// arrays and frees none; here a graph owns its device copies and frees them
// in asc_graph_destroy, and traversal state is allocated once per graph and
// reused by later runs.
#include <algorithm>

#include "internal.hpp"
#include "traverse.hpp"

namespace asc {
int pagerank(asc_graph* g, const asc_pr_opts* o, float* rank_f32, double* rank_f64, double* residual_out,
             asc_stats* st);

void graph_free_device(asc_graph* g) {
    if (g->d_off) (void)hipFree(g->d_off);
    if (g->d_edges) (void)hipFree(g->d_edges);
    if (g->d_outdeg) (void)hipFree(g->d_outdeg);
    g->d_off = g->d_edges = nullptr;
    g->d_outdeg = nullptr;
    g->device_bytes = 0;
    if (g->trav) trav_free(g->trav);
    g->trav = nullptr;
    if (g->pr) pr_free(g->pr);
    g->pr = nullptr;
    g->zc_tail.release();
    g->ptail.release();
    g->ptail_w.release();
    g->packed = false;
    g->pk_split = false;
}

int param_size_for(int kind) {
    // paramSize of CalculateOpt.cu:20 (bfs), :389 (sssp), :552 (pr)
    return kind == ASC_KIND_WCSR ? 15 : kind == ASC_KIND_CSC ? 17 : 13;
}

int graph_ready(const asc_graph* g) {
    if (g->usable) return ASC_OK;
    set_last_error("the graph lost its device copies in a failed re-plan; re-plan it again");
    return ASC_ERR_STATE;
}

// Plans the split with the graph's budget and uploads the static prefix.  The
// plan and its device size are checked before anything is freed, so a plan
// that cannot fit (asc_config.device_mem_limit caps the device bytes of one
// graph) leaves the current copies in place; a HIP allocation or copy that
// fails after the old copies were freed leaves the graph unusable.
int graph_upload(asc_graph* g) {
    asc_ctx* ctx = g->ctx;
    ASC_HIP(hipSetDevice(ctx->device));
    const uint64_t V = g->h.V, E = g->h.E;
    const size_t EB = g->h.edge_bytes();
    size_t free_b = 0, total_b = 0;
    ASC_HIP(hipMemGetInfo(&free_b, &total_b));
    // the graph's current device copies return to the pool when it is re-planned
    free_b += g->device_bytes;
    if (ctx->cfg.device_mem_limit) free_b = std::min<size_t>(free_b, ctx->cfg.device_mem_limit);
    asc_plan plan;
    ASC_TRY(asc_plan_partition(V, E, g->h.off.p, EB, free_b, param_size_for(g->h.kind), g->advise_rate,
                               g->edge_budget_bytes, ctx->cfg.part_bytes, &plan));
    const uint64_t need = (uint64_t)g->off_bytes * (V + 1) + std::max<uint64_t>(EB * plan.static_edges, 64) +
                          (g->h.kind == ASC_KIND_CSC ? 4 * V : 0);
    if (ctx->cfg.device_mem_limit && need > ctx->cfg.device_mem_limit) {
        set_last_error("the planned device copies need %llu bytes, above the %llu-byte device limit",
                       (unsigned long long)need, (unsigned long long)ctx->cfg.device_mem_limit);
        return ASC_ERR_NOMEM;
    }
    graph_free_device(g);
    g->usable = false;
    g->static_edges = plan.static_edges;
    g->static_vertices = plan.static_vertices;
    g->part_edges = plan.part_edges;
    g->advise_k = plan.advise_k;

    // offsets
    const size_t ob = g->off_bytes;
    hipError_t e = hipMalloc(&g->d_off, ob * (V + 1));
    if (e != hipSuccess) {
        set_last_error("device allocation of %zu offset bytes failed", ob * (V + 1));
        return ASC_ERR_NOMEM;
    }
    if (ob == 8) {
        ASC_HIP(hipMemcpy(g->d_off, g->h.off.p, 8 * (V + 1), hipMemcpyHostToDevice));
    } else {
        HostBuf<uint32_t> tmp;
        ASC_TRY(tmp.alloc(V + 1));
        host_pool().parallel_for(V + 1, [&](uint64_t b, uint64_t en) {
            for (uint64_t i = b; i < en; i++) tmp[i] = (uint32_t)g->h.off[i];
        });
        ASC_HIP(hipMemcpy(g->d_off, tmp.p, 4 * (V + 1), hipMemcpyHostToDevice));
    }
    // static edge prefix (GraphMeta.cuh:254-255)
    const size_t sbytes = std::max<size_t>(EB * g->static_edges, 64);
    e = hipMalloc(&g->d_edges, sbytes);
    if (e != hipSuccess) {
        set_last_error("device allocation of %zu static edge bytes failed", sbytes);
        return ASC_ERR_NOMEM;
    }
    if (g->static_edges)
        ASC_HIP(hipMemcpy(g->d_edges, g->h.edges.p, EB * g->static_edges, hipMemcpyHostToDevice));
    g->device_bytes = ob * (V + 1) + sbytes;
    if (g->h.kind == ASC_KIND_CSC) {
        ASC_HIP(hipMalloc(&g->d_outdeg, 4 * V));
        ASC_HIP(hipMemcpy(g->d_outdeg, g->h.outdeg.p, 4 * V, hipMemcpyHostToDevice));
        g->device_bytes += 4 * V;
    }
    // packed on-demand stream: one u32 per edge when the id and weight fit,
    // else a u32 id and a weight byte when the weights span at most 256 values
    // (stream_pack 2 forces the split form)
    g->packed = false;
    g->pk_split = false;
    if (ctx->cfg.stream_pack && g->h.kind == ASC_KIND_WCSR && g->static_edges < E) {
        const uint64_t S = g->static_edges, n = E - S;
        const uint32_t* ew = g->h.edges.p;
        std::mutex mu;
        uint32_t wlo = 0xFFFFFFFFu, whi = 0;
        host_pool().parallel_for(n, [&](uint64_t b, uint64_t en) {
            uint32_t lo = 0xFFFFFFFFu, hi = 0;
            for (uint64_t i = b; i < en; i++) {
                const uint32_t w = ew[2 * (S + i) + 1];
                lo = std::min(lo, w);
                hi = std::max(hi, w);
            }
            std::lock_guard<std::mutex> lk(mu);
            wlo = std::min(wlo, lo);
            whi = std::max(whi, hi);
        });
        const uint32_t dbits = V > 1 ? 64 - (uint32_t)__builtin_clzll(V - 1) : 1;
        const uint32_t span = whi - wlo;
        const uint32_t wbits = span ? 32 - (uint32_t)__builtin_clz(span) : 0;
        const uint64_t pb = ctx->cfg.part_bytes ? ctx->cfg.part_bytes : (uint64_t(64) << 20);
        if (dbits + wbits <= 32 && dbits < 32 && ctx->cfg.stream_pack != 2) {
            ASC_TRY(g->ptail.alloc_prefer_pinned(n, ctx->cfg.max_pinned_bytes));
            host_pool().parallel_for(n, [&](uint64_t b, uint64_t en) {
                for (uint64_t i = b; i < en; i++)
                    g->ptail[i] = ew[2 * (S + i)] | ((ew[2 * (S + i) + 1] - wlo) << dbits);
            });
            g->packed = true;
            g->pk_shift = dbits;
            g->pk_wmin = wlo;
            g->part_edges = std::max<uint64_t>(pb / 4, 1024);
        } else if (wbits <= 8) {
            // ids and weight bytes are page-locked together (5 bytes an edge
            // against the cap) or not at all
            const uint64_t cap = ctx->cfg.max_pinned_bytes;
            ASC_TRY(!cap || 5 * n <= cap ? g->ptail.alloc_prefer_pinned(n, 0) : g->ptail.alloc(n, false));
            ASC_TRY(g->ptail.pinned ? g->ptail_w.alloc_prefer_pinned(n, 0) : g->ptail_w.alloc(n, false));
            host_pool().parallel_for(n, [&](uint64_t b, uint64_t en) {
                for (uint64_t i = b; i < en; i++) {
                    g->ptail[i] = ew[2 * (S + i)];
                    g->ptail_w[i] = (uint8_t)(ew[2 * (S + i) + 1] - wlo);
                }
            });
            g->packed = true;
            g->pk_split = true;
            g->pk_shift = 0;
            g->pk_wmin = wlo;
            g->part_edges = std::max<uint64_t>(pb / 5, 1024);
        }
    }
    // zero-copy streaming reads the on-demand edges from mapped pinned memory;
    // PageRank streams whole vertex-range chunks of them by DMA every iteration;
    // gather streaming copies dense chunks whole by DMA, which needs pinned memory
    const bool pr_tail = g->h.kind == ASC_KIND_CSC;
    const bool dense = ctx->cfg.stream_mode == ASC_STREAM_HYBRID ||
                       (ctx->cfg.stream_mode == ASC_STREAM_GATHER && ctx->cfg.dense_threshold <= 1.0);
    if ((ctx->cfg.stream_mode == ASC_STREAM_ZEROCOPY || pr_tail || dense) && g->static_edges < E && !g->packed) {
        const uint64_t words = (E - g->static_edges) * g->h.edge_words();
        ASC_TRY(g->zc_tail.alloc(words, true));
        const uint32_t* src = g->h.edges.p + g->static_edges * g->h.edge_words();
        host_pool().parallel_for(words, [&](uint64_t b, uint64_t en) {
            std::memcpy(g->zc_tail.p + b, src + b, 4 * (en - b));
        });
    }
    // the host {dst, w} records are redundant once the on-demand edges sit in
    // their packed pinned form and the static prefix on the device: free them
    // on request (what later calls read from them is derived first)
    if (ctx->cfg.release_host_edges && g->packed && g->h.edges.p) {
        compute_mean_weight(g);
        g->fp = fingerprint(g);
        g->h.edges.release();
        g->host_edges_released = true;
    }
    // symmetry decides the resident fast paths (direction-optimizing BFS,
    // union-find CC); checked on the device copy when every edge is on it
    g->symmetric = -1;
    if (g->static_vertices == V) ASC_TRY(check_symmetric(g));
    g->usable = true;
    return ASC_OK;
}

void compute_max_degree(asc_graph* g) {
    uint32_t best = 0, at = 0;
    for (uint64_t v = 0; v < g->h.V; v++) {
        uint64_t d = g->h.off[v + 1] - g->h.off[v];
        if (d > best) {
            best = (uint32_t)d;
            at = (uint32_t)v;
        }
    }
    g->max_degree = best;
    g->max_degree_vertex = at;
}

static int finish_graph(asc_ctx* ctx, asc_graph* g, asc_graph** out) {
    g->ctx = ctx;
    g->edge_budget_bytes = ctx->cfg.edge_budget_bytes;
    g->advise_rate = ctx->cfg.advise_rate;
    if (g->h.E >= (uint64_t(1) << 32)) g->off_bytes = 8;
    g->heavy_threshold = ctx->cfg.heavy_threshold;
    g->pr_group = ctx->cfg.pr_group;
    g->sssp_delta = ctx->cfg.sssp_delta;
    compute_max_degree(g);
    int r = graph_upload(g);
    if (r != ASC_OK) {
        graph_free_device(g);
        delete g;
        return r;
    }
    *out = g;
    return ASC_OK;
}

}  // namespace asc

namespace asc {
__global__ void k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
}  // namespace asc

using namespace asc;

extern "C" {

int asc_copy_probe(asc_ctx* ctx, uint64_t bytes, double* ms) {
    if (!ctx || bytes < 16) return ASC_ERR_INVALID;
    ASC_HIP(hipSetDevice(ctx->device));
    const uint64_t n = bytes / 16;
    DevTmp<uint4> src, dst;
    if (hipMalloc(&src.p, 16 * n) != hipSuccess || hipMalloc(&dst.p, 16 * n) != hipSuccess) return ASC_ERR_NOMEM;
    ScopedEvent a, b;
    ASC_HIP(hipEventCreate(&a.e));
    ASC_HIP(hipEventCreate(&b.e));
    ASC_HIP(hipMemsetAsync(src.p, 1, 16 * n, ctx->s_main));
    ASC_HIP(hipEventRecord(a.e, ctx->s_main));
    hipLaunchKernelGGL(k_copy_probe, dim3(ctx->num_cus * 8), dim3(256), 0, ctx->s_main, (const uint4*)src.p, dst.p,
                       n);
    ASC_HIP(hipGetLastError());
    ASC_HIP(hipEventRecord(b.e, ctx->s_main));
    ASC_HIP(hipEventSynchronize(b.e));
    float t = 0;
    (void)hipEventElapsedTime(&t, a.e, b.e);
    if (ms) *ms = t;
    return ASC_OK;
}

int asc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int asc_ctx_create(const asc_config* cfg, asc_ctx** out) {
    if (!out) return ASC_ERR_INVALID;
    *out = nullptr;
    asc_config c;
    if (cfg) c = *cfg; else asc_config_init(&c);
    int n = asc_device_count();
    if (n <= 0) {
        set_last_error("no HIP device is visible");
        return ASC_ERR_NOGPU;
    }
    if (c.device < 0 || c.device >= n) return ASC_ERR_INVALID;
    ASC_HIP(hipSetDevice(c.device));
    hipDeviceProp_t prop;
    ASC_HIP(hipGetDeviceProperties(&prop, c.device));
    asc_ctx* ctx = new asc_ctx();
    ctx->cfg = c;
    ctx->device = c.device;
    ctx->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 1;
    ctx->total_mem = prop.totalGlobalMem;
    if (hipStreamCreateWithFlags(&ctx->s_main, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->s_copy, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        set_last_error("stream creation failed");
        return ASC_ERR_DEVICE;
    }
    int nt = c.gather_threads > 0 ? c.gather_threads : std::min(16, hardware_threads());
    ctx->pool = new ThreadPool(nt);
    const int rn = noalloc_init(c.device);   // once per device, with the device current
    if (rn != ASC_OK) {
        asc_ctx_destroy(ctx);
        return rn;
    }
    *out = ctx;
    return ASC_OK;
}

int asc_ctx_destroy(asc_ctx* ctx) {
    if (!ctx) return ASC_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    if (ctx->s_main) (void)hipStreamDestroy(ctx->s_main);
    if (ctx->s_copy) (void)hipStreamDestroy(ctx->s_copy);
    delete ctx->pool;
    delete ctx;
    return ASC_OK;
}

int asc_ctx_synchronize(asc_ctx* ctx) {
    if (!ctx) return ASC_ERR_INVALID;
    ASC_HIP(hipSetDevice(ctx->device));
    ASC_HIP(hipDeviceSynchronize());
    return ASC_OK;
}

int asc_measure_link(asc_ctx* ctx, uint64_t bytes, int reps, double* h2d_gbps, double* d2h_gbps) {
    if (!ctx || bytes == 0 || reps <= 0) return ASC_ERR_INVALID;
    ASC_HIP(hipSetDevice(ctx->device));
    HostBuf<char> host;
    ASC_TRY(host.alloc(bytes, true));
    std::memset(host.p, 1, bytes);
    DevTmp<char> dev;
    if (hipMalloc(&dev.p, bytes) != hipSuccess) return ASC_ERR_NOMEM;
    ScopedEvent a, b;
    int rc = ASC_OK;
    if (hipEventCreate(&a.e) != hipSuccess || hipEventCreate(&b.e) != hipSuccess) rc = ASC_ERR_DEVICE;
    for (int dir = 0; dir < 2 && rc == ASC_OK; dir++) {
        auto copy = [&] {
            return dir == 0 ? hipMemcpyAsync(dev.p, host.p, bytes, hipMemcpyHostToDevice, ctx->s_copy)
                            : hipMemcpyAsync(host.p, dev.p, bytes, hipMemcpyDeviceToHost, ctx->s_copy);
        };
        if (copy() != hipSuccess) rc = ASC_ERR_DEVICE;  // warm-up
        (void)hipEventRecord(a.e, ctx->s_copy);
        for (int i = 0; i < reps && rc == ASC_OK; i++)
            if (copy() != hipSuccess) rc = ASC_ERR_DEVICE;
        (void)hipEventRecord(b.e, ctx->s_copy);
        (void)hipEventSynchronize(b.e);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a.e, b.e);
        const double gbps = ms > 0 ? (double)bytes * reps / (ms * 1e-3) / 1e9 : 0;
        if (dir == 0 && h2d_gbps) *h2d_gbps = gbps;
        if (dir == 1 && d2h_gbps) *d2h_gbps = gbps;
    }
    return rc;
}

int asc_graph_load(asc_ctx* ctx, const char* path, int format, asc_graph** out) {
    if (!ctx || !path || !out) return ASC_ERR_INVALID;
    asc_graph* g = new asc_graph();
    int r = read_graph_file(path, format, g->h);
    if (r != ASC_OK) {
        delete g;
        return r;
    }
    g->off_bytes = g->h.offset_bytes;
    return finish_graph(ctx, g, out);
}

int asc_graph_from_csr(asc_ctx* ctx, const asc_csr_view* view, asc_graph** out) {
    if (!ctx || !view || !out) return ASC_ERR_INVALID;
    asc_graph* g = new asc_graph();
    int r = csr_from_view(view, g->h);
    if (r != ASC_OK) {
        delete g;
        return r;
    }
    g->off_bytes = g->h.offset_bytes;
    return finish_graph(ctx, g, out);
}

int asc_graph_borrow_csr(asc_ctx* ctx, const asc_csr_view* view, asc_graph** out) {
    if (!ctx || !view || !out) return ASC_ERR_INVALID;
    asc_graph* g = new asc_graph();
    int r = csr_from_view(view, g->h, true);
    if (r != ASC_OK) {
        delete g;
        return r;
    }
    g->off_bytes = g->h.offset_bytes;
    return finish_graph(ctx, g, out);
}

int asc_graph_adopt(asc_ctx* ctx, asc_host_csr* h, int offset_bytes, asc_graph** out) {
    if (!ctx || !h || !out || !h->off || (h->num_edges && !h->edges)) return ASC_ERR_INVALID;
    if (offset_bytes != 4 && offset_bytes != 8) return ASC_ERR_INVALID;
    asc_graph* g = new asc_graph();
    g->h.V = h->num_vertices;
    g->h.E = h->num_edges;
    g->h.kind = h->kind;
    g->h.offset_bytes = offset_bytes;
    g->h.off.p = h->off;
    g->h.off.n = h->num_vertices + 1;
    g->h.edges.p = h->edges;
    g->h.edges.n = h->num_edges * (h->kind == ASC_KIND_WCSR ? 2 : 1);
    g->h.outdeg.p = h->out_degree;
    g->h.outdeg.n = h->out_degree ? h->num_vertices : 0;
    std::memset(h, 0, sizeof(*h));
    int r = validate_csr(g->h);
    if (r != ASC_OK) {
        delete g;
        return r;
    }
    if (g->h.kind == ASC_KIND_CSC && !g->h.outdeg.p) {
        delete g;
        return ASC_ERR_INVALID;
    }
    g->off_bytes = offset_bytes;
    return finish_graph(ctx, g, out);
}

int asc_graph_destroy(asc_graph* g) {
    if (!g) return ASC_ERR_INVALID;
    if (g->servers.load() > 0) {
        set_last_error("a server still runs over this graph; destroy it first");
        return ASC_ERR_STATE;
    }
    (void)hipSetDevice(g->ctx->device);
    graph_free_device(g);
    delete g;
    return ASC_OK;
}

int asc_graph_get_info(const asc_graph* g, asc_graph_info* info) {
    if (!g || !info) return ASC_ERR_INVALID;
    std::memset(info, 0, sizeof(*info));
    info->num_vertices = g->h.V;
    info->num_edges = g->h.E;
    info->kind = g->h.kind;
    info->offset_bytes = g->off_bytes;
    info->static_edges = g->static_edges;
    info->static_vertices = g->static_vertices;
    info->ondemand_edges = g->h.E - g->static_edges;
    info->part_edges = g->part_edges;
    info->device_bytes = g->device_bytes;
    info->max_degree = g->max_degree;
    info->max_degree_vertex = g->max_degree_vertex;
    info->symmetric = g->symmetric;
    info->tail_pinned = g->static_edges < g->h.E && g->tail_pinned() ? 1 : 0;
    info->host_bytes = (g->h.off.borrowed ? 0 : 8 * (g->h.V + 1)) +
                       (g->h.edges.p && !g->h.edges.borrowed ? (uint64_t)g->h.E * g->h.edge_bytes() : 0) +
                       (g->h.outdeg.borrowed ? 0 : 4ull * g->h.outdeg.n) + 4ull * g->ptail.n + g->ptail_w.n +
                       4ull * g->zc_tail.n;
    return ASC_OK;
}

int asc_graph_replan(asc_graph* g, uint64_t edge_budget_bytes, double advise_rate) {
    if (!g) return ASC_ERR_INVALID;
    if (g->host_edges_released) {
        set_last_error("the graph's host edge copy was released (asc_config.release_host_edges); it cannot be "
                       "re-planned");
        return ASC_ERR_UNSUPPORTED;
    }
    if (g->servers.load() > 0) {
        set_last_error("a server still runs over this graph; destroy it before re-planning");
        return ASC_ERR_STATE;
    }
    const uint64_t old_budget = g->edge_budget_bytes;
    const double old_rate = g->advise_rate;
    g->edge_budget_bytes = edge_budget_bytes;
    g->advise_rate = advise_rate;
    const int r = graph_upload(g);
    if (r != ASC_OK && g->usable) {   // refused before anything changed: keep the old plan's inputs
        g->edge_budget_bytes = old_budget;
        g->advise_rate = old_rate;
    }
    return r;
}

int asc_bfs(asc_graph* g, uint32_t source, uint32_t* level_out, asc_stats* st) {
    if (!g) return ASC_ERR_INVALID;
    ASC_TRY(graph_ready(g));
    if (source >= g->h.V) {
        set_last_error("source %u >= V %llu", source, (unsigned long long)g->h.V);
        return ASC_ERR_INVALID;
    }
    if (g->h.kind == ASC_KIND_CSC) return ASC_ERR_UNSUPPORTED;
    return traverse(g, ALG_BFS, source, level_out, st);
}

int asc_sssp(asc_graph* g, uint32_t source, uint32_t* dist_out, asc_stats* st) {
    if (!g) return ASC_ERR_INVALID;
    ASC_TRY(graph_ready(g));
    if (source >= g->h.V) {
        set_last_error("source %u >= V %llu", source, (unsigned long long)g->h.V);
        return ASC_ERR_INVALID;
    }
    if (g->h.kind != ASC_KIND_WCSR) {
        set_last_error("SSSP needs a weighted graph");
        return ASC_ERR_UNSUPPORTED;
    }
    return traverse(g, ALG_SSSP, source, dist_out, st);
}

int asc_cc(asc_graph* g, uint32_t* label_out, asc_stats* st) {
    if (!g) return ASC_ERR_INVALID;
    ASC_TRY(graph_ready(g));
    if (g->h.kind == ASC_KIND_CSC) return ASC_ERR_UNSUPPORTED;
    return traverse(g, ALG_CC, 0, label_out, st);
}

int asc_pr(asc_graph* g, const asc_pr_opts* opts, float* rank_f32, double* rank_f64, double* residual_out,
           asc_stats* st) {
    if (!g) return ASC_ERR_INVALID;
    ASC_TRY(graph_ready(g));
    asc_pr_opts o;
    if (opts) o = *opts; else asc_pr_opts_init(&o);
    if (o.mode != ASC_PR_NORMALISED && o.mode != ASC_PR_REFERENCE) return ASC_ERR_INVALID;
    if (!(o.damping > 0 && o.damping < 1)) return ASC_ERR_INVALID;
    return pagerank(g, &o, rank_f32, rank_f64, residual_out, st);
}

}  // extern "C"
