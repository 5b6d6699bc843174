// PageRank over the in-CSR (BCSC input, GraphMeta.cuh:159-162).
//
// Normalised mode: r0 = 1/V; r' = (1-d)/V + d*(sum_{u->v} r[u]/outdeg[u] +
// dangling/V) for a fixed number of iterations.  Ranks and the gathered
// contributions c[u] = r[u]/outdeg[u] are FP32 (4 bytes per random gather);
// each vertex's in-edge sum, the residual sum |r'-r| and the dangling mass
// are accumulated in FP64.  The pull is coalesced: G lanes of a wave share a
// vertex and read its in-edge ids contiguously; vertices with in-degree >=
// the heavy threshold are summed by a whole workgroup instead.  The residual
// and dangling sums are produced inside the same kernel, so iterations run
// back-to-back without a host synchronisation.
//
// Reference mode (gpu_kernels.cuh:533-594, CalculateOpt.cu:547-721): FP64,
// 0.15 + 0.85*sum over active vertices, threshold 0.001, loop while any
// vertex is active.  Each active vertex's sum runs sequentially in CSR order
// and no multiply-add is fused, so results equal a sequential FP64
// evaluation of the same formula.
#include <algorithm>

#include "device_common.hpp"
#include "internal.hpp"

namespace asc {

struct PrState {
    uint64_t V = 0;
    float* rank[2] = {nullptr, nullptr};
    float* contrib[2] = {nullptr, nullptr};
    double* acc = nullptr;        // [0, MAXIT): residual; [MAXIT, 2*MAXIT+1): dangling
    uint32_t* heavy = nullptr;    // vertices with in-degree >= heavy threshold
    uint32_t n_heavy = 0;
    int heavy_thr = 0;
    int acc_iters = 0;
    // reference mode
    double* val64 = nullptr;
    double* sum64 = nullptr;
    double* contrib64 = nullptr;
    uint8_t* active = nullptr;
    uint32_t* count = nullptr;
    HostBuf<uint32_t> h_count;
    // on-demand in-edges: vertex-range chunks [v0, v1) whose in-edges
    // [off[v0], off[v1]) fit one part buffer, copied from the pinned tail
    struct Chunk {
        uint64_t v0, v1, e0, e1;
    };
    std::vector<Chunk> chunks;
    uint32_t* part[2] = {nullptr, nullptr};
    uint64_t part_edges = 0;
    hipEvent_t ev_copy[2] = {nullptr, nullptr}, ev_kern[2] = {nullptr, nullptr};
    uint64_t streamed_edges = 0;  // per pull sweep
};

void pr_free(PrState* p) {
    if (!p) return;
    auto f = [](void* x) {
        if (x) (void)hipFree(x);
    };
    f(p->rank[0]); f(p->rank[1]); f(p->contrib[0]); f(p->contrib[1]);
    f(p->acc); f(p->heavy); f(p->val64); f(p->sum64); f(p->contrib64); f(p->active); f(p->count);
    f(p->part[0]); f(p->part[1]);
    for (int i = 0; i < 2; i++) {
        if (p->ev_copy[i]) (void)hipEventDestroy(p->ev_copy[i]);
        if (p->ev_kern[i]) (void)hipEventDestroy(p->ev_kern[i]);
    }
    delete p;
}

// Plans the on-demand chunks of a partially resident PageRank graph (once).
static int pr_plan_chunks(asc_graph* g, PrState* p) {
    const uint64_t V = g->h.V, vs = g->static_vertices;
    if (vs >= V || !p->chunks.empty()) return ASC_OK;
    const uint64_t* off = g->h.off.p;
    uint64_t maxdeg = 0;
    for (uint64_t v = vs; v < V; v++) maxdeg = std::max<uint64_t>(maxdeg, off[v + 1] - off[v]);
    // a part holds at least the largest on-demand in-list
    p->part_edges = std::max<uint64_t>(g->part_edges, maxdeg);
    uint64_t v = vs;
    while (v < V) {
        uint64_t v1 = v;
        while (v1 < V && off[v1 + 1] - off[v] <= p->part_edges) v1++;
        if (v1 == v) v1 = v + 1;
        p->chunks.push_back({v, v1, off[v], off[v1]});
        v = v1;
    }
    p->streamed_edges = off[V] - off[vs];
    for (int i = 0; i < 2; i++) {
        ASC_HIP(hipMalloc(&p->part[i], 4 * std::max<uint64_t>(p->part_edges, 1)));
        ASC_HIP(hipEventCreateWithFlags(&p->ev_copy[i], hipEventDisableTiming));
        ASC_HIP(hipEventCreateWithFlags(&p->ev_kern[i], hipEventDisableTiming));
        // recorded once so that every copy can wait for the buffer's last reader
        ASC_HIP(hipEventRecord(p->ev_kern[i], g->ctx->s_main));
    }
    return ASC_OK;
}

// Enqueues, for every on-demand chunk, its copy from the pinned tail on the
// copy stream (double-buffered) and `pull(src, e_base, v0, v1)` on the compute
// stream.  No host synchronisation: every copy waits for the previous kernel
// that read its buffer, also across iterations.
template <class F>
static int pr_stream_chunks(asc_graph* g, PrState* p, F&& pull) {
    asc_ctx* ctx = g->ctx;
    const uint64_t S = g->static_edges;
    for (size_t c = 0; c < p->chunks.size(); c++) {
        const auto& ch = p->chunks[c];
        const int b = (int)(c & 1);
        ASC_HIP(hipStreamWaitEvent(ctx->s_copy, p->ev_kern[b], 0));
        ASC_HIP(hipMemcpyAsync(p->part[b], g->zc_tail.p + (ch.e0 - S), 4 * (ch.e1 - ch.e0), hipMemcpyHostToDevice,
                               ctx->s_copy));
        ASC_HIP(hipEventRecord(p->ev_copy[b], ctx->s_copy));
        ASC_HIP(hipStreamWaitEvent(ctx->s_main, p->ev_copy[b], 0));
        ASC_TRY(pull(p->part[b], ch.e0, ch.v0, ch.v1));
        ASC_HIP(hipEventRecord(p->ev_kern[b], ctx->s_main));
    }
    return ASC_OK;
}

// ---- normalised mode --------------------------------------------------------

__global__ void k_pr_init(float* rank, float* contrib, const uint32_t* outdeg, uint64_t V, double* dang0) {
    __shared__ double s_red[kWavesPerBlock];
    const float r0 = (float)(1.0 / (double)V);
    double dang = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t d = outdeg[v];
        rank[v] = r0;
        contrib[v] = d ? r0 / (float)d : 0.0f;
        if (!d) dang += (double)r0;
    }
    dang = wave_sum(dang);
    if (lane_id() == 0) s_red[wave_id()] = dang;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int i = 0; i < kWavesPerBlock; i++) t += s_red[i];
        if (t != 0) atomicAdd(dang0, t);
    }
}

struct PrArgs {
    const void* off;              // OffT[V+1]
    const uint32_t* src;          // in-edges: edge k at src[k - e_base]
    unsigned long long e_base;
    unsigned long long v_begin, v_end;  // vertices pulled by this launch
    int skip_heavy;               // leave in-degree >= heavy_thr to k_pr_pull_heavy
    const uint32_t* outdeg;
    const float* rank_in;
    const float* contrib_in;
    float* rank_out;
    float* contrib_out;
    const double* dang_in;        // dangling mass of rank_in
    double* dang_out;             // dangling mass of rank_out
    double* residual;
    uint64_t V;
    float damping;
    double damping64;
    uint32_t heavy_thr;
};

__device__ __forceinline__ void pr_finish_vertex(const PrArgs& a, uint64_t v, double s, double base, double& res,
                                                 double& dang) {
    const float nr = (float)(base + a.damping64 * s);
    const uint32_t d = a.outdeg[v];
    res += fabs((double)nr - (double)a.rank_in[v]);
    a.rank_out[v] = nr;
    a.contrib_out[v] = d ? nr / (float)d : 0.0f;
    if (!d) dang += (double)nr;
}

__device__ __forceinline__ void pr_block_reduce2(double res, double dang, double* out_res, double* out_dang) {
    __shared__ double s_r[kWavesPerBlock], s_d[kWavesPerBlock];
    res = wave_sum(res);
    dang = wave_sum(dang);
    if (lane_id() == 0) {
        s_r[wave_id()] = res;
        s_d[wave_id()] = dang;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0, d = 0;
        for (int i = 0; i < kWavesPerBlock; i++) {
            r += s_r[i];
            d += s_d[i];
        }
        if (r != 0) atomicAdd(out_res, r);
        if (d != 0) atomicAdd(out_dang, d);
    }
}

// G lanes per vertex; vertices with in-degree >= heavy_thr are skipped here.
template <int G, class OffT>
__global__ __launch_bounds__(kBlock) void k_pr_pull(PrArgs a) {
    const OffT* off = (const OffT*)a.off;
    const double base = (1.0 - a.damping64) / (double)a.V + a.damping64 * (*a.dang_in) / (double)a.V;
    const uint32_t sub = threadIdx.x % G;
    const uint64_t groups_per_grid = (uint64_t)gridDim.x * (blockDim.x / G);
    const uint32_t* src = a.src - a.e_base;
    double res = 0, dang = 0;
    for (uint64_t v = a.v_begin + blockIdx.x * (uint64_t)(blockDim.x / G) + threadIdx.x / G; v < a.v_end;
         v += groups_per_grid) {
        const OffT b = off[v], e = off[v + 1];
        const bool heavy = a.skip_heavy && (uint64_t)(e - b) >= a.heavy_thr;
        double s = 0;
        if (!heavy)
            for (OffT k = b + sub; k < e; k += G) s += (double)a.contrib_in[src[k]];
#pragma unroll
        for (int d = G / 2; d > 0; d >>= 1) s += __shfl_xor(s, d, G);
        if (sub == 0 && !heavy) pr_finish_vertex(a, v, s, base, res, dang);
    }
    pr_block_reduce2(res, dang, a.residual, a.dang_out);
}

template <class OffT>
__global__ __launch_bounds__(kBlock) void k_pr_pull_heavy(PrArgs a, const uint32_t* heavy, uint32_t n_heavy) {
    __shared__ double s_sum[kWavesPerBlock];
    const OffT* off = (const OffT*)a.off;
    const double base = (1.0 - a.damping64) / (double)a.V + a.damping64 * (*a.dang_in) / (double)a.V;
    double res = 0, dang = 0;
    for (uint32_t i = blockIdx.x; i < n_heavy; i += gridDim.x) {
        const uint32_t v = heavy[i];
        const OffT b = off[v], e = off[v + 1];
        double s = 0;
        for (OffT k = b + threadIdx.x; k < e; k += blockDim.x) s += (double)a.contrib_in[a.src[k]];
        s = wave_sum(s);
        if (lane_id() == 0) s_sum[wave_id()] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0;
            for (int j = 0; j < kWavesPerBlock; j++) t += s_sum[j];
            pr_finish_vertex(a, v, t, base, res, dang);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (res != 0) atomicAdd(a.residual, res);
        if (dang != 0) atomicAdd(a.dang_out, dang);
    }
}

// ---- reference mode --------------------------------------------------------

__global__ void k_prref_init(double* val, uint8_t* active, uint64_t V) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        val[v] = 1.0 / (double)V;  // GraphMeta.cuh:418
        active[v] = 1;
    }
}

__global__ void k_prref_contrib(const double* val, const uint32_t* outdeg, double* contrib, uint64_t V) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
        contrib[v] = outdeg[v] ? val[v] / (double)outdeg[v] : 0.0;
}

// prSumKernel_static / _dynamic (gpu_kernels.cuh:554-594): sequential FP64
// sum in CSR order for each active vertex.
template <class OffT>
__global__ void k_prref_sum(const OffT* off, const uint32_t* src, unsigned long long e_base, const double* contrib,
                            const uint8_t* active, double* sum, uint64_t v_begin, uint64_t v_end) {
    src -= e_base;
    for (uint64_t v = v_begin + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < v_end;
         v += (uint64_t)gridDim.x * blockDim.x) {
        if (!active[v]) continue;
        double s = 0;
        for (OffT k = off[v]; k < off[v + 1]; k++) s += contrib[src[k]];
        sum[v] = s;
    }
}

// prKernel_Opt (gpu_kernels.cuh:533-552).
__global__ void k_prref_update(double* val, double* sum, uint8_t* active, uint64_t V, double base, double damping,
                               double threshold, uint32_t* count) {
#pragma clang fp contract(off)
    uint32_t c = 0;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        if (!active[v]) continue;
        const double t = base + damping * sum[v];
        const double diff = t > val[v] ? (t - val[v]) : (val[v] - t);
        if (diff > threshold) {
            val[v] = t;
            c++;
        } else {
            active[v] = 0;
        }
        sum[v] = 0;
    }
    c = wave_sum(c);
    if (lane_id() == 0 && c) atomicAdd(count, c);
}

// ---- host -------------------------------------------------------------------

static int pr_ensure(asc_graph* g, int iters, bool ref) {
    PrState* p = g->pr;
    const uint64_t V = g->h.V;
    if (!p || p->V != V) {
        if (p) pr_free(p);
        p = new PrState();
        g->pr = p;
        p->V = V;
    }
    ASC_HIP(hipSetDevice(g->ctx->device));
    if (!ref && !p->rank[0]) {
        for (int i = 0; i < 2; i++) {
            ASC_HIP(hipMalloc(&p->rank[i], 4 * V));
            ASC_HIP(hipMalloc(&p->contrib[i], 4 * V));
        }
    }
    if (!ref && p->acc_iters < iters) {
        if (p->acc) (void)hipFree(p->acc);
        ASC_HIP(hipMalloc(&p->acc, sizeof(double) * (2 * (size_t)iters + 1)));
        p->acc_iters = iters;
    }
    const int thr = 4096;  // in-degree from which a workgroup sums one vertex
    if (!ref && p->heavy_thr != thr) {
        // resident vertices only: their in-edges are on the device
        std::vector<uint32_t> hv;
        for (uint64_t v = 0; v < g->static_vertices; v++)
            if (g->h.off[v + 1] - g->h.off[v] >= (uint64_t)thr) hv.push_back((uint32_t)v);
        p->n_heavy = (uint32_t)hv.size();
        p->heavy_thr = thr;
        if (!hv.empty()) {
            ASC_HIP(hipMalloc(&p->heavy, 4 * hv.size()));
            ASC_HIP(hipMemcpy(p->heavy, hv.data(), 4 * hv.size(), hipMemcpyHostToDevice));
        }
    }
    if (ref && !p->val64) {
        ASC_HIP(hipMalloc(&p->val64, 8 * V));
        ASC_HIP(hipMalloc(&p->sum64, 8 * V));
        ASC_HIP(hipMalloc(&p->contrib64, 8 * V));
        ASC_HIP(hipMalloc(&p->active, V));
        ASC_HIP(hipMalloc(&p->count, 4));
        ASC_TRY(p->h_count.alloc(1, true));
    }
    return pr_plan_chunks(g, p);
}

static int pr_group_for(const asc_graph* g) {
    if (g->pr_group > 0) return g->pr_group;
    const double avg = g->h.V ? (double)g->h.E / (double)g->h.V : 0;
    if (avg >= 48) return 32;
    if (avg >= 12) return 16;
    if (avg >= 6) return 8;
    return 4;
}

template <class OffT>
static int launch_pull(int G, int grid, hipStream_t s, const PrArgs& a) {
    switch (G) {
        case 4: hipLaunchKernelGGL((k_pr_pull<4, OffT>), dim3(grid), dim3(kBlock), 0, s, a); break;
        case 8: hipLaunchKernelGGL((k_pr_pull<8, OffT>), dim3(grid), dim3(kBlock), 0, s, a); break;
        case 16: hipLaunchKernelGGL((k_pr_pull<16, OffT>), dim3(grid), dim3(kBlock), 0, s, a); break;
        case 32: hipLaunchKernelGGL((k_pr_pull<32, OffT>), dim3(grid), dim3(kBlock), 0, s, a); break;
        default: hipLaunchKernelGGL((k_pr_pull<64, OffT>), dim3(grid), dim3(kBlock), 0, s, a); break;
    }
    ASC_HIP(hipGetLastError());
    return ASC_OK;
}

template <class OffT>
static int pr_normalised(asc_graph* g, const asc_pr_opts* o, float* rank_out, double* residual_out, asc_stats* st) {
    asc_ctx* ctx = g->ctx;
    const int iters = o->iterations > 0 ? o->iterations : 20;
    ASC_TRY(pr_ensure(g, iters, false));
    PrState* p = g->pr;
    const uint64_t V = g->h.V;
    hipStream_t s = ctx->s_main;
    const int grid = ctx->num_cus * 8;
    ASC_HIP(hipMemsetAsync(p->acc, 0, sizeof(double) * (2 * (size_t)iters + 1), s));
    double* resid = p->acc;
    double* dang = p->acc + iters;
    hipLaunchKernelGGL(k_pr_init, dim3(grid), dim3(kBlock), 0, s, p->rank[0], p->contrib[0], g->d_outdeg, V, dang);
    ASC_HIP(hipGetLastError());
    ASC_HIP(hipStreamSynchronize(s));
    const int G = pr_group_for(g);
    hipEvent_t e0, e1;
    ASC_HIP(hipEventCreate(&e0));
    ASC_HIP(hipEventCreate(&e1));
    const double t0 = now_ms();
    ASC_HIP(hipEventRecord(e0, s));
    for (int it = 0; it < iters; it++) {
        PrArgs a{};
        a.off = g->d_off;
        a.src = (const uint32_t*)g->d_edges;
        a.outdeg = g->d_outdeg;
        a.rank_in = p->rank[it & 1];
        a.contrib_in = p->contrib[it & 1];
        a.rank_out = p->rank[(it + 1) & 1];
        a.contrib_out = p->contrib[(it + 1) & 1];
        a.dang_in = dang + it;
        a.dang_out = dang + it + 1;
        a.residual = resid + it;
        a.V = V;
        a.damping = (float)o->damping;
        a.damping64 = o->damping;
        a.heavy_thr = (uint32_t)p->heavy_thr;
        // resident vertices from the device edges, heavy in-lists by workgroups
        a.e_base = 0;
        a.v_begin = 0;
        a.v_end = g->static_vertices;
        a.skip_heavy = 1;
        if (a.v_end > 0) ASC_TRY(launch_pull<OffT>(G, grid, s, a));
        if (p->n_heavy) {
            hipLaunchKernelGGL((k_pr_pull_heavy<OffT>), dim3(std::min<uint32_t>(p->n_heavy, grid)), dim3(kBlock), 0,
                               s, a, p->heavy, p->n_heavy);
            ASC_HIP(hipGetLastError());
        }
        // on-demand vertices from streamed chunks
        ASC_TRY(pr_stream_chunks(g, p, [&](const uint32_t* src, uint64_t e0, uint64_t v0, uint64_t v1) {
            PrArgs c = a;
            c.src = src;
            c.e_base = e0;
            c.v_begin = v0;
            c.v_end = v1;
            c.skip_heavy = 0;
            const int cgrid = (int)std::min<uint64_t>(grid, (v1 - v0 + 63) / 64 + 1);
            return launch_pull<OffT>(G, cgrid, s, c);
        }));
    }
    ASC_HIP(hipEventRecord(e1, s));
    ASC_HIP(hipStreamSynchronize(s));
    const double t1 = now_ms();
    float kms = 0;
    (void)hipEventElapsedTime(&kms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::vector<double> hres(iters);
    ASC_HIP(hipMemcpy(hres.data(), resid, sizeof(double) * iters, hipMemcpyDeviceToHost));
    if (residual_out) std::memcpy(residual_out, hres.data(), sizeof(double) * iters);
    if (rank_out) ASC_HIP(hipMemcpy(rank_out, p->rank[iters & 1], 4 * V, hipMemcpyDeviceToHost));
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->total_ms = t1 - t0;
        st->static_ms = kms;
        st->kernel_ms = kms;
        st->iterations = (uint32_t)iters;
        st->node_sum = V * (uint64_t)iters;
        st->edges_traversed = g->h.E * (uint64_t)iters;
        st->static_edges_traversed = st->edges_traversed;
        st->vertices_processed = V * (uint64_t)iters;
        st->reached = V;
        st->reached_degree_sum = g->h.E;
        st->pr_residual = hres[iters - 1];
        st->host_cores = hardware_threads();
        st->streamed_edges = p->streamed_edges * (uint64_t)iters;
        st->streamed_bytes = 4 * st->streamed_edges;
        st->parts = (uint32_t)(p->chunks.size() * (size_t)iters);
        st->kern[3].launches = (uint32_t)iters;
        st->kern[3].ms = kms;
        st->kern[3].edges = g->h.E * (uint64_t)iters;
        st->kern[3].vertices = V * (uint64_t)iters;
    }
    return ASC_OK;
}

template <class OffT>
static int pr_reference(asc_graph* g, const asc_pr_opts* o, double* rank_out, asc_stats* st) {
    asc_ctx* ctx = g->ctx;
    ASC_TRY(pr_ensure(g, 0, true));
    PrState* p = g->pr;
    const uint64_t V = g->h.V;
    hipStream_t s = ctx->s_main;
    const int grid = ctx->num_cus * 8;
    // 0.15 and 0.85 are the reference's literals (gpu_kernels.cuh:538)
    const double damping = o->damping;
    const double base = damping == 0.85 ? 0.15 : 1.0 - damping;
    const int maxit = o->max_iterations > 0 ? o->max_iterations : 10000;
    hipLaunchKernelGGL(k_prref_init, dim3(grid), dim3(256), 0, s, p->val64, p->active, V);
    ASC_HIP(hipGetLastError());
    ASC_HIP(hipStreamSynchronize(s));
    const double t0 = now_ms();
    uint64_t node_sum = V;
    uint32_t active = (uint32_t)V;
    int it = 0;
    while (active && it < maxit) {
        it++;
        hipLaunchKernelGGL(k_prref_contrib, dim3(grid), dim3(256), 0, s, p->val64, g->d_outdeg, p->contrib64, V);
        if (g->static_vertices)
            hipLaunchKernelGGL((k_prref_sum<OffT>), dim3(grid), dim3(256), 0, s, (const OffT*)g->d_off,
                               (const uint32_t*)g->d_edges, 0ull, p->contrib64, p->active, p->sum64, 0ull,
                               (unsigned long long)g->static_vertices);
        ASC_TRY(pr_stream_chunks(g, p, [&](const uint32_t* src, uint64_t e0, uint64_t v0, uint64_t v1) {
            const int cgrid = (int)std::min<uint64_t>(grid, (v1 - v0 + 255) / 256);
            hipLaunchKernelGGL((k_prref_sum<OffT>), dim3(cgrid), dim3(256), 0, s, (const OffT*)g->d_off, src,
                               (unsigned long long)e0, p->contrib64, p->active, p->sum64, (unsigned long long)v0,
                               (unsigned long long)v1);
            ASC_HIP(hipGetLastError());
            return ASC_OK;
        }));
        ASC_HIP(hipMemsetAsync(p->count, 0, 4, s));
        hipLaunchKernelGGL(k_prref_update, dim3(grid), dim3(256), 0, s, p->val64, p->sum64, p->active, V, base,
                           damping, o->threshold, p->count);
        ASC_HIP(hipGetLastError());
        ASC_HIP(hipMemcpyAsync(p->h_count.p, p->count, 4, hipMemcpyDeviceToHost, s));
        ASC_HIP(hipStreamSynchronize(s));
        active = p->h_count[0];
        node_sum += active;
    }
    const double t1 = now_ms();
    if (rank_out) ASC_HIP(hipMemcpy(rank_out, p->val64, 8 * V, hipMemcpyDeviceToHost));
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->total_ms = t1 - t0;
        st->iterations = (uint32_t)it;
        st->node_sum = node_sum;
        st->reached = V;
        st->reached_degree_sum = g->h.E;
        st->host_cores = hardware_threads();
    }
    return ASC_OK;
}

int pagerank(asc_graph* g, const asc_pr_opts* o, float* rank_f32, double* rank_f64, double* residual_out,
             asc_stats* st) {
    if (g->h.kind != ASC_KIND_CSC) {
        set_last_error("PageRank needs a BCSC graph (in-CSR + out-degree)");
        return ASC_ERR_UNSUPPORTED;
    }
    if (o->mode == ASC_PR_REFERENCE)
        return g->off_bytes == 4 ? pr_reference<uint32_t>(g, o, rank_f64, st) : pr_reference<uint64_t>(g, o, rank_f64, st);
    return g->off_bytes == 4 ? pr_normalised<uint32_t>(g, o, rank_f32, residual_out, st)
                             : pr_normalised<uint64_t>(g, o, rank_f32, residual_out, st);
}

}  // namespace asc
