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
    // propagation-blocked normalised mode (every edge resident): out-edge e
    // of source src_of[e] writes its contribution to slot pos[e] (or none,
    // kPbSkip, when its destination is heavy); the slots of destination bin
    // b and source range r form one region, slot k feeding bin-local vertex
    // dl[k]
    int pb_state = 0;             // 0 not built, 1 built, -1 not possible
    uint32_t pb_nbins = 0;
    uint64_t pb_edges = 0;        // out-edges (= E)
    uint32_t* pb_src = nullptr;
    uint32_t* pb_pos = nullptr;
    uint16_t* pb_dl = nullptr;
    float* pb_vals = nullptr;
    unsigned long long* pb_bin_start = nullptr;   // region bases [b * R + r]
    uint32_t* pb_cnt = nullptr;                   // region sizes [b * R + r]
    // source-segmented normalised mode (every edge resident): the light
    // in-edges regrouped into tiles (destination block d, source segment s),
    // tile (d, s) at tl_edges[tl_off[d * tl_ns + s] ...], one u32 word per
    // edge (block-local destination << kTlSrcBits | segment-local source)
    int tl_state = 0;             // 0 not built, 1 built, -1 not possible
    uint32_t tl_sb = 0, tl_db = 0, tl_nd = 0, tl_ns = 0;
    uint64_t tl_edges_n = 0;
    uint32_t* tl_edges = nullptr;
    unsigned long long* tl_off = nullptr;
    uint32_t* tl_heavy = nullptr; // bitmap of destinations left to k_pr_pull_heavy
};

void pr_free(PrState* p) {
    if (!p) return;
    auto f = [](void* x) {
        if (x) (void)hipFree(x);
    };
    f(p->rank[0]); f(p->rank[1]); f(p->contrib[0]); f(p->contrib[1]);
    f(p->acc); f(p->heavy); f(p->val64); f(p->sum64); f(p->contrib64); f(p->active); f(p->count);
    f(p->part[0]); f(p->part[1]);
    f(p->pb_src); f(p->pb_pos); f(p->pb_dl); f(p->pb_vals); f(p->pb_bin_start); f(p->pb_cnt);
    f(p->tl_edges); f(p->tl_off); f(p->tl_heavy);
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
        if (!heavy) {
            // four edges per lane per pass: their source ids, then their
            // contributions, are loaded back to back (a lane past the end
            // re-reads the list's first edge and drops it)
            for (OffT k = b + sub; k < e; k += 4 * G) {
                uint32_t u[4];
#pragma unroll
                for (int j = 0; j < 4; j++) u[j] = src[k + (OffT)(j * G) < e ? k + (OffT)(j * G) : b];
                float c[4];
#pragma unroll
                for (int j = 0; j < 4; j++) c[j] = a.contrib_in[u[j]];
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (k + (OffT)(j * G) < e) s += (double)c[j];
            }
        }
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

// ---- propagation-blocked normalised mode -----------------------------------
// The pull gathers one random 4-byte contribution per in-edge and is bound by
// the lines those gathers move.  This schedule streams instead (selected by
// ASC_PR_BLOCKED; on Erdos-Renyi 2^26 / 2^31 edges it measured 46 ms scatter +
// 11.5 ms gather per iteration against 42 ms for the pull, the scattered
// 4-byte slot writes evicting partly filled lines):
//   scatter  out-edge e of source u writes contrib[u] to its slot pos[e]
//            (sources in order, so contrib is read sequentially; the slots of
//            one source range and destination bin are contiguous);
//   gather   one workgroup per bin of 32768 destinations adds its slots into
//            LDS (FP32) and finishes those vertices.
// The slot layout is fixed by the graph and built once.  Destinations with
// in-degree >= the heavy threshold keep the FP64 pull (k_pr_pull_heavy).
constexpr int kPbBinBits = 15;
constexpr uint32_t kPbBin = 1u << kPbBinBits;   // destinations per bin (128 KB of sums)
constexpr uint32_t kPbSkip = 0xFFFFFFFFu;       // out-edge into a heavy destination
constexpr uint32_t kPbRanges = 2048;            // source ranges of the slot layout
constexpr int kPbBlock = 1024;                  // gather workgroup
constexpr uint32_t kPbMaxBins = 8192;           // build histograms live in LDS

// out_dst[cursor[u]++] = v for every in-edge (u -> v): the out-CSR of a
// graph given as in-CSR (order within an out-list is arbitrary).
template <class OffT>
__global__ __launch_bounds__(kBlock) void k_pb_transpose(const OffT* in_off, const uint32_t* in_src, uint64_t V,
                                                         unsigned long long* cursor, uint32_t* out_dst) {
    for (uint64_t v = blockIdx.x * (uint64_t)kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock)
        for (OffT k = in_off[v]; k < in_off[v + 1]; k++)
            out_dst[atomicAdd(cursor + in_src[k], 1ull)] = (uint32_t)v;
}

// src[e] = u for the out-edges e of u.
template <class OffT>
__global__ __launch_bounds__(kBlock) void k_pb_src(const OffT* out_off, uint64_t V, uint32_t* src) {
    for (uint64_t u = blockIdx.x * (uint64_t)kBlock + threadIdx.x; u < V; u += (uint64_t)gridDim.x * kBlock)
        for (OffT e = out_off[u]; e < out_off[u + 1]; e++) src[e] = (uint32_t)u;
}

__device__ __forceinline__ uint64_t pb_range_begin(uint64_t E, uint32_t r) { return E * r / kPbRanges; }

// cnt[b * kPbRanges + r]: out-edges of range r (edges [E r / R, E (r+1) / R))
// into light destinations of bin b.
template <class OffT>
__global__ __launch_bounds__(kBlock) void k_pb_count(const uint32_t* out_dst, uint64_t E, const OffT* in_off,
                                                     uint32_t heavy_thr, uint32_t nb, uint32_t* cnt) {
    __shared__ uint32_t h[kPbMaxBins];
    for (uint32_t r = blockIdx.x; r < kPbRanges; r += gridDim.x) {
        for (uint32_t b = threadIdx.x; b < nb; b += kBlock) h[b] = 0;
        __syncthreads();
        for (uint64_t e = pb_range_begin(E, r) + threadIdx.x; e < pb_range_begin(E, r + 1); e += kBlock) {
            const uint32_t v = out_dst[e];
            if ((uint64_t)(in_off[v + 1] - in_off[v]) < heavy_thr) atomicAdd(&h[v >> kPbBinBits], 1u);
        }
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < nb; b += kBlock) cnt[(size_t)b * kPbRanges + r] = h[b];
        __syncthreads();
    }
}

// pos[e] for every out-edge (slots of (bin, range) handed out from
// base[b * kPbRanges + r]) and dl[slot] = bin-local destination.
template <class OffT>
__global__ __launch_bounds__(kBlock) void k_pb_assign(const uint32_t* out_dst, uint64_t E, const OffT* in_off,
                                                      uint32_t heavy_thr, uint32_t nb,
                                                      const unsigned long long* base, uint32_t* pos, uint16_t* dl) {
    __shared__ unsigned long long cur[kPbMaxBins];
    for (uint32_t r = blockIdx.x; r < kPbRanges; r += gridDim.x) {
        for (uint32_t b = threadIdx.x; b < nb; b += kBlock) cur[b] = base[(size_t)b * kPbRanges + r];
        __syncthreads();
        for (uint64_t e = pb_range_begin(E, r) + threadIdx.x; e < pb_range_begin(E, r + 1); e += kBlock) {
            const uint32_t v = out_dst[e];
            if ((uint64_t)(in_off[v + 1] - in_off[v]) < heavy_thr) {
                const unsigned long long k = atomicAdd(&cur[v >> kPbBinBits], 1ull);
                pos[e] = (uint32_t)k;
                dl[k] = (uint16_t)(v & (kPbBin - 1));
            } else {
                pos[e] = kPbSkip;
            }
        }
        __syncthreads();
    }
}

// Workgroups are dealt to the 8 XCDs round-robin (blockIdx % 8), and each
// XCD walks the source ranges x, x + 8, ... with all of its workgroups on one
// range at a time: a range writes 2048 slot regions (one per bin), whose
// partly filled lines then stay in that XCD's L2 until they are whole instead
// of being written back from every XCD.  The grid is a multiple of 8.
__global__ __launch_bounds__(kBlock) void k_pb_scatter(const uint32_t* src, const uint32_t* pos, uint64_t E,
                                                       const float* contrib, float* vals) {
    const uint32_t xcd = blockIdx.x & 7u, local = blockIdx.x >> 3, nloc = gridDim.x >> 3;
    for (uint32_t r = xcd; r < kPbRanges; r += 8) {
        const uint64_t e1 = pb_range_begin(E, r + 1);
        for (uint64_t e = pb_range_begin(E, r) + (uint64_t)local * kBlock + threadIdx.x; e < e1;
             e += (uint64_t)nloc * kBlock) {
            const uint32_t k = pos[e];
            if (k != kPbSkip) vals[k] = contrib[src[e]];
        }
    }
}

// The slots of bin b are the regions (b, r) of every source range r, at
// reg_base[b * R + r] with cnt[b * R + r] slots; a wave adds one region at a
// time.
template <class OffT>
__global__ __launch_bounds__(kPbBlock) void k_pb_gather(PrArgs a, const unsigned long long* reg_base,
                                                        const uint32_t* cnt, const uint16_t* dl, const float* vals) {
    __shared__ float acc[kPbBin];
    __shared__ double s_r[kPbBlock / kWave], s_d[kPbBlock / kWave];
    const OffT* off = (const OffT*)a.off;
    const uint32_t b = blockIdx.x;
    for (uint32_t i = threadIdx.x; i < kPbBin; i += kPbBlock) acc[i] = 0.0f;
    __syncthreads();
    for (uint32_t r = threadIdx.x / kWave; r < kPbRanges; r += kPbBlock / kWave) {
        const unsigned long long k0 = reg_base[(size_t)b * kPbRanges + r];
        const uint32_t n = cnt[(size_t)b * kPbRanges + r];
        for (uint32_t k = lane_id(); k < n; k += kWave) atomicAdd(&acc[dl[k0 + k]], vals[k0 + k]);
    }
    __syncthreads();
    const double base = (1.0 - a.damping64) / (double)a.V + a.damping64 * (*a.dang_in) / (double)a.V;
    double res = 0, dang = 0;
    for (uint32_t i = threadIdx.x; i < kPbBin; i += kPbBlock) {
        const uint64_t v = (uint64_t)b * kPbBin + i;
        if (v >= a.V || (uint64_t)(off[v + 1] - off[v]) >= a.heavy_thr) continue;
        pr_finish_vertex(a, v, (double)acc[i], base, res, dang);
    }
    res = wave_sum(res);
    dang = wave_sum(dang);
    if (lane_id() == 0) {
        s_r[threadIdx.x / kWave] = res;
        s_d[threadIdx.x / kWave] = dang;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0, d = 0;
        for (int i = 0; i < kPbBlock / kWave; i++) {
            r += s_r[i];
            d += s_d[i];
        }
        if (r != 0) atomicAdd(a.residual, r);
        if (d != 0) atomicAdd(a.dang_out, d);
    }
}

// ---- source-segmented normalised mode --------------------------------------
// The pull gathers one random 4-byte contribution per in-edge from an array
// of 4V bytes (256 MB at V = 2^26, beyond the 4 MB L2 of an XCD and as large
// as the MALL), so almost every gather moves a whole line from HBM.  Here the
// light in-edges are regrouped once into tiles by (destination block of 2^db
// vertices, source segment of 2^sb vertices).  One workgroup owns a
// destination block at a time: it keeps the block's sums in LDS (FP32) and
// walks the segments in order, so the gathers of a tile read one 4 * 2^sb
// byte slice of the contributions, and the workgroups resident on an XCD,
// walking the segments in the same order, share that slice in the XCD's L2.
// A tile edge is one u32 word (block-local destination << sb | segment-local
// source).  The grid is persistent: workgroup b takes blocks b, b + G, ...,
// so the resident workgroups start together and stay near each other's
// segment.  The last pass over a block finishes its vertices (rank,
// contribution, residual, dangling mass) as the pull does.  Destinations with
// in-degree >= the heavy threshold keep the FP64 pull (k_pr_pull_heavy).
constexpr int kTlMaxDstBits = 15;                          // LDS sums: up to 32768 FP32 (128 KB)
constexpr uint32_t kTlMaxSeg = 8192;                       // build histograms in LDS
constexpr int kTlBlock = 1024;                             // workgroup of the tiled pull

struct TileShape {
    uint32_t sb, db;            // source segment / destination block bits
    uint32_t ns, nd;            // segments, blocks
};

// cnt[d * ns + s] = light in-edges of destination block d from segment s.
template <class OffT>
__global__ __launch_bounds__(kBlock) void k_tl_count(const OffT* off, const uint32_t* src, uint64_t V,
                                                     uint32_t heavy_thr, TileShape ts, uint32_t* cnt) {
    __shared__ uint32_t h[kTlMaxSeg];
    const uint64_t d = blockIdx.x;
    for (uint32_t i = threadIdx.x; i < ts.ns; i += kBlock) h[i] = 0;
    __syncthreads();
    const uint64_t v0 = d << ts.db, v1 = min<uint64_t>(V, v0 + (1ull << ts.db));
    for (uint64_t v = v0 + threadIdx.x; v < v1; v += kBlock) {
        const OffT b = off[v], e = off[v + 1];
        if ((uint64_t)(e - b) >= heavy_thr) continue;
        for (OffT k = b; k < e; k++) atomicAdd(&h[src[k] >> ts.sb], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < ts.ns; i += kBlock) cnt[d * ts.ns + i] = h[i];
}

// Writes the tile words of destination block d (positions within a tile in
// arrival order) and marks its heavy destinations in `heavy_bits`.
template <class OffT>
__global__ __launch_bounds__(kBlock) void k_tl_fill(const OffT* off, const uint32_t* src, uint64_t V,
                                                    uint32_t heavy_thr, TileShape ts,
                                                    const unsigned long long* toff, uint32_t* tiles,
                                                    uint32_t* heavy_bits) {
    __shared__ uint32_t cur[kTlMaxSeg];
    const uint64_t d = blockIdx.x;
    for (uint32_t i = threadIdx.x; i < ts.ns; i += kBlock) cur[i] = 0;
    __syncthreads();
    const uint64_t v0 = d << ts.db, v1 = min<uint64_t>(V, v0 + (1ull << ts.db));
    const unsigned long long* t = toff + d * ts.ns;
    const uint32_t smask = (1u << ts.sb) - 1;
    for (uint64_t v = v0 + threadIdx.x; v < v1; v += kBlock) {
        const OffT b = off[v], e = off[v + 1];
        if ((uint64_t)(e - b) >= heavy_thr) {
            atomicOr(&heavy_bits[v >> 5], 1u << (v & 31));
            continue;
        }
        const uint32_t dl = (uint32_t)(v - v0) << ts.sb;
        for (OffT k = b; k < e; k++) {
            const uint32_t u = src[k];
            const uint32_t s = u >> ts.sb;
            tiles[t[s] + atomicAdd(&cur[s], 1u)] = dl | (u & smask);
        }
    }
}

// Persistent: workgroup b takes destination blocks b, b + G, b + 2G, ...
// (G = gridDim.x).
// DB: the largest destination block (bits) its LDS sums hold; 2^14 blocks
// (64 KB) leave room for two workgroups per CU, 2^15 (128 KB) for one.
// Tiles are walked one by one (a flat walk across tile boundaries, each
// word's segment found by a binary search in LDS, measured 30% slower: the
// gathers of one pass then spread over several segments).  A tile is
// usually one pass (ER 2^26: ~512 words for 1024 lanes), so a pass is two
// dependent memory round trips (its words, then their gathers): the words of
// the next tile's first pass are loaded before the current tile's gathers,
// which leaves one round trip per tile.
template <int DB>
__global__ __launch_bounds__(kTlBlock) void k_pr_tiled(PrArgs a, const uint32_t* __restrict__ tiles,
                                                       const unsigned long long* __restrict__ toff, TileShape ts,
                                                       const uint32_t* heavy_bits) {
    __shared__ float acc[1u << DB];
    __shared__ double s_r[kTlBlock / kWave], s_d[kTlBlock / kWave];
    const uint32_t G = gridDim.x;
    const uint32_t nb = blockIdx.x < ts.nd ? (ts.nd - blockIdx.x + G - 1) / G : 0;   // blocks of this workgroup
    const uint32_t dsz = 1u << ts.db, smask = (1u << ts.sb) - 1;
    const double base = (1.0 - a.damping64) / (double)a.V + a.damping64 * (*a.dang_in) / (double)a.V;
    double res = 0, dang = 0;
    for (uint32_t r = 0; r < nb; r++) {
        const uint64_t d = blockIdx.x + (uint64_t)r * G;
        for (uint32_t i = threadIdx.x; i < dsz; i += kTlBlock) acc[i] = 0.0f;
        __syncthreads();
        const unsigned long long* t = toff + d * ts.ns;
        unsigned long long k0 = t[0], k1 = t[1];
        // first pass of a tile: four words per lane, lanes past its end load nothing
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const unsigned long long k = k0 + threadIdx.x + j * kTlBlock;
            w[j] = k < k1 ? tiles[k] : 0u;
        }
        for (uint32_t s = 0; s < ts.ns; s++) {
            const bool more = s + 1 < ts.ns;
            const unsigned long long k2 = more ? t[s + 2] : k1;
            uint32_t wn[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const unsigned long long k = k1 + threadIdx.x + j * kTlBlock;
                wn[j] = k < k2 ? tiles[k] : 0u;
            }
            const float* cseg = a.contrib_in + ((uint64_t)s << ts.sb);
            float c[4];
#pragma unroll
            for (int j = 0; j < 4; j++) c[j] = k0 + threadIdx.x + j * kTlBlock < k1 ? cseg[w[j] & smask] : 0.0f;
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (k0 + threadIdx.x + j * kTlBlock < k1) atomicAdd(&acc[w[j] >> ts.sb], c[j]);
            // further passes of a tile longer than one: the words, then their
            // gathers, then the LDS adds (lanes past the end re-read its first word)
            for (unsigned long long k = k0 + threadIdx.x + 4 * kTlBlock; k < k1; k += 4 * kTlBlock) {
                uint32_t x[4];
#pragma unroll
                for (int j = 0; j < 4; j++) x[j] = tiles[k + j * kTlBlock < k1 ? k + j * kTlBlock : k0];
                float cx[4];
#pragma unroll
                for (int j = 0; j < 4; j++) cx[j] = cseg[x[j] & smask];
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (k + j * kTlBlock < k1) atomicAdd(&acc[x[j] >> ts.sb], cx[j]);
            }
#pragma unroll
            for (int j = 0; j < 4; j++) w[j] = wn[j];
            k0 = k1;
            k1 = k2;
        }
        __syncthreads();
        const uint64_t v0 = d << ts.db;
        for (uint32_t i = threadIdx.x; i < dsz; i += kTlBlock) {
            const uint64_t v = v0 + i;
            if (v >= a.V) break;
            if (heavy_bits && (heavy_bits[v >> 5] >> (v & 31) & 1u)) continue;
            pr_finish_vertex(a, v, (double)acc[i], base, res, dang);
        }
        __syncthreads();
    }
    res = wave_sum(res);
    dang = wave_sum(dang);
    if (lane_id() == 0) {
        s_r[threadIdx.x / kWave] = res;
        s_d[threadIdx.x / kWave] = dang;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double rr = 0, dd = 0;
        for (int i = 0; i < kTlBlock / kWave; i++) {
            rr += s_r[i];
            dd += s_d[i];
        }
        if (rr != 0) atomicAdd(a.residual, rr);
        if (dd != 0) atomicAdd(a.dang_out, dd);
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

// Builds the propagation-blocked slot layout once (every edge resident,
// E < 2^32, at most kPbMaxBins bins, slot arrays within free device memory);
// pb_state = -1 when it cannot be built.
template <class OffT>
static int pr_blocked_build(asc_graph* g, PrState* p) {
    if (p->pb_state != 0) return ASC_OK;
    p->pb_state = -1;
    const uint64_t V = g->h.V, E = g->h.E;
    const uint64_t nb = (V + kPbBin - 1) / kPbBin;
    if (g->static_vertices != V || E == 0 || E >= 0xFFFFFFFFull || nb > kPbMaxBins) return ASC_OK;
    const bool sym = g->symmetric == 1;
    size_t free_b = 0, total_b = 0;
    ASC_HIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t keep = E * (4 + 4 + 2 + 4) + 12 * (uint64_t)kPbRanges * nb;
    const uint64_t temp = 8 * (uint64_t)kPbRanges * nb + 8 * (V + 1) + (sym ? 0 : 4 * E + 8 * (V + 1));
    if (keep + temp + (uint64_t(1) << 30) > free_b) return ASC_OK;
    hipStream_t s = g->ctx->s_main;
    const int grid = g->ctx->num_cus * 8;
    const OffT* in_off = (const OffT*)g->d_off;
    // out-CSR: the in-CSR itself when symmetric, else a transpose
    HostBuf<uint64_t> h_out;
    ASC_TRY(h_out.alloc(V + 1));
    h_out[0] = 0;
    for (uint64_t u = 0; u < V; u++) h_out[u + 1] = h_out[u] + g->h.outdeg[u];
    if (h_out[V] != E) {
        set_last_error("out-degrees sum to %llu, not E = %llu", (unsigned long long)h_out[V], (unsigned long long)E);
        return ASC_ERR_FORMAT;
    }
    // set-up temporaries (freed on every return): out-offsets, the transpose
    // of a directed graph and its fill cursors, the slot counts
    DevTmp<unsigned long long> d_out_off, cursor;
    DevTmp<uint32_t> transposed, d_cnt;
    ASC_HIP(hipMalloc(&d_out_off.p, 8 * (V + 1)));
    ASC_HIP(hipMemcpy(d_out_off.p, h_out.p, 8 * (V + 1), hipMemcpyHostToDevice));
    const uint32_t* out_dst = (const uint32_t*)g->d_edges;
    if (!sym) {
        ASC_HIP(hipMalloc(&transposed.p, 4 * E));
        ASC_HIP(hipMalloc(&cursor.p, 8 * (V + 1)));
        ASC_HIP(hipMemcpyAsync(cursor.p, d_out_off.p, 8 * (V + 1), hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL((k_pb_transpose<OffT>), dim3(grid), dim3(kBlock), 0, s, in_off,
                           (const uint32_t*)g->d_edges, V, cursor.p, transposed.p);
        ASC_HIP(hipGetLastError());
        ASC_HIP(hipStreamSynchronize(s));
        out_dst = transposed.p;
    }
    ASC_HIP(hipMalloc(&p->pb_src, 4 * E));
    hipLaunchKernelGGL((k_pb_src<unsigned long long>), dim3(grid), dim3(kBlock), 0, s, d_out_off.p, V, p->pb_src);
    ASC_HIP(hipGetLastError());
    // slot layout: counts per (bin, range), exclusive scan in bin-major order
    const uint32_t thr = (uint32_t)p->heavy_thr;
    ASC_HIP(hipMalloc(&d_cnt.p, 4 * (size_t)kPbRanges * nb));
    hipLaunchKernelGGL((k_pb_count<OffT>), dim3(std::min<uint32_t>(kPbRanges, grid)), dim3(kBlock), 0, s, out_dst, E,
                       in_off, thr, (uint32_t)nb, d_cnt.p);
    ASC_HIP(hipGetLastError());
    std::vector<uint32_t> cnt((size_t)kPbRanges * nb);
    ASC_HIP(hipMemcpyAsync(cnt.data(), d_cnt.p, 4 * cnt.size(), hipMemcpyDeviceToHost, s));
    ASC_HIP(hipStreamSynchronize(s));
    // range-major slots: the regions of one source range are adjacent, so
    // the scatter of a range writes one contiguous stretch
    std::vector<unsigned long long> base(cnt.size());
    unsigned long long acc = 0;
    for (uint32_t r = 0; r < kPbRanges; r++)
        for (uint64_t b = 0; b < nb; b++) {
            base[b * kPbRanges + r] = acc;
            acc += cnt[b * kPbRanges + r];
        }
    ASC_HIP(hipMalloc(&p->pb_bin_start, 8 * base.size()));   // region bases, [b * R + r]
    ASC_HIP(hipMemcpy(p->pb_bin_start, base.data(), 8 * base.size(), hipMemcpyHostToDevice));
    ASC_HIP(hipMalloc(&p->pb_cnt, 4 * cnt.size()));
    ASC_HIP(hipMemcpy(p->pb_cnt, cnt.data(), 4 * cnt.size(), hipMemcpyHostToDevice));
    ASC_HIP(hipMalloc(&p->pb_pos, 4 * E));
    ASC_HIP(hipMalloc(&p->pb_dl, 2 * std::max<unsigned long long>(acc, 1)));
    ASC_HIP(hipMalloc(&p->pb_vals, 4 * std::max<unsigned long long>(acc, 1)));
    hipLaunchKernelGGL((k_pb_assign<OffT>), dim3(std::min<uint32_t>(kPbRanges, grid)), dim3(kBlock), 0, s, out_dst,
                       E, in_off, thr, (uint32_t)nb, p->pb_bin_start, p->pb_pos, p->pb_dl);
    ASC_HIP(hipGetLastError());
    ASC_HIP(hipStreamSynchronize(s));
    p->pb_nbins = (uint32_t)nb;
    p->pb_edges = E;
    p->pb_state = 1;
    return ASC_OK;
}

// Builds the tiles of the source-segmented mode for segments of 2^sb and
// blocks of 2^db vertices (every edge resident, at most kTlMaxSeg segments,
// sb + db <= 32, tile words within free device memory); rebuilt when the
// shape changes.  tl_state = -1 when they cannot be built.
template <class OffT>
static int pr_tiled_build(asc_graph* g, PrState* p, uint32_t sb, uint32_t db) {
    if (p->tl_state == 1 && p->tl_sb == sb && p->tl_db == db) return ASC_OK;
    auto f = [](void* x) {
        if (x) (void)hipFree(x);
    };
    f(p->tl_edges); f(p->tl_off); f(p->tl_heavy);
    p->tl_edges = nullptr;
    p->tl_off = nullptr;
    p->tl_heavy = nullptr;
    p->tl_state = -1;
    const uint64_t V = g->h.V, E = g->h.E;
    if (db < 6 || db > (uint32_t)kTlMaxDstBits || sb < 8 || sb + db > 32) return ASC_OK;
    const uint64_t ns = (V + (1ull << sb) - 1) >> sb, nd = (V + (1ull << db) - 1) >> db;
    if (g->static_vertices != V || E == 0 || ns > kTlMaxSeg || nd > 0xFFFFFFFFull) return ASC_OK;
    size_t free_b = 0, total_b = 0;
    ASC_HIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t need = 4 * E + 12 * nd * ns + 8 + (V + 31) / 32 * 4;
    if (need + (uint64_t(1) << 30) > free_b) return ASC_OK;
    hipStream_t s = g->ctx->s_main;
    const OffT* off = (const OffT*)g->d_off;
    const uint32_t* src = (const uint32_t*)g->d_edges;
    const uint32_t thr = (uint32_t)p->heavy_thr;
    const TileShape ts{sb, db, (uint32_t)ns, (uint32_t)nd};
    DevTmp<uint32_t> d_cnt;
    ASC_HIP(hipMalloc(&d_cnt.p, 4 * nd * ns));
    hipLaunchKernelGGL((k_tl_count<OffT>), dim3((uint32_t)nd), dim3(kBlock), 0, s, off, src, V, thr, ts, d_cnt.p);
    ASC_HIP(hipGetLastError());
    std::vector<uint32_t> cnt(nd * ns);
    ASC_HIP(hipMemcpyAsync(cnt.data(), d_cnt.p, 4 * cnt.size(), hipMemcpyDeviceToHost, s));
    ASC_HIP(hipStreamSynchronize(s));
    std::vector<unsigned long long> toff(nd * ns + 1);
    unsigned long long acc = 0;
    for (size_t i = 0; i < cnt.size(); i++) {
        toff[i] = acc;
        acc += cnt[i];
    }
    toff[cnt.size()] = acc;
    ASC_HIP(hipMalloc(&p->tl_off, 8 * toff.size()));
    ASC_HIP(hipMemcpy(p->tl_off, toff.data(), 8 * toff.size(), hipMemcpyHostToDevice));
    ASC_HIP(hipMalloc(&p->tl_edges, 4 * std::max<unsigned long long>(acc, 1)));
    ASC_HIP(hipMalloc(&p->tl_heavy, 4 * ((V + 31) / 32)));
    ASC_HIP(hipMemsetAsync(p->tl_heavy, 0, 4 * ((V + 31) / 32), s));
    hipLaunchKernelGGL((k_tl_fill<OffT>), dim3((uint32_t)nd), dim3(kBlock), 0, s, off, src, V, thr, ts,
                       (const unsigned long long*)p->tl_off, p->tl_edges, p->tl_heavy);
    ASC_HIP(hipGetLastError());
    ASC_HIP(hipStreamSynchronize(s));
    p->tl_sb = sb;
    p->tl_db = db;
    p->tl_nd = (uint32_t)nd;
    p->tl_ns = (uint32_t)ns;
    p->tl_edges_n = acc;
    p->tl_state = 1;
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
    if (o->method == ASC_PR_BLOCKED) ASC_TRY(pr_blocked_build<OffT>(g, p));
    const bool blocked = o->method == ASC_PR_BLOCKED && p->pb_state == 1;
    if (o->method == ASC_PR_BLOCKED && !blocked) {
        set_last_error("propagation-blocked PageRank needs every edge resident, E < 2^32 and room for its slots");
        return ASC_ERR_UNSUPPORTED;
    }
    // tile shape: 2^15-vertex segments (128 KB contribution slices) measured
    // fastest on Erdos-Renyi 2^26 / 2^31 edges (20 iterations: 496 ms with
    // 2^14-vertex blocks against 557-833 ms for 2^14 or 2^16-2^20-vertex
    // segments, and 821 ms for the pull); 2^15-vertex blocks (one workgroup
    // per CU, 128 KB of LDS sums, tiles twice as long) 468 ms, taken once
    // there are at least four blocks per CU; the graph's tuned shape or the
    // caller's wins
    const uint32_t sb = o->tile_src_bits > 0 ? (uint32_t)o->tile_src_bits
                        : g->pr_tile_src_bits > 0 ? (uint32_t)g->pr_tile_src_bits : 15;
    const uint32_t db_auto = (V >> 15) >= 4 * (uint64_t)ctx->num_cus ? 15u : 14u;
    const uint32_t db = o->tile_dst_bits > 0 ? (uint32_t)o->tile_dst_bits
                        : g->pr_tile_dst_bits > 0 ? (uint32_t)g->pr_tile_dst_bits : db_auto;
    // AUTO: the tiled pull once the contributions outgrow an XCD's L2
    const bool want_tiled = o->method == ASC_PR_TILED ||
                            (o->method == ASC_PR_AUTO && g->static_vertices == V && V >= (uint64_t(1) << 20));
    if (want_tiled) ASC_TRY(pr_tiled_build<OffT>(g, p, sb, db));
    const bool tiled = want_tiled && p->tl_state == 1;
    if (o->method == ASC_PR_TILED && !tiled) {
        set_last_error("source-segmented PageRank needs every edge resident, V <= 2^31 and room for its tiles");
        return ASC_ERR_UNSUPPORTED;
    }
    ScopedEvent e0, e1;
    ASC_HIP(hipEventCreate(&e0.e));
    ASC_HIP(hipEventCreate(&e1.e));
    const double t0 = now_ms();
    ASC_HIP(hipEventRecord(e0.e, s));
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
        if (tiled) {
            const TileShape ts{p->tl_sb, p->tl_db, p->tl_ns, p->tl_nd};
            // 1024-lane workgroups: two per CU with blocks up to 2^14
            // destinations (64 KB of LDS each), one with 2^15 (128 KB)
            const bool big = p->tl_db > 14;
            const uint32_t tgrid = (uint32_t)std::min<uint64_t>((uint64_t)ctx->num_cus * (big ? 1 : 2), p->tl_nd);
            const uint32_t* hb = p->n_heavy ? (const uint32_t*)p->tl_heavy : nullptr;
            if (big)
                hipLaunchKernelGGL(k_pr_tiled<15>, dim3(tgrid), dim3(kTlBlock), 0, s, a,
                                   (const uint32_t*)p->tl_edges, (const unsigned long long*)p->tl_off, ts, hb);
            else
                hipLaunchKernelGGL(k_pr_tiled<14>, dim3(tgrid), dim3(kTlBlock), 0, s, a,
                                   (const uint32_t*)p->tl_edges, (const unsigned long long*)p->tl_off, ts, hb);
            ASC_HIP(hipGetLastError());
        } else if (blocked) {
            hipLaunchKernelGGL(k_pb_scatter, dim3(grid), dim3(kBlock), 0, s, p->pb_src, p->pb_pos, p->pb_edges,
                               a.contrib_in, p->pb_vals);
            hipLaunchKernelGGL((k_pb_gather<OffT>), dim3(p->pb_nbins), dim3(kPbBlock), 0, s, a, p->pb_bin_start,
                               p->pb_cnt, p->pb_dl, p->pb_vals);
            ASC_HIP(hipGetLastError());
        } else if (a.v_end > 0) {
            ASC_TRY(launch_pull<OffT>(G, grid, s, a));
        }
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
    ASC_HIP(hipEventRecord(e1.e, s));
    ASC_HIP(hipStreamSynchronize(s));
    const double t1 = now_ms();
    float kms = 0;
    (void)hipEventElapsedTime(&kms, e0.e, e1.e);
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
    asc_stats local;
    if (!st && g->ctx->cfg.verbose) st = &local;
    int r;
    if (o->mode == ASC_PR_REFERENCE)
        r = g->off_bytes == 4 ? pr_reference<uint32_t>(g, o, rank_f64, st) : pr_reference<uint64_t>(g, o, rank_f64, st);
    else
        r = g->off_bytes == 4 ? pr_normalised<uint32_t>(g, o, rank_f32, residual_out, st)
                              : pr_normalised<uint64_t>(g, o, rank_f32, residual_out, st);
    // per-phase timings in the reference's wording (CalculateOpt.cu:700-720)
    if (r == ASC_OK && g->ctx->cfg.verbose && st)
        printf("totalProcess time is %.3f ms\nstaticProcess time is %.3f ms\noverloadProcess time is %.3f ms\n"
               "iterations : %u\n",
               st->total_ms, st->static_ms, st->overload_ms, st->iterations);
    return r;
}

}  // namespace asc
