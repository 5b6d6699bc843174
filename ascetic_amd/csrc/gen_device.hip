// Synthetic graphs built on the GPU: the CSR the host generator
// (generate.cpp) builds, bit for bit, from the same samples
// (gen_sampler.hpp), in seconds instead of minutes at R-MAT scale 26.
//
// 1. k_gen_count: raw out-degree of every vertex over all samples (both
//    directions of a symmetrized sample, self-loops dropped).
// 2. The vertices are cut into ranges whose raw edges fit the device (two
//    8-byte keys per raw edge); per range:
//    k_gen_keys   every raw edge (u, v) with u in the range as the key
//                 (u - lo) << vbits | v, appended with one atomic per wave;
//    radix sort   of the keys over their rbits + vbits bits (rocPRIM);
//    unique       drops duplicate pairs (rocPRIM), which leaves each
//                 adjacency sorted and deduplicated;
//    k_gen_bounds first and last position of every vertex's run -> degrees;
//    k_gen_emit   the records ({dst, w} with the pair weight, or dst alone),
//                 copied to the host CSR at the range's offset.
// Every sample is drawn again per range (a pure function of its index), so
// nothing of the sample list is stored.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>

#include "device_common.hpp"
#include "gen_sampler.hpp"
#include "internal.hpp"

namespace asc {

namespace {

constexpr int kGenBlock = 256;

__global__ __launch_bounds__(kGenBlock) void k_gen_count(Sampler s, uint64_t M, int sym, uint32_t* cnt) {
    for (uint64_t i = blockIdx.x * (uint64_t)kGenBlock + threadIdx.x; i < M; i += (uint64_t)gridDim.x * kGenBlock) {
        uint32_t u, v;
        s.sample(i, u, v);
        if (u == v) continue;
        atomicAdd(&cnt[u], 1u);
        if (sym) atomicAdd(&cnt[v], 1u);
    }
}

// Keys are staged per workgroup in LDS and flushed to the global array with
// one atomic per flush: a single global counter taking one atomic per wave
// serialised the pass (measured: 6 s per pass over 2^34 samples, against
// 2.5 s for the count pass).  Keys past cap are counted but not stored (the
// host then reports the mismatch).
constexpr uint32_t kGenStage = 4096;

__device__ __forceinline__ void gen_stage(bool want, unsigned long long key, unsigned long long* buf,
                                          uint32_t* n) {
    const uint64_t m = wave_ballot(want);
    if (!m) return;
    const uint32_t leader = (uint32_t)__builtin_ctzll(m);
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(n, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader, kWave);
    if (want) buf[base + lanes_below(m)] = key;
}

__device__ __forceinline__ void gen_flush(unsigned long long* buf, uint32_t* n, unsigned long long* s_base,
                                          unsigned long long* keys, unsigned long long* count,
                                          unsigned long long cap) {
    __syncthreads();
    const uint32_t cnt = *n;
    if (threadIdx.x == 0) *s_base = cnt ? atomicAdd(count, (unsigned long long)cnt) : 0;
    __syncthreads();
    const unsigned long long base = *s_base;
    for (uint32_t i = threadIdx.x; i < cnt; i += kGenBlock)
        if (base + i < cap) keys[base + i] = buf[i];
    __syncthreads();
    if (threadIdx.x == 0) *n = 0;
    __syncthreads();
}

__global__ __launch_bounds__(kGenBlock) void k_gen_keys(Sampler s, uint64_t M, int sym, uint32_t lo, uint32_t hi,
                                                        uint32_t vbits, unsigned long long* keys,
                                                        unsigned long long* count, unsigned long long cap) {
    __shared__ unsigned long long buf[kGenStage];
    __shared__ uint32_t n;
    __shared__ unsigned long long s_base;
    if (threadIdx.x == 0) n = 0;
    __syncthreads();
    // the loop bound is uniform per workgroup, so every lane takes part in
    // the ballots and barriers; an iteration adds at most 2 * kGenBlock keys
    for (uint64_t i0 = blockIdx.x * (uint64_t)kGenBlock; i0 < M; i0 += (uint64_t)gridDim.x * kGenBlock) {
        const uint64_t i = i0 + threadIdx.x;
        uint32_t u = 0, v = 0;
        bool fwd = false, rev = false;
        if (i < M) {
            s.sample(i, u, v);
            if (u != v) {
                fwd = u >= lo && u < hi;
                rev = sym && v >= lo && v < hi;
            }
        }
        gen_stage(fwd, ((unsigned long long)(u - lo) << vbits) | v, buf, &n);
        gen_stage(rev, ((unsigned long long)(v - lo) << vbits) | u, buf, &n);
        __syncthreads();
        if (n > kGenStage - 2 * kGenBlock) gen_flush(buf, &n, &s_base, keys, count, cap);
    }
    gen_flush(buf, &n, &s_base, keys, count, cap);
}

// first[r] / last[r]: start and end of range vertex r's run in the sorted
// unique keys (both 0 for a vertex without edges).
__global__ __launch_bounds__(kGenBlock) void k_gen_bounds(const unsigned long long* keys, unsigned long long n,
                                                          uint32_t vbits, unsigned long long* first,
                                                          unsigned long long* last) {
    for (uint64_t i = blockIdx.x * (uint64_t)kGenBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kGenBlock) {
        const unsigned long long r = keys[i] >> vbits;
        if (i == 0 || (keys[i - 1] >> vbits) != r) first[r] = i;
        if (i + 1 == n || (keys[i + 1] >> vbits) != r) last[r] = i + 1;
    }
}

__global__ __launch_bounds__(kGenBlock) void k_gen_degrees(const unsigned long long* first,
                                                           const unsigned long long* last, uint32_t n,
                                                           uint32_t* deg) {
    for (uint64_t r = blockIdx.x * (uint64_t)kGenBlock + threadIdx.x; r < n; r += (uint64_t)gridDim.x * kGenBlock)
        deg[r] = (uint32_t)(last[r] - first[r]);
}

// Records of the sorted unique keys: {dst, w(u, dst)} (words = 2) or dst.
__global__ __launch_bounds__(kGenBlock) void k_gen_emit(const unsigned long long* keys, unsigned long long n,
                                                        uint32_t vbits, uint32_t lo, int weighted,
                                                        unsigned long long wkey, uint32_t wmin, uint32_t span,
                                                        uint32_t* out) {
    const unsigned long long vmask = (1ull << vbits) - 1;
    for (uint64_t i = blockIdx.x * (uint64_t)kGenBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kGenBlock) {
        const unsigned long long k = keys[i];
        const uint32_t v = (uint32_t)(k & vmask);
        if (weighted) {
            const uint32_t u = lo + (uint32_t)(k >> vbits);
            out[2 * i] = v;
            out[2 * i + 1] = pair_weight(wkey, u, v, wmin, span);
        } else {
            out[i] = v;
        }
    }
}

// Records of a generated graph: edge g = base + i goes to the device's static
// prefix when g < S, else to the tail staging buffer at g - stage_base, as a
// packed word dst | (w - wmin) << pk_shift, or (split) the id and a weight
// byte.  The weights are summed into *wsum (one atomic per wave).
__global__ __launch_bounds__(kGenBlock) void k_gen_emit_graph(
    const unsigned long long* keys, unsigned long long n, uint32_t vbits, uint32_t lo, unsigned long long base,
    unsigned long long S, WEdge* d_edges, uint32_t* stage32, uint8_t* stage8, unsigned long long stage_base,
    unsigned long long wkey, uint32_t wmin, uint32_t span, uint32_t pk_shift, unsigned long long* wsum) {
    const unsigned long long vmask = (1ull << vbits) - 1;
    unsigned long long ws = 0;
    for (uint64_t i0 = blockIdx.x * (uint64_t)kGenBlock; i0 < n; i0 += (uint64_t)gridDim.x * kGenBlock) {
        const uint64_t i = i0 + threadIdx.x;
        if (i < n) {
            const unsigned long long k = keys[i];
            const uint32_t v = (uint32_t)(k & vmask);
            const uint32_t u = lo + (uint32_t)(k >> vbits);
            const uint32_t w = pair_weight(wkey, u, v, wmin, span);
            ws += w;
            const unsigned long long g = base + i;
            if (g < S) {
                d_edges[g] = WEdge{v, w};
            } else if (stage8) {
                stage32[g - stage_base] = v;
                stage8[g - stage_base] = (uint8_t)(w - wmin);
            } else {
                stage32[g - stage_base] = v | ((w - wmin) << pk_shift);
            }
        }
    }
    ws = wave_sum(ws);
    if (lane_id() == 0 && ws) atomicAdd(wsum, ws);
}

uint32_t bits_for(uint64_t n) { return n > 1 ? 64 - (uint32_t)__builtin_clzll(n - 1) : 1; }

}  // namespace

// Per-range machinery shared by the host-CSR and the graph consumers.
class DeviceGen {
public:
    asc_ctx* ctx = nullptr;
    hipStream_t st = nullptr;
    Sampler s{};
    uint64_t M = 0, V = 0, raw_total = 0;
    bool sym = false;
    uint32_t vbits = 1;
    int grid = 0;
    HostBuf<uint32_t> cnt;            // raw out-degree per vertex
    std::vector<uint32_t> degh;       // deduplicated degrees of the last range
    std::vector<uint64_t> cuts;       // range boundaries

    int init(asc_ctx* c, const asc_gen_params* p) {
        ctx = c;
        ASC_TRY(make_sampler(p, s, M));
        sym = p->symmetrize != 0;
        V = s.V;
        vbits = bits_for(V);
        ASC_HIP(hipSetDevice(ctx->device));
        st = ctx->s_main;
        grid = ctx->num_cus * 8;
        return ASC_OK;
    }

    // Raw out-degrees of every vertex (k_gen_count) into `cnt`.
    int count() {
        ASC_TRY(cnt.alloc(V));
        DevTmp<uint32_t> d;
        if (hipMalloc(&d.p, 4 * V) != hipSuccess) {
            set_last_error("device allocation of %llu count bytes failed", (unsigned long long)(4 * V));
            return ASC_ERR_NOMEM;
        }
        ASC_HIP(hipMemsetAsync(d.p, 0, 4 * V, st));
        hipLaunchKernelGGL(k_gen_count, dim3(grid), dim3(kGenBlock), 0, st, s, M, (int)sym, d.p);
        ASC_HIP(hipGetLastError());
        ASC_HIP(hipMemcpyAsync(cnt.p, d.p, 4 * V, hipMemcpyDeviceToHost, st));
        ASC_HIP(hipStreamSynchronize(st));
        raw_total = 0;
        for (uint64_t v = 0; v < V; v++) raw_total += cnt[v];
        return ASC_OK;
    }

    // Cuts the vertices into ranges whose keys (two 8-byte buffers) fit
    // three quarters of the free device memory (capped by the context's
    // device_mem_limit) once the per-vertex bound arrays are set aside, and
    // allocates the buffers of the largest range.
    int plan_and_alloc() {
        release();
        size_t free_b = 0, total_b = 0;
        ASC_HIP(hipMemGetInfo(&free_b, &total_b));
        if (ctx->cfg.device_mem_limit) free_b = std::min<size_t>(free_b, ctx->cfg.device_mem_limit);
        const uint64_t per_vertex = 20 * V + (64u << 20);
        const uint64_t usable = free_b > per_vertex ? (free_b - per_vertex) / 4 * 3 : 0;
        const uint64_t cap_keys = std::max<uint64_t>(usable / 16, 1);
        cuts.assign(1, 0);
        uint64_t acc = 0;
        for (uint64_t v = 0; v < V; v++) {
            if (cnt[v] > cap_keys) {
                set_last_error("vertex %llu has %u raw edges, more than fit the device at once",
                               (unsigned long long)v, cnt[v]);
                return ASC_ERR_NOMEM;
            }
            if (acc + cnt[v] > cap_keys) {
                cuts.push_back(v);
                acc = 0;
            }
            acc += cnt[v];
        }
        cuts.push_back(V);
        uint64_t max_raw = 0, max_nv = 0;
        for (size_t r = 0; r + 1 < cuts.size(); r++) {
            max_raw = std::max(max_raw, raw_in(r));
            max_nv = std::max(max_nv, cuts[r + 1] - cuts[r]);
        }
        const uint64_t kbytes = 8 * std::max<uint64_t>(max_raw, 1);
        if (hipMalloc(&ka.p, kbytes) != hipSuccess || hipMalloc(&kb.p, kbytes) != hipSuccess ||
            hipMalloc(&first.p, 8 * max_nv) != hipSuccess || hipMalloc(&last.p, 8 * max_nv) != hipSuccess ||
            hipMalloc(&deg.p, 4 * max_nv) != hipSuccess || hipMalloc(&dcount.p, 16) != hipSuccess) {
            set_last_error("device allocation of the generator's %llu key bytes failed",
                           (unsigned long long)(2 * kbytes));
            return ASC_ERR_NOMEM;
        }
        key_cap = kbytes / 8;
        size_t sort_bytes = 0, uniq_bytes = 0;
        rocprim::double_buffer<unsigned long long> db(ka.p, kb.p);
        ASC_HIP(rocprim::radix_sort_keys(nullptr, sort_bytes, db, (size_t)max_raw, 0, 64, st));
        ASC_HIP(rocprim::unique(nullptr, uniq_bytes, ka.p, kb.p, dcount.p + 1, (size_t)max_raw,
                                rocprim::equal_to<unsigned long long>(), st));
        tmp_bytes = std::max<size_t>(std::max(sort_bytes, uniq_bytes), 16);
        if (hipMalloc(&tmp.p, tmp_bytes) != hipSuccess) {
            set_last_error("device allocation of %zu sort bytes failed", tmp_bytes);
            return ASC_ERR_NOMEM;
        }
        degh.assign(max_nv, 0);
        return ASC_OK;
    }

    size_t ranges() const { return cuts.size() - 1; }
    uint64_t raw_in(size_t r) const {
        uint64_t n = 0;
        for (uint64_t v = cuts[r]; v < cuts[r + 1]; v++) n += cnt[v];
        return n;
    }

    // Range r: its sorted unique keys (*uniq, *nu), the degrees of its
    // vertices in degh, and a spare buffer of the same capacity (*spare,
    // 8 bytes per raw key) for records.
    int range(size_t r, unsigned long long** uniq, unsigned long long** spare, uint64_t* nu) {
        const uint64_t lo = cuts[r], hi = cuts[r + 1], nv = hi - lo, nraw = raw_in(r);
        *nu = 0;
        *uniq = ka.p;
        *spare = kb.p;
        if (!nraw) {
            std::fill(degh.begin(), degh.begin() + nv, 0u);
            return ASC_OK;
        }
        ASC_HIP(hipMemsetAsync(dcount.p, 0, 16, st));
        hipLaunchKernelGGL(k_gen_keys, dim3(grid), dim3(kGenBlock), 0, st, s, M, (int)sym, (uint32_t)lo,
                           (uint32_t)hi, vbits, ka.p, dcount.p, (unsigned long long)nraw);
        ASC_HIP(hipGetLastError());
        unsigned long long got = 0;
        ASC_HIP(hipMemcpyAsync(&got, dcount.p, 8, hipMemcpyDeviceToHost, st));
        ASC_HIP(hipStreamSynchronize(st));
        if (got != nraw) {
            set_last_error("generator: %llu keys drawn for a range counted at %llu", got, (unsigned long long)nraw);
            return ASC_ERR_DEVICE;
        }
        rocprim::double_buffer<unsigned long long> db(ka.p, kb.p);
        size_t b = tmp_bytes;
        ASC_HIP(rocprim::radix_sort_keys(tmp.p, b, db, (size_t)nraw, 0, bits_for(nv) + vbits, st));
        unsigned long long* sorted = db.current();
        unsigned long long* out = sorted == ka.p ? kb.p : ka.p;
        b = tmp_bytes;
        ASC_HIP(rocprim::unique(tmp.p, b, sorted, out, dcount.p + 1, (size_t)nraw,
                                rocprim::equal_to<unsigned long long>(), st));
        unsigned long long n = 0;
        ASC_HIP(hipMemcpyAsync(&n, dcount.p + 1, 8, hipMemcpyDeviceToHost, st));
        ASC_HIP(hipMemsetAsync(first.p, 0, 8 * nv, st));
        ASC_HIP(hipMemsetAsync(last.p, 0, 8 * nv, st));
        ASC_HIP(hipStreamSynchronize(st));
        hipLaunchKernelGGL(k_gen_bounds, dim3(grid), dim3(kGenBlock), 0, st, out, n, vbits, first.p, last.p);
        hipLaunchKernelGGL(k_gen_degrees, dim3(grid), dim3(kGenBlock), 0, st, first.p, last.p, (uint32_t)nv, deg.p);
        ASC_HIP(hipGetLastError());
        ASC_HIP(hipMemcpyAsync(degh.data(), deg.p, 4 * nv, hipMemcpyDeviceToHost, st));
        ASC_HIP(hipStreamSynchronize(st));
        uint64_t sum = 0;
        for (uint64_t i = 0; i < nv; i++) sum += degh[i];
        if (sum != n) {
            set_last_error("generator: degrees of a range sum to %llu, not %llu", (unsigned long long)sum, n);
            return ASC_ERR_DEVICE;
        }
        *uniq = out;
        *spare = sorted;
        *nu = n;
        return ASC_OK;
    }

    uint64_t key_capacity() const { return key_cap; }

    void release() {
        for (DevTmp<unsigned long long>* d : {&ka, &kb, &first, &last, &dcount}) {
            if (d->p) (void)hipFree(d->p);
            d->p = nullptr;
        }
        if (deg.p) (void)hipFree(deg.p);
        deg.p = nullptr;
        if (tmp.p) (void)hipFree(tmp.p);
        tmp.p = nullptr;
    }

private:
    DevTmp<unsigned long long> ka, kb, first, last, dcount;
    DevTmp<uint32_t> deg;
    DevTmp<unsigned char> tmp;
    size_t tmp_bytes = 0;
    uint64_t key_cap = 0;
};

int generate_on_device(asc_ctx* ctx, const asc_gen_params* p, HostCSR& out) {
    if (!p) return ASC_ERR_INVALID;
    if (p->csc && !p->symmetrize) {
        set_last_error("device generation builds the in-CSR of symmetrized graphs only");
        return ASC_ERR_UNSUPPORTED;
    }
    DeviceGen gen;
    ASC_TRY(gen.init(ctx, p));
    const double t0 = now_ms();
    ASC_TRY(gen.count());
    ASC_TRY(gen.plan_and_alloc());
    const uint64_t V = gen.V;
    const int words = p->csc ? 1 : (p->weighted ? 2 : 1);
    out.V = V;
    out.kind = p->csc ? ASC_KIND_CSC : (p->weighted ? ASC_KIND_WCSR : ASC_KIND_CSR);
    ASC_TRY(out.off.alloc(V + 1));
    // raw_total bounds the deduplicated edges; pages past E are never touched
    ASC_TRY(out.edges.alloc((size_t)gen.raw_total * words + 1));
    const uint32_t span = p->weighted ? p->wmax - p->wmin + 1 : 1;
    const uint64_t wkey = weight_key(p->seed);
    uint64_t E = 0;
    out.off[0] = 0;
    for (size_t r = 0; r < gen.ranges(); r++) {
        const uint64_t lo = gen.cuts[r], nv = gen.cuts[r + 1] - lo;
        unsigned long long *uniq, *spare;
        uint64_t nu = 0;
        ASC_TRY(gen.range(r, &uniq, &spare, &nu));
        if (nu) {
            uint32_t* rec = (uint32_t*)spare;
            hipLaunchKernelGGL(k_gen_emit, dim3(gen.grid), dim3(kGenBlock), 0, gen.st, uniq, (unsigned long long)nu,
                               gen.vbits, (uint32_t)lo, (int)(words == 2), (unsigned long long)wkey, p->wmin, span,
                               rec);
            ASC_HIP(hipGetLastError());
            ASC_HIP(hipMemcpyAsync(out.edges.p + E * words, rec, 4 * words * nu, hipMemcpyDeviceToHost, gen.st));
            ASC_HIP(hipStreamSynchronize(gen.st));
        }
        for (uint64_t i = 0; i < nv; i++) {
            E += gen.degh[i];
            out.off[lo + i + 1] = E;
        }
    }
    out.E = E;
    if (p->csc) {
        ASC_TRY(out.outdeg.alloc(V));
        for (uint64_t v = 0; v < V; v++) out.outdeg[v] = (uint32_t)(out.off[v + 1] - out.off[v]);
    }
    if (ctx->cfg.verbose)
        printf("device generation: V=%llu E=%llu (raw %llu) in %zu range(s), %.1f ms\n", (unsigned long long)V,
               (unsigned long long)E, (unsigned long long)gen.raw_total, gen.ranges(), now_ms() - t0);
    return ASC_OK;
}

// A weighted graph built straight into the device's static prefix and the
// pinned, packed on-demand tail: the host never holds its {dst, w} records.
// Two sweeps over the ranges: the first yields the deduplicated degrees (the
// offsets, then the static/on-demand plan), the second the records.
int graph_generate(asc_ctx* ctx, const asc_gen_params* p, asc_graph** out) {
    if (!p->weighted || p->csc) {
        set_last_error("asc_graph_generate builds weighted out-CSR graphs");
        return ASC_ERR_UNSUPPORTED;
    }
    DeviceGen gen;
    ASC_TRY(gen.init(ctx, p));
    const double t0 = now_ms();
    ASC_TRY(gen.count());
    ASC_TRY(gen.plan_and_alloc());
    struct Owner {
        asc_graph* g = new asc_graph();
        ~Owner() {
            if (g) {
                graph_free_device(g);
                delete g;
            }
        }
    } own;
    asc_graph* g = own.g;
    const uint64_t V = gen.V;
    g->ctx = ctx;
    g->h.V = V;
    g->h.kind = ASC_KIND_WCSR;
    ASC_TRY(g->h.off.alloc(V + 1));
    // sweep 1: degrees -> offsets
    uint64_t E = 0;
    g->h.off[0] = 0;
    for (size_t r = 0; r < gen.ranges(); r++) {
        unsigned long long *uniq, *spare;
        uint64_t nu = 0;
        ASC_TRY(gen.range(r, &uniq, &spare, &nu));
        const uint64_t lo = gen.cuts[r], nv = gen.cuts[r + 1] - lo;
        for (uint64_t i = 0; i < nv; i++) {
            E += gen.degh[i];
            g->h.off[lo + i + 1] = E;
        }
    }
    gen.release();
    const double t_deg = now_ms();
    g->h.E = E;
    g->off_bytes = (E >= (uint64_t(1) << 32) || p->offset_bytes == 8) ? 8 : 4;
    g->edge_budget_bytes = ctx->cfg.edge_budget_bytes;
    g->advise_rate = ctx->cfg.advise_rate;
    g->heavy_threshold = ctx->cfg.heavy_threshold;
    g->pr_group = ctx->cfg.pr_group;
    g->sssp_delta = ctx->cfg.sssp_delta;
    compute_max_degree(g);

    // plan, as graph_upload does for a host graph
    size_t free_b = 0, total_b = 0;
    ASC_HIP(hipMemGetInfo(&free_b, &total_b));
    if (ctx->cfg.device_mem_limit) free_b = std::min<size_t>(free_b, ctx->cfg.device_mem_limit);
    asc_plan plan;
    ASC_TRY(asc_plan_partition(V, E, g->h.off.p, 8, free_b, param_size_for(ASC_KIND_WCSR), g->advise_rate,
                               g->edge_budget_bytes, ctx->cfg.part_bytes, &plan));
    const uint64_t S = plan.static_edges, n_tail = E - S;
    const uint64_t need = (uint64_t)g->off_bytes * (V + 1) + std::max<uint64_t>(8 * S, 64);
    if (ctx->cfg.device_mem_limit && need > ctx->cfg.device_mem_limit) {
        set_last_error("the planned device copies need %llu bytes, above the %llu-byte device limit",
                       (unsigned long long)need, (unsigned long long)ctx->cfg.device_mem_limit);
        return ASC_ERR_NOMEM;
    }
    g->static_edges = S;
    g->static_vertices = plan.static_vertices;
    g->part_edges = plan.part_edges;
    g->advise_k = plan.advise_k;
    // the tail's packed form (as graph_upload chooses it)
    const uint32_t dbits = V > 1 ? 64 - (uint32_t)__builtin_clzll(V - 1) : 1;
    const uint32_t wspan = p->wmax - p->wmin;
    const uint32_t wbits = wspan ? 32 - (uint32_t)__builtin_clz(wspan) : 0;
    bool split = false;
    if (n_tail) {
        if (!ctx->cfg.stream_pack) {
            set_last_error("a generated graph keeps its on-demand edges packed; stream_pack 0 is not available");
            return ASC_ERR_UNSUPPORTED;
        }
        if (dbits + wbits <= 32 && dbits < 32 && ctx->cfg.stream_pack != 2) split = false;
        else if (wbits <= 8) split = true;
        else {
            set_last_error("ids and weights of the generated graph fit neither packed form");
            return ASC_ERR_UNSUPPORTED;
        }
    }
    // device copies
    const size_t ob = g->off_bytes;
    if (hipMalloc(&g->d_off, ob * (V + 1)) != hipSuccess) {
        set_last_error("device allocation of %zu offset bytes failed", ob * (V + 1));
        return ASC_ERR_NOMEM;
    }
    if (ob == 8) {
        ASC_HIP(hipMemcpy(g->d_off, g->h.off.p, 8 * (V + 1), hipMemcpyHostToDevice));
    } else {
        std::vector<uint32_t> o32(V + 1);
        for (uint64_t i = 0; i <= V; i++) o32[i] = (uint32_t)g->h.off[i];
        ASC_HIP(hipMemcpy(g->d_off, o32.data(), 4 * (V + 1), hipMemcpyHostToDevice));
    }
    const size_t sbytes = std::max<size_t>(8 * S, 64);
    if (hipMalloc(&g->d_edges, sbytes) != hipSuccess) {
        set_last_error("device allocation of %zu static edge bytes failed", sbytes);
        return ASC_ERR_NOMEM;
    }
    g->device_bytes = ob * (V + 1) + sbytes;
    if (n_tail) {
        // split form: ids and weight bytes page-locked together (5 bytes an
        // edge against the cap) or not at all
        const uint64_t cap = ctx->cfg.max_pinned_bytes;
        if (split) {
            ASC_TRY(!cap || 5 * n_tail <= cap ? g->ptail.alloc_prefer_pinned(n_tail, 0) : g->ptail.alloc(n_tail, false));
            ASC_TRY(g->ptail.pinned ? g->ptail_w.alloc_prefer_pinned(n_tail, 0) : g->ptail_w.alloc(n_tail, false));
        } else {
            ASC_TRY(g->ptail.alloc_prefer_pinned(n_tail, cap));
        }
        const uint64_t pb = ctx->cfg.part_bytes ? ctx->cfg.part_bytes : (uint64_t(64) << 20);
        g->packed = true;
        g->pk_split = split;
        g->pk_shift = split ? 0 : dbits;
        g->pk_wmin = p->wmin;
        g->part_edges = std::max<uint64_t>(pb / (split ? 5 : 4), 1024);
    }
    const double t_alloc = now_ms();
    // sweep 2: records
    ASC_TRY(gen.plan_and_alloc());
    DevTmp<unsigned long long> wsum;
    ASC_HIP(hipMalloc(&wsum.p, 8));
    ASC_HIP(hipMemsetAsync(wsum.p, 0, 8, gen.st));
    const uint32_t span = p->wmax - p->wmin + 1;
    const uint64_t wkey = weight_key(p->seed);
    for (size_t r = 0; r < gen.ranges(); r++) {
        unsigned long long *uniq, *spare;
        uint64_t nu = 0;
        ASC_TRY(gen.range(r, &uniq, &spare, &nu));
        const uint64_t lo = gen.cuts[r];
        const uint64_t base = g->h.off[lo];
        if (base + nu != g->h.off[gen.cuts[r + 1]]) {
            set_last_error("generator: the second sweep drew a different range");
            return ASC_ERR_DEVICE;
        }
        if (!nu) continue;
        const uint64_t sb = std::max(base, S), nt = base + nu > sb ? base + nu - sb : 0;
        uint32_t* stage32 = (uint32_t*)spare;
        uint8_t* stage8 = split ? (uint8_t*)(stage32 + nt) : nullptr;
        hipLaunchKernelGGL(k_gen_emit_graph, dim3(gen.grid), dim3(kGenBlock), 0, gen.st, uniq,
                           (unsigned long long)nu, gen.vbits, (uint32_t)lo, (unsigned long long)base,
                           (unsigned long long)S, (WEdge*)g->d_edges, stage32, stage8, (unsigned long long)sb,
                           (unsigned long long)wkey, p->wmin, span, g->pk_shift, wsum.p);
        ASC_HIP(hipGetLastError());
        if (nt) {
            ASC_HIP(hipMemcpyAsync(g->ptail.p + (sb - S), stage32, 4 * nt, hipMemcpyDeviceToHost, gen.st));
            if (split)
                ASC_HIP(hipMemcpyAsync(g->ptail_w.p + (sb - S), stage8, nt, hipMemcpyDeviceToHost, gen.st));
        }
        ASC_HIP(hipStreamSynchronize(gen.st));
    }
    unsigned long long ws = 0;
    ASC_HIP(hipMemcpy(&ws, wsum.p, 8, hipMemcpyDeviceToHost));
    gen.release();
    g->mean_weight = E ? (double)ws / (double)E : 1.0;
    g->host_edges_released = true;
    g->fp = mix64(mix64(V ^ (E << 1) ^ ((uint64_t)p->type << 62)) ^ mix64(p->seed ^ ((uint64_t)p->wmax << 40) ^
                                                                           ((uint64_t)p->edge_factor << 20)));
    if (!g->fp) g->fp = 1;
    g->symmetric = -1;
    if (g->static_vertices == V) ASC_TRY(check_symmetric(g));
    g->usable = true;
    if (ctx->cfg.verbose)
        printf("graph generation: V=%llu E=%llu static %llu, tail %llu (%s) in %zu range(s): degrees %.1f ms, "
               "allocation %.1f ms, records %.1f ms\n",
               (unsigned long long)V, (unsigned long long)E, (unsigned long long)S, (unsigned long long)n_tail,
               n_tail ? (split ? "5-byte split" : "4-byte packed") : "none", gen.ranges(), t_deg - t0,
               t_alloc - t_deg, now_ms() - t_alloc);
    *out = g;
    own.g = nullptr;
    return ASC_OK;
}

}  // namespace asc

extern "C" int asc_graph_generate(asc_ctx* ctx, const asc_gen_params* p, asc_graph** out) {
    if (!ctx || !p || !out) return ASC_ERR_INVALID;
    *out = nullptr;
    return asc::graph_generate(ctx, p, out);
}

extern "C" int asc_generate_device(asc_ctx* ctx, const asc_gen_params* p, asc_host_csr* out) {
    if (!ctx || !p || !out) return ASC_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    asc::HostCSR h;
    ASC_TRY(asc::generate_on_device(ctx, p, h));
    out->num_vertices = h.V;
    out->num_edges = h.E;
    out->kind = h.kind;
    out->off = h.off.p;
    out->edges = h.edges.p;
    out->out_degree = h.outdeg.p;
    h.off.p = nullptr;
    h.edges.p = nullptr;
    h.outdeg.p = nullptr;
    return ASC_OK;
}
