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

// Appends key when want: one atomic per wave.  Keys past cap are counted
// but not stored (the host then reports the mismatch).
__device__ __forceinline__ void gen_append(bool want, unsigned long long key, unsigned long long* keys,
                                           unsigned long long* count, unsigned long long cap) {
    const uint64_t m = wave_ballot(want);
    if (!m) return;
    const uint32_t leader = (uint32_t)__builtin_ctzll(m);
    unsigned long long base = 0;
    if (lane_id() == leader) base = atomicAdd(count, (unsigned long long)__popcll(m));
    base = __shfl(base, (int)leader, kWave);
    const unsigned long long pos = base + lanes_below(m);
    if (want && pos < cap) keys[pos] = key;
}

__global__ __launch_bounds__(kGenBlock) void k_gen_keys(Sampler s, uint64_t M, int sym, uint32_t lo, uint32_t hi,
                                                        uint32_t vbits, unsigned long long* keys,
                                                        unsigned long long* count, unsigned long long cap) {
    // the loop bound is uniform per workgroup, so every lane takes part in
    // the ballots
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
        gen_append(fwd, ((unsigned long long)(u - lo) << vbits) | v, keys, count, cap);
        gen_append(rev, ((unsigned long long)(v - lo) << vbits) | u, keys, count, cap);
    }
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

uint32_t bits_for(uint64_t n) { return n > 1 ? 64 - (uint32_t)__builtin_clzll(n - 1) : 1; }

}  // namespace

int generate_on_device(asc_ctx* ctx, const asc_gen_params* p, HostCSR& out) {
    Sampler s;
    uint64_t M = 0;
    ASC_TRY(make_sampler(p, s, M));
    const bool sym = p->symmetrize != 0;
    if (p->csc && !sym) {
        set_last_error("device generation builds the in-CSR of symmetrized graphs only");
        return ASC_ERR_UNSUPPORTED;
    }
    ASC_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->s_main;
    const uint64_t V = s.V;
    const int grid = ctx->num_cus * 8;
    const double t0 = now_ms();

    // 1. raw out-degrees
    HostBuf<uint32_t> cnt_h;
    ASC_TRY(cnt_h.alloc(V));
    {
        DevTmp<uint32_t> cnt;
        if (hipMalloc(&cnt.p, 4 * V) != hipSuccess) {
            set_last_error("device allocation of %llu count bytes failed", (unsigned long long)(4 * V));
            return ASC_ERR_NOMEM;
        }
        ASC_HIP(hipMemsetAsync(cnt.p, 0, 4 * V, st));
        hipLaunchKernelGGL(k_gen_count, dim3(grid), dim3(kGenBlock), 0, st, s, M, (int)sym, cnt.p);
        ASC_HIP(hipGetLastError());
        ASC_HIP(hipMemcpyAsync(cnt_h.p, cnt.p, 4 * V, hipMemcpyDeviceToHost, st));
        ASC_HIP(hipStreamSynchronize(st));
    }
    uint64_t raw_total = 0;
    for (uint64_t v = 0; v < V; v++) raw_total += cnt_h[v];
    const double t_count = now_ms();

    // 2. ranges whose keys (two 8-byte buffers) fit three quarters of the
    // free memory, after the per-vertex bound arrays of a whole-graph range
    size_t free_b = 0, total_b = 0;
    ASC_HIP(hipMemGetInfo(&free_b, &total_b));
    if (ctx->cfg.device_mem_limit) free_b = std::min<size_t>(free_b, ctx->cfg.device_mem_limit);
    const uint64_t per_vertex = 20 * V + (64u << 20);
    const uint64_t usable = free_b > per_vertex ? (free_b - per_vertex) / 4 * 3 : 0;
    const uint64_t cap_keys = std::max<uint64_t>(usable / 16, 1);
    std::vector<uint64_t> cuts{0};
    {
        uint64_t acc = 0;
        for (uint64_t v = 0; v < V; v++) {
            if (cnt_h[v] > cap_keys) {
                set_last_error("vertex %llu has %u raw edges, more than fit the device at once",
                               (unsigned long long)v, cnt_h[v]);
                return ASC_ERR_NOMEM;
            }
            if (acc + cnt_h[v] > cap_keys) {
                cuts.push_back(v);
                acc = 0;
            }
            acc += cnt_h[v];
        }
        cuts.push_back(V);
    }
    uint64_t max_raw = 0, max_vertices = 0;
    for (size_t r = 0; r + 1 < cuts.size(); r++) {
        uint64_t n = 0;
        for (uint64_t v = cuts[r]; v < cuts[r + 1]; v++) n += cnt_h[v];
        max_raw = std::max(max_raw, n);
        max_vertices = std::max(max_vertices, cuts[r + 1] - cuts[r]);
    }

    const int words = p->csc ? 1 : (p->weighted ? 2 : 1);
    out.V = V;
    out.kind = p->csc ? ASC_KIND_CSC : (p->weighted ? ASC_KIND_WCSR : ASC_KIND_CSR);
    ASC_TRY(out.off.alloc(V + 1));
    // raw_total bounds the deduplicated edges; pages past E are never touched
    ASC_TRY(out.edges.alloc((size_t)raw_total * words + 1));

    DevTmp<unsigned long long> ka, kb, first, last, dcount;
    DevTmp<uint32_t> deg;
    DevTmp<unsigned char> tmp;
    const uint64_t kbytes = 8 * std::max<uint64_t>(max_raw, 1);
    if (hipMalloc(&ka.p, kbytes) != hipSuccess || hipMalloc(&kb.p, kbytes) != hipSuccess ||
        hipMalloc(&first.p, 8 * max_vertices) != hipSuccess || hipMalloc(&last.p, 8 * max_vertices) != hipSuccess ||
        hipMalloc(&deg.p, 4 * max_vertices) != hipSuccess || hipMalloc(&dcount.p, 16) != hipSuccess) {
        set_last_error("device allocation of the generator's %llu key bytes failed",
                       (unsigned long long)(2 * kbytes));
        return ASC_ERR_NOMEM;
    }
    const uint32_t vbits = bits_for(V);
    size_t sort_bytes = 0, uniq_bytes = 0;
    {
        rocprim::double_buffer<unsigned long long> db(ka.p, kb.p);
        ASC_HIP(rocprim::radix_sort_keys(nullptr, sort_bytes, db, (size_t)max_raw, 0, 64, st));
        ASC_HIP(rocprim::unique(nullptr, uniq_bytes, ka.p, kb.p, dcount.p + 1, (size_t)max_raw,
                                rocprim::equal_to<unsigned long long>(), st));
    }
    const size_t tmp_bytes = std::max<size_t>(std::max(sort_bytes, uniq_bytes), 16);
    if (hipMalloc(&tmp.p, tmp_bytes) != hipSuccess) {
        set_last_error("device allocation of %zu sort bytes failed", tmp_bytes);
        return ASC_ERR_NOMEM;
    }

    std::vector<uint32_t> degh(max_vertices);
    const uint32_t span = p->weighted ? p->wmax - p->wmin + 1 : 1;
    const uint64_t wkey = weight_key(p->seed);
    uint64_t E = 0;
    out.off[0] = 0;
    for (size_t r = 0; r + 1 < cuts.size(); r++) {
        const uint64_t lo = cuts[r], hi = cuts[r + 1], nv = hi - lo;
        uint64_t nraw = 0;
        for (uint64_t v = lo; v < hi; v++) nraw += cnt_h[v];
        uint64_t nu = 0;
        if (nraw) {
            ASC_HIP(hipMemsetAsync(dcount.p, 0, 16, st));
            hipLaunchKernelGGL(k_gen_keys, dim3(grid), dim3(kGenBlock), 0, st, s, M, (int)sym, (uint32_t)lo,
                               (uint32_t)hi, vbits, ka.p, dcount.p, (unsigned long long)nraw);
            ASC_HIP(hipGetLastError());
            unsigned long long got = 0;
            ASC_HIP(hipMemcpyAsync(&got, dcount.p, 8, hipMemcpyDeviceToHost, st));
            ASC_HIP(hipStreamSynchronize(st));
            if (got != nraw) {
                set_last_error("generator: %llu keys drawn for a range counted at %llu", got,
                               (unsigned long long)nraw);
                return ASC_ERR_DEVICE;
            }
            rocprim::double_buffer<unsigned long long> db(ka.p, kb.p);
            const uint32_t kbits = bits_for(nv) + vbits;
            size_t b = tmp_bytes;
            ASC_HIP(rocprim::radix_sort_keys(tmp.p, b, db, (size_t)nraw, 0, kbits, st));
            unsigned long long* sorted = db.current();
            unsigned long long* uniq = sorted == ka.p ? kb.p : ka.p;
            b = tmp_bytes;
            ASC_HIP(rocprim::unique(tmp.p, b, sorted, uniq, dcount.p + 1, (size_t)nraw,
                                    rocprim::equal_to<unsigned long long>(), st));
            unsigned long long nuniq = 0;
            ASC_HIP(hipMemcpyAsync(&nuniq, dcount.p + 1, 8, hipMemcpyDeviceToHost, st));
            ASC_HIP(hipMemsetAsync(first.p, 0, 8 * nv, st));
            ASC_HIP(hipMemsetAsync(last.p, 0, 8 * nv, st));
            ASC_HIP(hipStreamSynchronize(st));
            nu = nuniq;
            hipLaunchKernelGGL(k_gen_bounds, dim3(grid), dim3(kGenBlock), 0, st, uniq, (unsigned long long)nu, vbits,
                               first.p, last.p);
            hipLaunchKernelGGL(k_gen_degrees, dim3(grid), dim3(kGenBlock), 0, st, first.p, last.p, (uint32_t)nv,
                               deg.p);
            // records over the sorted keys' buffer, which unique has consumed
            uint32_t* rec = (uint32_t*)sorted;
            hipLaunchKernelGGL(k_gen_emit, dim3(grid), dim3(kGenBlock), 0, st, uniq, (unsigned long long)nu, vbits,
                               (uint32_t)lo, (int)(words == 2), (unsigned long long)wkey, p->wmin, span, rec);
            ASC_HIP(hipGetLastError());
            ASC_HIP(hipMemcpyAsync(degh.data(), deg.p, 4 * nv, hipMemcpyDeviceToHost, st));
            ASC_HIP(hipMemcpyAsync(out.edges.p + E * words, rec, 4 * words * nu, hipMemcpyDeviceToHost, st));
            ASC_HIP(hipStreamSynchronize(st));
        } else {
            std::fill(degh.begin(), degh.begin() + nv, 0u);
        }
        uint64_t run = E;
        for (uint64_t i = 0; i < nv; i++) {
            run += degh[i];
            out.off[lo + i + 1] = run;
        }
        if (run != E + nu) {
            set_last_error("generator: degrees of a range sum to %llu, not %llu", (unsigned long long)(run - E),
                           (unsigned long long)nu);
            return ASC_ERR_DEVICE;
        }
        E = run;
    }
    out.E = E;
    if (p->csc) {
        ASC_TRY(out.outdeg.alloc(V));
        for (uint64_t v = 0; v < V; v++) out.outdeg[v] = (uint32_t)(out.off[v + 1] - out.off[v]);
    }
    if (ctx->cfg.verbose)
        printf("device generation: V=%llu E=%llu (raw %llu) in %zu range(s): count %.1f ms, total %.1f ms\n",
               (unsigned long long)V, (unsigned long long)E, (unsigned long long)raw_total, cuts.size() - 1,
               t_count - t0, now_ms() - t0);
    return ASC_OK;
}

}  // namespace asc

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
