// BFS that allocates nothing: asc_bfs_noalloc.
//
// asc_bfs keeps its frontier queues, visited bitmap, level counters and heavy
// list in traversal state that the library allocates on a graph's first run.
// Callers that cannot allocate device memory at call time (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc) get
// this variant instead: the same top-down frontier kernels (frontier.hpp),
// with that state carved out of `scratch`, one statically sized device array
// per GPU, and the levels written into a device buffer the caller owns.
// The graph must be fully resident and have at most kNaMaxVertices vertices.
//
// `scratch` is a host-side stand-in in host code: its device address is taken
// once per device with hipGetSymbolAddress (noalloc_init, called by
// asc_ctx_create with the device current) and only that address reaches the
// kernels.  Every call on one GPU shares that GPU's array: calls that overlap
// on one GPU race (one call at a time per device).
#include <algorithm>

#include "frontier.hpp"
#include "internal.hpp"
#include "traverse.hpp"

namespace asc {

constexpr uint64_t kNaMaxVertices = uint64_t(1) << 21;   // 2,097,152
constexpr uint32_t kNaHeavyCap = 1u << 16;               // heavy-list entries per level
constexpr int kNaMaxDevices = 64;
constexpr int kNaLevelsPerCheck = 8;                     // levels queued between counter reads

// Layout of the scratch array for kNaMaxVertices vertices (byte offsets,
// each 256-aligned): two frontier queues, the visited bitmap, the level
// counter ring, kernel statistics and the heavy list.
constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }
constexpr uint64_t kNaQ0 = 0;
constexpr uint64_t kNaQ1 = kNaQ0 + align256(4 * kNaMaxVertices);
constexpr uint64_t kNaVisited = kNaQ1 + align256(4 * kNaMaxVertices);
constexpr uint64_t kNaSlots = kNaVisited + align256(kNaMaxVertices / 8);
constexpr uint64_t kNaStats = kNaSlots + align256(4 * sizeof(Slot));
constexpr uint64_t kNaHeavyU = kNaStats + align256(sizeof(DevStats));
constexpr uint64_t kNaHeavySv = kNaHeavyU + align256(4 * (uint64_t)kNaHeavyCap);
constexpr uint64_t kNaHeavyBase = kNaHeavySv + align256(4 * (uint64_t)kNaHeavyCap);
constexpr uint64_t kScratchBytes = kNaHeavyBase + align256(8 * (uint64_t)kNaHeavyCap);   // ~17.3 MB

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device (null until noalloc_init).
static unsigned char* g_scratch_ptr[kNaMaxDevices] = {};
static std::mutex g_scratch_mu;

int noalloc_init(int device) {
    if (device < 0 || device >= kNaMaxDevices) return ASC_OK;   // entry point then reports the device unsupported
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return ASC_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_last_error("hipGetSymbolAddress of the no-allocation scratch array failed: %s", hipGetErrorString(e));
        return ASC_ERR_DEVICE;
    }
    g_scratch_ptr[device] = (unsigned char*)p;
    return ASC_OK;
}

__global__ void k_na_fill(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Level 0: the source alone (value 1, GraphMeta.cuh:432-446); counter ring
// and statistics start at zero.
__global__ void k_na_source(uint32_t* val, uint32_t* visited, uint32_t* q, Slot* slots, DevStats* stats,
                            uint32_t src) {
    if (threadIdx.x < 4) {
        slots[threadIdx.x].n_res = 0;
        slots[threadIdx.x].heavy = 0;
    }
    if (threadIdx.x == 0) {
        *stats = DevStats{};
        val[src] = 1;
        visited[src >> 5] |= 1u << (src & 31);
        q[0] = src;
        slots[0].n_res = 1;
    }
}

// ASC_VALUES_ZERO_BASED: source 0, unreached 0xFFFFFFFF.
__global__ void k_na_zero_based(uint32_t* val, uint64_t V) {
    const uint32_t unreached = (uint32_t)(V + 1);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < V; i += (uint64_t)gridDim.x * blockDim.x)
        val[i] = val[i] == unreached ? 0xFFFFFFFFu : val[i] - 1;
}

template <class OffT, class EdgeT>
static int bfs_noalloc(asc_graph* g, uint32_t src, uint32_t* d_val, hipStream_t s, unsigned char* base) {
    asc_ctx* ctx = g->ctx;
    const uint64_t V = g->h.V, E = g->h.E;
    const int grid = ctx->num_cus * kBlocksPerCU;
    const int fill_grid = ctx->num_cus * 4;
    uint32_t* visited = (uint32_t*)(base + kNaVisited);
    Slot* slots = (Slot*)(base + kNaSlots);
    TravArgs<OffT, EdgeT> a{};
    a.off = (const OffT*)g->d_off;
    a.edges = (const EdgeT*)g->d_edges;
    a.val = d_val;
    a.visited = visited;
    a.slots = slots;
    a.heavy_u = (uint32_t*)(base + kNaHeavyU);
    a.heavy_sv = (uint32_t*)(base + kNaHeavySv);
    a.heavy_base = (unsigned long long*)(base + kNaHeavyBase);
    a.stats = (DevStats*)(base + kNaStats);
    a.vs = (uint32_t)V;   // fully resident: nothing is routed to an on-demand pile
    // a level's heavy vertices each have >= thr edges, so at most E / thr <= kNaHeavyCap of them
    const uint64_t cap_thr = (E + kNaHeavyCap - 1) / kNaHeavyCap;
    a.heavy_thr = (uint32_t)std::max<uint64_t>((uint64_t)effective_heavy_threshold(g), cap_thr);
    a.theta = 0xFFFFFFFFu;

    hipLaunchKernelGGL(k_na_fill, dim3(fill_grid), dim3(kBlock), 0, s, d_val, V, (uint32_t)(V + 1));
    hipLaunchKernelGGL(k_na_fill, dim3(fill_grid), dim3(kBlock), 0, s, visited, (V + 31) / 32, 0u);
    hipLaunchKernelGGL(k_na_source, dim3(1), dim3(64), 0, s, d_val, visited, (uint32_t*)(base + kNaQ0), slots,
                       a.stats, src);
    ASC_HIP(hipGetLastError());
    // The kernels read each level's frontier size from the counter ring on
    // the device, so levels are queued kNaLevelsPerCheck at a time (a level
    // whose frontier is empty does nothing) and the host reads one counter
    // per batch to learn whether the traversal ended.
    uint32_t level = 0;
    for (;;) {
        for (int i = 0; i < kNaLevelsPerCheck; i++, level++) {
            a.q_in = (uint32_t*)(base + ((level & 1) ? kNaQ1 : kNaQ0));
            a.q_out = (uint32_t*)(base + ((level & 1) ? kNaQ0 : kNaQ1));
            hipLaunchKernelGGL((k_expand_local<ALG_BFS, OffT, EdgeT>), dim3(grid), dim3(kBlock), 0, s, a, level);
            hipLaunchKernelGGL((k_expand_heavy<ALG_BFS, OffT, EdgeT>), dim3(grid), dim3(kBlock), 0, s, a, level);
        }
        ASC_HIP(hipGetLastError());
        uint32_t next = 0;
        ASC_HIP(hipMemcpyAsync(&next, &slots[level & 3].n_res, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        ASC_HIP(hipStreamSynchronize(s));
        if (next == 0) break;
        if (level > V + kNaLevelsPerCheck) {   // a BFS has at most V levels
            set_last_error("BFS did not end after %u levels", level);
            return ASC_ERR_STATE;
        }
    }
    if (ctx->cfg.value_convention == ASC_VALUES_ZERO_BASED) {
        hipLaunchKernelGGL(k_na_zero_based, dim3(fill_grid), dim3(kBlock), 0, s, d_val, V);
        ASC_HIP(hipGetLastError());
    }
    ASC_HIP(hipStreamSynchronize(s));
    return ASC_OK;
}

}  // namespace asc

using namespace asc;

extern "C" int asc_bfs_noalloc(asc_graph* g, uint32_t source, uint32_t* d_level_out, void* stream) {
    if (!g || !d_level_out) return ASC_ERR_INVALID;
    ASC_TRY(graph_ready(g));
    const uint64_t V = g->h.V;
    if (source >= V) {
        set_last_error("source %u >= V %llu", source, (unsigned long long)V);
        return ASC_ERR_INVALID;
    }
    if (g->h.kind == ASC_KIND_CSC) return ASC_ERR_UNSUPPORTED;
    if (g->static_vertices < V) {
        set_last_error("asc_bfs_noalloc needs a fully resident graph");
        return ASC_ERR_UNSUPPORTED;
    }
    if (V > kNaMaxVertices) {
        set_last_error("asc_bfs_noalloc handles at most %llu vertices (graph has %llu)",
                       (unsigned long long)kNaMaxVertices, (unsigned long long)V);
        return ASC_ERR_RANGE;
    }
    const int dev = g->ctx->device;
    if (dev < 0 || dev >= kNaMaxDevices) return ASC_ERR_UNSUPPORTED;
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = g_scratch_ptr[dev];
    }
    if (!base) {
        set_last_error("no-allocation scratch of device %d not initialised", dev);
        return ASC_ERR_STATE;
    }
    ASC_HIP(hipSetDevice(dev));
    hipStream_t s = stream ? (hipStream_t)stream : g->ctx->s_main;
    const bool w = g->h.kind == ASC_KIND_WCSR, o4 = g->off_bytes == 4;
    if (w) return o4 ? bfs_noalloc<uint32_t, WEdge>(g, source, d_level_out, s, base)
                     : bfs_noalloc<uint64_t, WEdge>(g, source, d_level_out, s, base);
    return o4 ? bfs_noalloc<uint32_t, uint32_t>(g, source, d_level_out, s, base)
              : bfs_noalloc<uint64_t, uint32_t>(g, source, d_level_out, s, base);
}
