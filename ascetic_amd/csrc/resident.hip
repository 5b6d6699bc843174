// Resident fast paths for graphs whose edges all sit on the device and whose
// adjacency is symmetric (the symmetrized R-MAT and Erdos-Renyi inputs):
//
//   direction-optimizing BFS  top-down steps (the frontier kernels of
//       traverse.hip) switch to bottom-up steps over frontier bitmaps when the
//       frontier's edges exceed the unexplored edges / kAlpha, and back once
//       the frontier shrinks below V / kBeta (Beamer, Asanovic and Patterson,
//       SC'12).  A bottom-up step lets every unvisited vertex scan its own
//       adjacency for a frontier member and stop at the first one, so the wide
//       middle levels read a fraction of their edges.  BFS levels are unique,
//       so the result is the reference's top-down bfs_opt result
//       (CalculateOpt.cu:14-208).
//   union-find CC  lock-free hooking of the larger root under the smaller one
//       (comp[x] <= x always holds, so the root of a tree is its smallest
//       vertex): first over two sampled neighbours of every vertex, then over
//       the remaining edges of vertices outside the most common component,
//       then full path compression (the Afforest scheme of Sutton, Ben-Nun and
//       Barak, IPDPS'18).  On a symmetric graph the root is the smallest id of
//       the component, which is where cc_opt's min-label propagation
//       (CalculateOpt.cu:209-382) converges.
// Symmetry is decided once per upload from order-independent 64-bit hash
// sums of (u, v) and of (v, u) over all edges, under two seeds.
#include <algorithm>
#include <unordered_map>

#include "device_common.hpp"
#include "internal.hpp"
#include "traverse.hpp"
#include "frontier.hpp"

namespace asc {

constexpr uint64_t kAlpha = 15;   // top-down -> bottom-up when m_frontier > m_unexplored / kAlpha
constexpr uint64_t kBeta = 18;    // bottom-up -> top-down when n_frontier < V / kBeta (and shrinking)
constexpr int kSamples = 1024;    // vertices sampled for the most common component
constexpr int kSampleRounds = 2;  // neighbours linked per vertex before sampling

// ---------------------------------------------------------------------------
// symmetry check
// ---------------------------------------------------------------------------

// SplitMix64 output function (Steele, Lea and Flood, OOPSLA'14).
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

constexpr uint64_t kSeed0 = 0x9e3779b97f4a7c15ull, kSeed1 = 0xd1b54a32d192ed03ull;

// acc[0..1] += h_s(u, v), acc[2..3] += h_s(v, u) over every edge u -> v.  Each
// workgroup takes tiles of kTile edges and finds every edge's owner u by a
// binary search of the offsets between the tile's first and last owners.
template <class OffT, class EdgeT>
__global__ __launch_bounds__(kBlock) void k_sym_hash(const OffT* off, const EdgeT* edges, uint64_t V, uint64_t E,
                                                     unsigned long long* acc) {
    __shared__ uint32_t s_u[2];
    __shared__ unsigned long long s_red[4][kWavesPerBlock];
    unsigned long long f0 = 0, f1 = 0, r0 = 0, r1 = 0;
    const uint32_t tid = threadIdx.x;
    for (uint64_t t0 = (uint64_t)blockIdx.x * kTile; t0 < E; t0 += (uint64_t)gridDim.x * kTile) {
        const uint64_t t1 = min(t0 + (uint64_t)kTile, E);
        if (tid < 2) {
            // largest u with off[u] <= key (off[V] = E > key)
            const uint64_t key = tid == 0 ? t0 : t1 - 1;
            uint64_t lo = 0, hi = V - 1;
            while (lo < hi) {
                const uint64_t mid = (lo + hi + 1) >> 1;
                if ((uint64_t)off[mid] <= key) lo = mid; else hi = mid - 1;
            }
            s_u[tid] = (uint32_t)lo;
        }
        __syncthreads();
        const uint32_t ua = s_u[0], ub = s_u[1];
        for (uint64_t e = t0 + tid; e < t1; e += kBlock) {
            uint32_t lo = ua, hi = ub;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if ((uint64_t)off[mid] <= e) lo = mid; else hi = mid - 1;
            }
            const uint64_t u = lo, v = edge_dst(edges[e]);
            const uint64_t fw = (u << 32) | v, rv = (v << 32) | u;
            f0 += mix64(fw + kSeed0);
            f1 += mix64(fw + kSeed1);
            r0 += mix64(rv + kSeed0);
            r1 += mix64(rv + kSeed1);
        }
        __syncthreads();
    }
    f0 = wave_sum(f0);
    f1 = wave_sum(f1);
    r0 = wave_sum(r0);
    r1 = wave_sum(r1);
    if (lane_id() == 0) {
        s_red[0][wave_id()] = f0;
        s_red[1][wave_id()] = f1;
        s_red[2][wave_id()] = r0;
        s_red[3][wave_id()] = r1;
    }
    __syncthreads();
    if (tid < 4) {
        unsigned long long x = 0;
        for (int i = 0; i < kWavesPerBlock; i++) x += s_red[tid][i];
        atomicAdd(acc + tid, x);
    }
}

template <class OffT, class EdgeT>
static int sym_hash(asc_graph* g, unsigned long long* d_acc) {
    const uint64_t V = g->h.V, E = g->h.E;
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((E + kTile - 1) / kTile,
                                                                   (uint64_t)g->ctx->num_cus * 8));
    hipLaunchKernelGGL((k_sym_hash<OffT, EdgeT>), dim3(grid), dim3(kBlock), 0, g->ctx->s_main,
                       (const OffT*)g->d_off, (const EdgeT*)g->d_edges, V, E, d_acc);
    ASC_HIP(hipGetLastError());
    return ASC_OK;
}

int check_symmetric(asc_graph* g) {
    g->symmetric = -1;
    const uint64_t V = g->h.V, E = g->h.E;
    if (V == 0 || V > 0xFFFFFFFFull || g->static_edges != E) return ASC_OK;
    if (E == 0) {
        g->symmetric = 1;
        return ASC_OK;
    }
    hipStream_t s = g->ctx->s_main;
    unsigned long long* d_acc = nullptr;
    ASC_HIP(hipMalloc(&d_acc, 4 * sizeof(unsigned long long)));
    int r = hipMemsetAsync(d_acc, 0, 4 * sizeof(unsigned long long), s) == hipSuccess ? ASC_OK : ASC_ERR_DEVICE;
    if (r == ASC_OK) {
        const bool w = g->h.kind == ASC_KIND_WCSR, o4 = g->off_bytes == 4;
        r = w ? (o4 ? sym_hash<uint32_t, WEdge>(g, d_acc) : sym_hash<uint64_t, WEdge>(g, d_acc))
              : (o4 ? sym_hash<uint32_t, uint32_t>(g, d_acc) : sym_hash<uint64_t, uint32_t>(g, d_acc));
    }
    unsigned long long h[4] = {0, 1, 2, 3};
    if (r == ASC_OK && hipMemcpyAsync(h, d_acc, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess) r = ASC_ERR_DEVICE;
    if (r == ASC_OK && hipStreamSynchronize(s) != hipSuccess) r = ASC_ERR_DEVICE;
    (void)hipFree(d_acc);
    if (r != ASC_OK) {
        set_last_error("symmetry check failed on the device");
        return r;
    }
    g->symmetric = (h[0] == h[2] && h[1] == h[3]) ? 1 : 0;
    return ASC_OK;
}

bool resident_symmetric(const asc_graph* g) {
    return g->symmetric == 1 && g->static_vertices == g->h.V && g->static_edges == g->h.E &&
           g->h.kind != ASC_KIND_CSC && g->h.V > 0;
}

// ---------------------------------------------------------------------------
// direction-optimizing BFS
// ---------------------------------------------------------------------------

// Block-wide sum of a per-thread counter, added to *dst by thread 0.
__device__ __forceinline__ void block_add(unsigned long long x, unsigned long long* s_tmp, unsigned long long* dst) {
    x = wave_sum(x);
    if (lane_id() == 0) s_tmp[wave_id()] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < kWavesPerBlock; i++) t += s_tmp[i];
        if (t) atomicAdd(dst, t);
    }
    __syncthreads();
}

// One bottom-up level: every unvisited vertex v scans its adjacency for a
// member of `front` (the level-`level` frontier bitmap) and, at the first,
// takes level + 2 (values are level + 1, unreached V + 1).  Each wave owns 64
// consecutive vertices, i.e. two bitmap words, and writes them whole: `next`
// needs no clearing and `visited` no atomics (no other wave writes them).
// A lane scans the first kBuLane edges of its own vertex (most vertices find
// a parent among them); the rest of any longer adjacency is then scanned by
// the whole wave, 128 edges per pass with coalesced loads, so a wave no longer
// waits for one lane walking a long list edge by edge.
constexpr int kBuLane = 8;

// Body of one bottom-up level for workgroup `bid` of `nblk`.  With q_out set,
// the vertices found are also appended to that queue (count in *q_count, one
// atomic per wave), so a top-down level can follow without a conversion.
template <class OffT, class EdgeT>
__device__ __forceinline__ void bottom_up_body(const OffT* off, const EdgeT* edges, uint32_t* val, uint32_t* visited,
                                               const uint32_t* front, uint32_t* next, uint32_t V, uint32_t level,
                                               DoCounters* c, DevStats* stats, uint32_t* q_out, uint32_t* q_count,
                                               uint32_t bid, uint32_t nblk) {
    __shared__ unsigned long long s_tmp[kWavesPerBlock];
    const uint32_t lane = lane_id();
    const uint32_t nwaves = nblk * kWavesPerBlock;
    unsigned long long n_found = 0, m_found = 0, examined = 0, scanned = 0;
    for (uint64_t wv = (uint64_t)bid * kWavesPerBlock + wave_id(); wv * kWave < V; wv += nwaves) {
        const uint32_t base = (uint32_t)(wv * kWave);
        const uint32_t v = base + lane;
        bool found = false;
        uint64_t deg = 0;
        OffT b = 0, e = 0;
        if (v < V && !((visited[v >> 5] >> (v & 31)) & 1u)) {
            b = off[v];
            e = off[v + 1];
            deg = (uint64_t)(e - b);
            scanned++;
            // the first kBuLane edges: four loads in flight per lane before
            // the first bitmap probe
            const OffT el = deg > (uint64_t)kBuLane ? b + (OffT)kBuLane : e;
            for (OffT i = b; i < el && !found; i += 4) {
                uint32_t u[4];
#pragma unroll
                for (int j = 0; j < 4; j++) u[j] = (i + (OffT)j < el) ? edge_dst(edges[i + j]) : 0xFFFFFFFFu;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (!found && u[j] != 0xFFFFFFFFu) {
                        examined++;
                        if ((front[u[j] >> 5] >> (u[j] & 31)) & 1u) found = true;
                    }
                }
            }
        }
        // longer adjacencies not yet settled: one vertex at a time, whole wave
        uint64_t todo = wave_ballot(!found && deg > (uint64_t)kBuLane);
        while (todo) {
            const int l = __builtin_ctzll(todo);
            todo &= todo - 1;
            const OffT lb = __shfl(b, l, kWave) + (OffT)kBuLane, le = __shfl(e, l, kWave);
            bool hit = false;
            for (OffT i0 = lb; i0 < le; i0 += 2 * kWave) {
                const OffT i1 = i0 + (OffT)lane, i2 = i1 + (OffT)kWave;
                const uint32_t u1 = i1 < le ? edge_dst(edges[i1]) : 0xFFFFFFFFu;
                const uint32_t u2 = i2 < le ? edge_dst(edges[i2]) : 0xFFFFFFFFu;
                const bool h1 = u1 != 0xFFFFFFFFu && ((front[u1 >> 5] >> (u1 & 31)) & 1u);
                const bool h2 = u2 != 0xFFFFFFFFu && ((front[u2 >> 5] >> (u2 & 31)) & 1u);
                examined += (u1 != 0xFFFFFFFFu) + (u2 != 0xFFFFFFFFu);
                if (wave_ballot(h1 || h2)) {
                    hit = true;
                    break;
                }
            }
            if ((int)lane == l) found = hit;
        }
        const uint64_t m = wave_ballot(found);
        if (found) {
            val[v] = level + 2;
            n_found++;
            m_found += deg;
        }
        if ((lane & 31) == 0) {
            const uint32_t wi = (base >> 5) + (lane >> 5);
            if ((uint64_t)wi * 32 < V) {
                const uint32_t word = (uint32_t)(m >> lane);
                next[wi] = word;
                if (word) visited[wi] |= word;
            }
        }
        if (q_out && m) {
            uint32_t at = 0;
            if (lane == 0) at = atomicAdd(q_count, (uint32_t)__popcll(m));
            at = __shfl(at, 0, kWave);
            if (found) q_out[at + lanes_below(m)] = v;
        }
    }
    block_add(n_found, s_tmp, &c->n_found);
    block_add(m_found, s_tmp, &c->m_found);
    block_add(examined, s_tmp, &stats->edges_bu);
    block_add(scanned, s_tmp, &stats->verts_bu);
}

template <class OffT, class EdgeT>
__global__ __launch_bounds__(kBlock) void k_bfs_bottom_up(const OffT* off, const EdgeT* edges, uint32_t* val,
                                                          uint32_t* visited, const uint32_t* front, uint32_t* next,
                                                          uint32_t V, uint32_t level, DoCounters* c,
                                                          DevStats* stats) {
    bottom_up_body(off, edges, val, visited, front, next, V, level, c, stats, (uint32_t*)nullptr,
                   (uint32_t*)nullptr, blockIdx.x, gridDim.x);
}

// Frontier queue (count in slot->n_res) -> bitmap (cleared beforehand).
__global__ __launch_bounds__(kBlock) void k_queue_to_bitmap(const uint32_t* q, const Slot* slot, uint32_t* bm) {
    const uint32_t n = slot->n_res;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint32_t v = q[i];
        atomicOr(bm + (v >> 5), 1u << (v & 31));
    }
}

// Bitmap -> frontier queue; *count (zero beforehand) gets the entries.  A
// thread owns one 32-vertex word; a workgroup scans its 256 words' popcounts
// and reserves their queue space with a single atomic.
__global__ __launch_bounds__(kBlock) void k_bitmap_to_queue(const uint32_t* bm, uint32_t V, uint32_t* q,
                                                            uint32_t* count) {
    __shared__ uint32_t s_scan[kWavesPerBlock];
    __shared__ uint32_t s_base;
    const uint32_t words = (V + 31) / 32;
    for (uint32_t w0 = blockIdx.x * kBlock; w0 < words; w0 += gridDim.x * kBlock) {
        const uint32_t wi = w0 + threadIdx.x;
        uint32_t word = 0;
        if (wi < words) {
            word = bm[wi];
            const uint32_t valid = V - wi * 32;
            if (valid < 32) word &= (1u << valid) - 1u;
        }
        uint32_t total;
        const uint32_t pre = block_scan_excl((uint32_t)__popc(word), s_scan, total);
        if (threadIdx.x == 0) s_base = total ? atomicAdd(count, total) : 0;
        __syncthreads();
        uint32_t at = s_base + pre;
        while (word) {
            const uint32_t b = __ffs(word) - 1;
            q[at++] = wi * 32 + b;
            word &= word - 1;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_do_fill(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) p[i] = v;
}

// End of a level: the level's counters (and the next top-down frontier's
// size, in `pad`) are written to mapped host memory, and the device counters
// are cleared for the next level (no copy or memset dispatch per level).
__global__ void k_do_publish(DoCounters* c, const Slot* next_slot, DoCounters* host, unsigned long long seq) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const DoCounters v = *c;
    host->n_found = v.n_found;
    host->m_found = v.m_found;
    host->m_frontier = v.m_frontier;
    host->pad = next_slot ? next_slot->n_res : 0;
    __threadfence_system();
    host[1].n_found = seq;   // the counters above are complete (the host polls this word)
    c->n_found = 0;
    c->m_found = 0;
    c->m_frontier = 0;
    c->pad = 0;
}

// Level-0 frontier: the source alone, value 1 (level + 1).
// ---- persistent direction-optimizing BFS -----------------------------------
// The whole traversal in one cooperative launch: every level runs between
// grid barriers, and every workgroup takes the same direction decision from
// counters that are final at the barrier, so the host only launches the
// kernel and waits for it once.  A level is either
//   top-down: expand_local_body over the frontier queue (heavy vertices to
//     the heavy list), a barrier, and, when the list is not empty,
//     expand_heavy_body and a second barrier;
//   bottom-up: bottom_up_body over the frontier bitmap, which also appends
//     the vertices it finds to the next queue, and a barrier.
// Switching to bottom-up converts the current queue into its bitmap (one
// more barrier); switching back needs nothing, the queue is already there.
// Three bitmaps rotate: level L reads bm[L % 3], writes bm[(L + 1) % 3] and
// clears bm[(L + 2) % 3].  The decision is the host-stepped one: both
// directions add the degrees of the vertices they reach to a ring of
// per-level counters, so at the start of a level every workgroup knows the
// frontier's vertices and edges and the edges not yet explored.
//
// Barrier: workgroups arrive in 16 groups (one counter each, on its own
// 128-byte line); the last of a group arrives at the top counter, and the
// last there advances the generation the others poll.  Arrival publishes
// the workgroup's writes with one agent-scope release fence, polls are
// relaxed loads, and one acquire fence follows the wait: an acquire load per
// poll would invalidate the L2 each time (measured: 57 us per barrier over
// 1024 workgroups, 16 us with relaxed polls).  Waits are bounded: a
// workgroup that waits past kBarSpin polls sets the error word, every
// workgroup then leaves at its next barrier, and the host reports a device
// error.
constexpr uint32_t kBarSpin = 1u << 22;
constexpr int kBarLine = 32;                 // u32 words per 128-byte line
constexpr int kBarGroups = 16;
// control block, in 128-byte lines: [0] generation, [1] top count, [2] error,
// [3] levels run, one line per arrival group, then the ring of four
// per-level DoCounters
constexpr int kRingWord = (4 + kBarGroups) * kBarLine;
constexpr int kPctlWords = kRingWord + kBarLine;
constexpr int kPersistPerCU = 2;             // workgroups per CU of the persistent grid

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_agent64(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool grid_barrier(uint32_t* ctl, uint32_t nblk) {
    __shared__ uint32_t s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ok = 1;
        uint32_t* gen_p = ctl;
        uint32_t* top_p = ctl + kBarLine;
        uint32_t* err_p = ctl + 2 * kBarLine;
        const uint32_t gen = ld_agent(gen_p);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t g = blockIdx.x % kBarGroups;
        const uint32_t gsize = nblk / kBarGroups + (g < nblk % kBarGroups ? 1u : 0u);
        const uint32_t ngroups = nblk < (uint32_t)kBarGroups ? nblk : (uint32_t)kBarGroups;
        uint32_t* grp = ctl + (4 + g) * kBarLine;
        bool last = false;
        if (__hip_atomic_fetch_add(grp, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == gsize) {
            __hip_atomic_store(grp, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__hip_atomic_fetch_add(top_p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == ngroups) {
                __hip_atomic_store(top_p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                last = true;
            }
        }
        if (last) {
            __hip_atomic_fetch_add(gen_p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0;
            while (ld_agent(gen_p) == gen) {
                if (ld_agent(err_p) || ++spins > kBarSpin) {
                    __hip_atomic_store(err_p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (ld_agent(err_p)) ok = 0;
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// mring[L % 4]: summed degrees of level L's frontier (mring[0] = the
// source's degree, set by the host).
template <class OffT, class EdgeT>
__global__ __launch_bounds__(kBlock) void k_bfs_persistent(TravArgs<OffT, EdgeT> a, uint32_t* q0, uint32_t* q1,
                                                           uint32_t* bm0, uint32_t* bm1, uint32_t* bm2,
                                                           DoCounters* cring, uint32_t* ctl, uint32_t V,
                                                           unsigned long long E, uint32_t max_levels) {
    const uint32_t bid = blockIdx.x, nblk = gridDim.x;
    uint32_t* bm[3] = {bm0, bm1, bm2};
    uint32_t* q[2] = {q0, q1};
    const uint32_t words = (V + 31) / 32;
    bool bu = false;
    uint32_t prev_nf = 0, level = 0;
    unsigned long long m_unexplored = E;
    for (; level < max_levels; level++) {
        Slot* cur = a.slots + (level & 3);
        const uint32_t nf = ld_agent(&cur->n_res);
        if (nf == 0) break;
        const unsigned long long mf = ld_agent64(&cring[level & 3].m_found);
        m_unexplored = m_unexplored > mf ? m_unexplored - mf : 0;
        const bool to_bu = !bu && level > 0 && mf > m_unexplored / kAlpha;
        if (bu && nf < V / (uint32_t)kBeta && nf < prev_nf) bu = false;
        prev_nf = nf;
        DoCounters* nxt_c = cring + ((level + 1) & 3);
        if (bid == 0 && threadIdx.x == 0) {
            DoCounters* z = cring + ((level + 2) & 3);
            z->n_found = 0;
            z->m_found = 0;
        }
        // the bitmap level + 1 writes is cleared now (a bottom-up level
        // writes whole words, a top-down one sets single bits of it)
        uint32_t* bz = bm[(level + 2) % 3];
        for (uint32_t i = bid * kBlock + threadIdx.x; i < words; i += nblk * kBlock) bz[i] = 0;
        if (to_bu) {
            // frontier queue -> bitmap bm[level % 3] (cleared at level - 2)
            for (uint32_t i = bid * kBlock + threadIdx.x; i < nf; i += nblk * kBlock) {
                const uint32_t v = q[level & 1][i];
                atomicOr(bm[level % 3] + (v >> 5), 1u << (v & 31));
            }
            bu = true;
            if (!grid_barrier(ctl, nblk)) return;
        }
        if (bu) {
            if (bid == 0 && threadIdx.x == 0) {
                Slot* z = a.slots + ((level + 2) & 3);
                z->n_res = 0;
                z->heavy = 0;
            }
            bottom_up_body(a.off, a.edges, a.val, a.visited, bm[level % 3], bm[(level + 1) % 3], V, level, nxt_c,
                           a.stats, q[(level + 1) & 1], &a.slots[(level + 1) & 3].n_res, bid, nblk);
            if (!grid_barrier(ctl, nblk)) return;
        } else {
            TravArgs<OffT, EdgeT> b = a;
            b.q_in = q[level & 1];
            b.q_out = q[(level + 1) & 1];
            b.m_next = &nxt_c->m_found;
            expand_local_body<ALG_BFS>(b, level, bid, nblk);
            if (!grid_barrier(ctl, nblk)) return;
            const unsigned long long h = ld_agent64(&cur->heavy);
            if (h >> kHeavyShift) {
                expand_heavy_body<ALG_BFS>(b, level, bid, nblk);
                if (!grid_barrier(ctl, nblk)) return;
            }
            // a top-down level's output becomes a bitmap only if bottom-up follows
        }
    }
    if (bid == 0 && threadIdx.x == 0) ctl[3 * kBarLine] = level;
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

__global__ void k_do_source(uint32_t* val, uint32_t* visited, uint32_t* q, Slot* slot, uint32_t src) {
    val[src] = 1;
    visited[src >> 5] |= 1u << (src & 31);
    q[0] = src;
    slot->n_res = 1;
}

static int ensure_fast_state(asc_graph* g) {
    ASC_TRY(trav_ensure(g));
    TravState* t = g->trav;
    const uint64_t words = (g->h.V + 31) / 32 + 2;
    if (!t->bm[0]) {
        ASC_HIP(hipMalloc(&t->bm[0], 4 * words));
        ASC_HIP(hipMalloc(&t->bm[1], 4 * words));
        ASC_HIP(hipMalloc(&t->bm[2], 4 * words));
        ASC_HIP(hipMalloc(&t->pctl, 4 * kPctlWords));
    }
    if (!t->dcnt) {
        ASC_HIP(hipMalloc(&t->dcnt, sizeof(DoCounters)));
        ASC_TRY(t->h_dcnt.alloc(2, true));   // [0] counters, [1].n_found the publish sequence
        t->h_dcnt[1].n_found = 0;
        t->pub_seq = 0;
        ASC_HIP(hipHostGetDevicePointer((void**)&t->d_hdcnt, t->h_dcnt.p, 0));
        ASC_TRY(t->h_sample.alloc(kSamples, true));
    }
    return ASC_OK;
}

template <class OffT, class EdgeT>
static int run_bfs_do(asc_graph* g, uint32_t src, uint32_t* out, asc_stats* st) {
    asc_ctx* ctx = g->ctx;
    ASC_HIP(hipSetDevice(ctx->device));
    ASC_TRY(ensure_fast_state(g));
    TravState* t = g->trav;
    const uint64_t V = g->h.V, E = g->h.E;
    const uint64_t words = (V + 31) / 32;
    hipStream_t s = ctx->s_main;
    const bool prof = ctx->cfg.profile != 0;
    const int fill_grid = ctx->num_cus * 4;
    const int grid = ctx->num_cus * (g->blocks_per_cu > 0 ? g->blocks_per_cu : kBlocksPerCU);
    const int bu_grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((V + kBlock - 1) / kBlock, (uint64_t)grid));

    hipLaunchKernelGGL(k_do_fill, dim3(1), dim3(kBlock), 0, s, (uint32_t*)t->slots, sizeof(Slot) * 4 / 4, 0u);
    hipLaunchKernelGGL(k_do_fill, dim3(1), dim3(kBlock), 0, s, (uint32_t*)t->stats, sizeof(DevStats) / 4, 0u);
    hipLaunchKernelGGL(k_do_fill, dim3(1), dim3(kBlock), 0, s, (uint32_t*)t->dcnt, sizeof(DoCounters) / 4, 0u);
    hipLaunchKernelGGL(k_do_fill, dim3(fill_grid), dim3(kBlock), 0, s, t->val, V, (uint32_t)(V + 1));
    hipLaunchKernelGGL(k_do_fill, dim3(fill_grid), dim3(kBlock), 0, s, t->visited, words, 0u);
    hipLaunchKernelGGL(k_do_source, dim3(1), dim3(1), 0, s, t->val, t->visited, t->q[0], t->slots, src);
    ASC_HIP(hipGetLastError());
    ASC_HIP(hipStreamSynchronize(s));

    TravArgs<OffT, EdgeT> a{};
    a.off = (const OffT*)g->d_off;
    a.edges = (const EdgeT*)g->d_edges;
    a.val = t->val;
    a.visited = t->visited;
    a.slots = t->slots;
    a.heavy_u = t->heavy_u;
    a.heavy_sv = t->heavy_sv;
    a.heavy_base = t->heavy_base;
    a.stats = t->stats;
    a.vs = (uint32_t)V;   // every vertex is resident: nothing goes to a pile
    a.heavy_thr = (uint32_t)effective_heavy_threshold(g);
    a.theta = 0xFFFFFFFFu;
    // top-down levels add the degrees of the vertices they claim here: the
    // next frontier's edges, for the direction rule
    a.m_next = &t->dcnt->m_frontier;

    RunCounters rc{};
    t->timed.clear();
    t->ev_next = 0;
    const uint64_t dsrc = g->h.off[src + 1] - g->h.off[src];
    uint64_t nf = 1, mf = dsrc, prev_nf = 0;
    uint64_t m_unexplored = E - dsrc;
    int cur = 0;          // bitmap holding the frontier while bottom-up
    bool bottom_up = false;
    uint32_t level = 0;
    const double t0 = now_ms();
    while (nf > 0) {
        hipEvent_t e0 = nullptr;
        if (!prof) {
            e0 = next_event(t);
            if (!e0) return ASC_ERR_DEVICE;
            ASC_HIP(hipEventRecord(e0, s));
        }
        if (!bottom_up && level > 0 && mf > m_unexplored / kAlpha) {
            // queue -> bitmap
            hipLaunchKernelGGL(k_do_fill, dim3(fill_grid), dim3(kBlock), 0, s, t->bm[cur], words, 0u);
            hipLaunchKernelGGL(k_queue_to_bitmap, dim3(std::max<uint64_t>(1, std::min<uint64_t>((nf + 255) / 256, grid))),
                               dim3(kBlock), 0, s, t->q[level & 1], t->slots + (level & 3), t->bm[cur]);
            bottom_up = true;
        } else if (bottom_up && nf < V / kBeta && nf < prev_nf) {
            // bitmap -> queue, slot ring restarted at this level
            hipLaunchKernelGGL(k_do_fill, dim3(1), dim3(kBlock), 0, s, (uint32_t*)t->slots, sizeof(Slot) * 4 / 4, 0u);
            hipLaunchKernelGGL(k_bitmap_to_queue, dim3(std::max<uint64_t>(1, std::min<uint64_t>((words + kBlock - 1) / kBlock, grid))),
                               dim3(kBlock), 0, s, t->bm[cur], (uint32_t)V,
                               t->q[level & 1], &t->slots[level & 3].n_res);
            bottom_up = false;
        }
        if (bottom_up) {
            ASC_TRY(timed_launch(t, prof, KERN_BU, s, [&] {
                hipLaunchKernelGGL((k_bfs_bottom_up<OffT, EdgeT>), dim3(bu_grid), dim3(kBlock), 0, s, a.off, a.edges,
                                   t->val, t->visited, t->bm[cur], t->bm[cur ^ 1], (uint32_t)V, level, t->dcnt,
                                   t->stats);
            }));
            cur ^= 1;
        } else {
            rc.node_sum += nf;
            a.q_in = t->q[level & 1];
            a.q_out = t->q[(level + 1) & 1];
            const int lgrid = (int)std::max<uint64_t>(1, std::min<uint64_t>(grid, (nf + kBlock - 1) / kBlock));
            ASC_TRY(timed_launch(t, prof, KERN_LOCAL, s, [&] {
                hipLaunchKernelGGL((k_expand_local<ALG_BFS, OffT, EdgeT>), dim3(lgrid), dim3(kBlock), 0, s, a, level);
            }));
            ASC_TRY(timed_launch(t, prof, KERN_HEAVY, s, [&] {
                hipLaunchKernelGGL((k_expand_heavy<ALG_BFS, OffT, EdgeT>), dim3(grid), dim3(kBlock), 0, s, a, level);
            }));
        }
        if (!prof) {
            hipEvent_t e1 = next_event(t);
            if (!e1) return ASC_ERR_DEVICE;
            ASC_HIP(hipEventRecord(e1, s));
            t->timed.push_back({-1, e0, e1});
        }
        const unsigned long long want = ++t->pub_seq;
        hipLaunchKernelGGL(k_do_publish, dim3(1), dim3(64), 0, s, t->dcnt,
                           bottom_up ? (const Slot*)nullptr : t->slots + ((level + 1) & 3), t->d_hdcnt, want);
        ASC_HIP(hipGetLastError());
        ASC_TRY(wait_host_seq(&t->h_dcnt.p[1].n_found, want, s));
        prev_nf = nf;
        const volatile DoCounters* pub = t->h_dcnt.p;   // written by k_do_publish
        if (bottom_up) {
            nf = pub->n_found;
            mf = pub->m_found;
        } else {
            nf = pub->pad;
            mf = pub->m_frontier;
        }
        m_unexplored = m_unexplored > mf ? m_unexplored - mf : 0;
        if (ctx->cfg.verbose > 1)
            printf("bfs level %u %s: next frontier %llu vertices, %llu edges (%.3f ms)\n", level,
                   bottom_up ? "bottom-up" : "top-down", (unsigned long long)nf, (unsigned long long)mf,
                   now_ms() - t0);
        level++;
    }
    const double t1 = now_ms();
    return finish_traversal(g, ALG_BFS, PATH_DIRECTION_OPT, out, st, rc, t1 - t0, level);
}

// The persistent form: initial state, one cooperative launch, one wait.
// Returns ASC_ERR_UNSUPPORTED (nothing run) when the device refuses a
// cooperative launch of the kernel, so the caller steps levels from the host.
template <class OffT, class EdgeT>
static int run_bfs_persistent(asc_graph* g, uint32_t src, uint32_t* out, asc_stats* st) {
    asc_ctx* ctx = g->ctx;
    ASC_HIP(hipSetDevice(ctx->device));
    ASC_TRY(ensure_fast_state(g));
    TravState* t = g->trav;
    const uint64_t V = g->h.V;
    const uint64_t words = (V + 31) / 32;
    hipStream_t s = ctx->s_main;
    const bool prof = ctx->cfg.profile != 0;
    const int fill_grid = ctx->num_cus * 4;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bfs_persistent<OffT, EdgeT>, kBlock, 0) !=
            hipSuccess ||
        per_cu < 1)
        return ASC_ERR_UNSUPPORTED;
    const int grid = ctx->num_cus * std::min(per_cu, kPersistPerCU);
    DoCounters* ring = (DoCounters*)(t->pctl + kRingWord);

    hipLaunchKernelGGL(k_do_fill, dim3(1), dim3(kBlock), 0, s, (uint32_t*)t->slots, sizeof(Slot) * 4 / 4, 0u);
    hipLaunchKernelGGL(k_do_fill, dim3(1), dim3(kBlock), 0, s, (uint32_t*)t->stats, sizeof(DevStats) / 4, 0u);
    hipLaunchKernelGGL(k_do_fill, dim3(1), dim3(kBlock), 0, s, (uint32_t*)t->dcnt, sizeof(DoCounters) / 4, 0u);
    hipLaunchKernelGGL(k_do_fill, dim3(1), dim3(kBlock), 0, s, t->pctl, (uint64_t)kPctlWords, 0u);
    hipLaunchKernelGGL(k_do_fill, dim3(fill_grid), dim3(kBlock), 0, s, t->val, V, (uint32_t)(V + 1));
    hipLaunchKernelGGL(k_do_fill, dim3(fill_grid), dim3(kBlock), 0, s, t->visited, words, 0u);
    for (int i = 0; i < 3; i++)
        hipLaunchKernelGGL(k_do_fill, dim3(fill_grid), dim3(kBlock), 0, s, t->bm[i], words, 0u);
    hipLaunchKernelGGL(k_do_source, dim3(1), dim3(1), 0, s, t->val, t->visited, t->q[0], t->slots, src);
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, &ring[0].m_found,
                       (unsigned long long)(g->h.off[src + 1] - g->h.off[src]));
    ASC_HIP(hipGetLastError());
    ASC_HIP(hipStreamSynchronize(s));

    TravArgs<OffT, EdgeT> a{};
    a.off = (const OffT*)g->d_off;
    a.edges = (const EdgeT*)g->d_edges;
    a.val = t->val;
    a.visited = t->visited;
    a.slots = t->slots;
    a.heavy_u = t->heavy_u;
    a.heavy_sv = t->heavy_sv;
    a.heavy_base = t->heavy_base;
    a.stats = t->stats;
    a.vs = (uint32_t)V;
    a.heavy_thr = (uint32_t)effective_heavy_threshold(g);
    a.theta = 0xFFFFFFFFu;
    unsigned long long E64 = g->h.E;
    uint32_t* q0 = t->q[0];
    uint32_t* q1 = t->q[1];
    uint32_t* b0 = t->bm[0];
    uint32_t* b1 = t->bm[1];
    uint32_t* b2 = t->bm[2];
    uint32_t* ctl = t->pctl;
    uint32_t V32 = (uint32_t)V;
    uint32_t max_levels = (uint32_t)std::min<uint64_t>(V + 1, 0xFFFFFFFFull);
    void* args[] = {&a, &q0, &q1, &b0, &b1, &b2, &ring, &ctl, &V32, &E64, &max_levels};

    RunCounters rc{};
    t->timed.clear();
    t->ev_next = 0;
    hipEvent_t e0 = next_event(t), e1 = next_event(t);
    if (!e0 || !e1) return ASC_ERR_DEVICE;
    const double t0 = now_ms();
    ASC_HIP(hipEventRecord(e0, s));
    if (hipLaunchCooperativeKernel((const void*)k_bfs_persistent<OffT, EdgeT>, dim3(grid), dim3(kBlock), args, 0,
                                   s) != hipSuccess) {
        (void)hipGetLastError();
        return ASC_ERR_UNSUPPORTED;
    }
    ASC_HIP(hipEventRecord(e1, s));
    ASC_HIP(hipStreamSynchronize(s));
    const double t1 = now_ms();
    t->timed.push_back({prof ? (int)KERN_PERSIST : -1, e0, e1});
    uint32_t h_ctl[4 * kBarLine];
    ASC_HIP(hipMemcpy(h_ctl, ctl, sizeof(h_ctl), hipMemcpyDeviceToHost));
    if (h_ctl[2 * kBarLine]) {
        set_last_error("persistent BFS: a grid barrier wait exceeded its bound");
        return ASC_ERR_DEVICE;
    }
    return finish_traversal(g, ALG_BFS, PATH_DIRECTION_OPT, out, st, rc, t1 - t0, h_ctl[3 * kBarLine]);
}

int bfs_direction_opt(asc_graph* g, uint32_t src, uint32_t* out, asc_stats* st) {
    const bool w = g->h.kind == ASC_KIND_WCSR, o4 = g->off_bytes == 4;
    if (g->ctx->cfg.bfs_persistent) {
        int r;
        if (w) r = o4 ? run_bfs_persistent<uint32_t, WEdge>(g, src, out, st)
                      : run_bfs_persistent<uint64_t, WEdge>(g, src, out, st);
        else r = o4 ? run_bfs_persistent<uint32_t, uint32_t>(g, src, out, st)
                    : run_bfs_persistent<uint64_t, uint32_t>(g, src, out, st);
        if (r != ASC_ERR_UNSUPPORTED) return r;
    }
    if (w) return o4 ? run_bfs_do<uint32_t, WEdge>(g, src, out, st) : run_bfs_do<uint64_t, WEdge>(g, src, out, st);
    return o4 ? run_bfs_do<uint32_t, uint32_t>(g, src, out, st) : run_bfs_do<uint64_t, uint32_t>(g, src, out, st);
}

// ---------------------------------------------------------------------------
// union-find CC
// ---------------------------------------------------------------------------

// Loads that observe other workgroups' hooks (device-coherent, no L1 reuse).
__device__ __forceinline__ uint32_t ld_comp(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Joins the trees of u and v: the root with the larger id is hooked under the
// smaller one by compare-and-swap; a failed or stale attempt re-reads and
// retries from the current parents.  Every write lowers a value, so comp[x]
// <= x holds throughout and the loop ends.
__device__ __forceinline__ void uf_link(uint32_t* comp, uint32_t u, uint32_t v) {
    uint32_t p1 = ld_comp(comp + u), p2 = ld_comp(comp + v);
    while (p1 != p2) {
        const uint32_t hi = max(p1, p2), lo = min(p1, p2);
        const uint32_t ph = ld_comp(comp + hi);
        if (ph == lo) break;
        if (ph == hi && atomicCAS(comp + hi, hi, lo) == hi) break;
        p1 = ld_comp(comp + ld_comp(comp + hi));
        p2 = ld_comp(comp + lo);
    }
}

__global__ __launch_bounds__(kBlock) void k_uf_init(uint32_t* comp, uint32_t V) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < V; v += gridDim.x * kBlock) comp[v] = v;
}

// Links every vertex with its r-th neighbour (when it has one).
template <class OffT, class EdgeT>
__global__ __launch_bounds__(kBlock) void k_uf_sample(const OffT* off, const EdgeT* edges, uint32_t* comp, uint32_t V,
                                                      uint32_t r, DevStats* stats) {
    __shared__ unsigned long long s_tmp[kWavesPerBlock];
    unsigned long long n = 0;
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < V; v += gridDim.x * kBlock) {
        const OffT b = off[v];
        if ((uint64_t)(off[v + 1] - b) > r) {
            uf_link(comp, v, edge_dst(edges[b + r]));
            n++;
        }
    }
    block_add(n, s_tmp, &stats->edges_uf);
}

// Points every vertex at its root.  No hooks run concurrently, so every value
// read (even an older one) is an ancestor and the chase ends at the root.
__global__ __launch_bounds__(kBlock) void k_uf_compress(uint32_t* comp, uint32_t V) {
    for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < V; v += gridDim.x * kBlock) {
        uint32_t p = comp[v], pp = comp[p];
        while (p != pp) {
            p = pp;
            pp = comp[p];
        }
        comp[v] = p;
    }
}

__global__ void k_uf_gather(const uint32_t* comp, const uint32_t* idx, uint32_t n, uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = comp[idx[i]];
}

// Links the edges from position `skip` on of every vertex whose component
// is not `giant`, edge-balanced inside each workgroup of 256 consecutive
// vertices (LDS exclusive scan of the degrees + per-lane binary search).
// Edges of giant-component vertices are covered from their other endpoint.
template <class OffT, class EdgeT>
__global__ __launch_bounds__(kBlock) void k_uf_rest(const OffT* off, const EdgeT* edges, uint32_t* comp, uint32_t V,
                                                    uint32_t giant, uint32_t skip, DevStats* stats) {
    __shared__ unsigned long long s_pref[kBlock];
    __shared__ OffT s_beg[kBlock];
    __shared__ unsigned long long s_scan[kWavesPerBlock];
    __shared__ unsigned long long s_tmp[kWavesPerBlock];
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
    unsigned long long linked = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * kBlock; base < V; base += (uint64_t)gridDim.x * kBlock) {
        const uint64_t v = base + tid;
        unsigned long long deg = 0;
        OffT beg = 0;
        if (v < V && ld_comp(comp + v) != giant) {
            const OffT b = off[v], e = off[v + 1];
            if ((uint64_t)(e - b) > skip) {
                beg = b + (OffT)skip;
                deg = (uint64_t)(e - beg);
            }
        }
        unsigned long long total;
        const unsigned long long pre = block_scan_excl(deg, s_scan, total);
        s_pref[tid] = pre;
        s_beg[tid] = beg;
        __syncthreads();
        const uint32_t cnt = (uint32_t)min((uint64_t)kBlock, V - base);
        for (unsigned long long eb = w * kWave; eb < total; eb += kBlock) {
            const unsigned long long e = eb + lane;
            if (e < total) {
                uint32_t k = 0;
#pragma unroll
                for (uint32_t step = kBlock / 2; step > 0; step >>= 1)
                    if (k + step < cnt && s_pref[k + step] <= e) k += step;
                uf_link(comp, (uint32_t)(base + k), edge_dst(edges[s_beg[k] + (OffT)(e - s_pref[k])]));
            }
        }
        linked += total;
        __syncthreads();
    }
    block_add(linked, s_tmp, &stats->edges_uf);
}

template <class OffT, class EdgeT>
static int run_cc_uf(asc_graph* g, uint32_t* out, asc_stats* st) {
    asc_ctx* ctx = g->ctx;
    ASC_HIP(hipSetDevice(ctx->device));
    ASC_TRY(ensure_fast_state(g));
    TravState* t = g->trav;
    const uint64_t V = g->h.V;
    hipStream_t s = ctx->s_main;
    const bool prof = ctx->cfg.profile != 0;
    const int grid = ctx->num_cus * (g->blocks_per_cu > 0 ? g->blocks_per_cu : kBlocksPerCU);
    const int vgrid = (int)std::max<uint64_t>(1, std::min<uint64_t>((V + kBlock - 1) / kBlock, (uint64_t)grid));
    const OffT* off = (const OffT*)g->d_off;
    const EdgeT* edges = (const EdgeT*)g->d_edges;
    uint32_t* comp = t->val;

    // sampled vertices: a fixed SplitMix64 sequence, so runs are repeatable
    const uint32_t ns = (uint32_t)std::min<uint64_t>(kSamples, V);
    uint64_t x = 0x243f6a8885a308d3ull;
    for (uint32_t i = 0; i < ns; i++) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        t->h_sample[i] = (uint32_t)(z % V);
    }
    uint32_t* d_idx = t->heavy_u;              // kHeavyCap >= kSamples * 2 entries of scratch
    uint32_t* d_lab = t->heavy_u + kSamples;
    ASC_HIP(hipMemcpyAsync(d_idx, t->h_sample.p, 4 * (size_t)ns, hipMemcpyHostToDevice, s));
    ASC_HIP(hipMemsetAsync(t->stats, 0, sizeof(DevStats), s));
    ASC_HIP(hipStreamSynchronize(s));

    RunCounters rc{};
    t->timed.clear();
    t->ev_next = 0;
    const double t0 = now_ms();
    hipEvent_t e0 = nullptr;
    if (!prof) {
        e0 = next_event(t);
        if (!e0) return ASC_ERR_DEVICE;
        ASC_HIP(hipEventRecord(e0, s));
    }
    hipLaunchKernelGGL(k_uf_init, dim3(vgrid), dim3(kBlock), 0, s, comp, (uint32_t)V);
    for (int r = 0; r < kSampleRounds; r++) {
        ASC_TRY(timed_launch(t, prof, KERN_UF, s, [&] {
            hipLaunchKernelGGL((k_uf_sample<OffT, EdgeT>), dim3(vgrid), dim3(kBlock), 0, s, off, edges, comp,
                               (uint32_t)V, (uint32_t)r, t->stats);
        }));
        hipLaunchKernelGGL(k_uf_compress, dim3(vgrid), dim3(kBlock), 0, s, comp, (uint32_t)V);
    }
    hipLaunchKernelGGL(k_uf_gather, dim3((ns + 255) / 256), dim3(256), 0, s, comp, d_idx, ns, d_lab);
    ASC_HIP(hipGetLastError());
    ASC_HIP(hipMemcpyAsync(t->h_sample.p, d_lab, 4 * (size_t)ns, hipMemcpyDeviceToHost, s));
    ASC_HIP(hipStreamSynchronize(s));
    // most common sampled label (ties: the smallest)
    std::unordered_map<uint32_t, uint32_t> freq;
    uint32_t giant = t->h_sample[0], best = 0;
    for (uint32_t i = 0; i < ns; i++) {
        const uint32_t c = ++freq[t->h_sample[i]];
        if (c > best || (c == best && t->h_sample[i] < giant)) {
            best = c;
            giant = t->h_sample[i];
        }
    }
    ASC_TRY(timed_launch(t, prof, KERN_UF, s, [&] {
        hipLaunchKernelGGL((k_uf_rest<OffT, EdgeT>), dim3(vgrid), dim3(kBlock), 0, s, off, edges, comp, (uint32_t)V,
                           giant, (uint32_t)kSampleRounds, t->stats);
    }));
    hipLaunchKernelGGL(k_uf_compress, dim3(vgrid), dim3(kBlock), 0, s, comp, (uint32_t)V);
    ASC_HIP(hipGetLastError());
    if (!prof) {
        hipEvent_t e1 = next_event(t);
        if (!e1) return ASC_ERR_DEVICE;
        ASC_HIP(hipEventRecord(e1, s));
        t->timed.push_back({-1, e0, e1});
    }
    ASC_HIP(hipStreamSynchronize(s));
    const double t1 = now_ms();
    rc.node_sum = V;
    return finish_traversal(g, ALG_CC, PATH_UNION_FIND, out, st, rc, t1 - t0, kSampleRounds + 1);
}

int cc_union_find(asc_graph* g, uint32_t* out, asc_stats* st) {
    const bool w = g->h.kind == ASC_KIND_WCSR, o4 = g->off_bytes == 4;
    if (w) return o4 ? run_cc_uf<uint32_t, WEdge>(g, out, st) : run_cc_uf<uint64_t, WEdge>(g, out, st);
    return o4 ? run_cc_uf<uint32_t, uint32_t>(g, out, st) : run_cc_uf<uint64_t, uint32_t>(g, out, st);
}

}  // namespace asc
