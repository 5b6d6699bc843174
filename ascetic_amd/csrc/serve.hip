// Batched query serving: continuous batching of BFS and SSSP queries over a
// resident graph.  No counterpart in the reference, which runs one traversal
// per process (main.cu:4-22).
//
// Two engines, each with up to 64 lanes (one query per lane):
//   BFS   bitset-per-vertex multi-source BFS: seen[v] and the frontier words
//         X[v] / Y[v] are 64-bit, bit q = lane q, so one scan of an edge
//         (u, v) advances every lane whose frontier holds u:
//         new = X[u] & ~seen[v].  Levels are kept vertex-major in 8 bits,
//         lev[v][q] = hops + 1 up to 254 (a row of 64 lanes is one 64-byte
//         line, so the lanes that reach a vertex in one step write one line);
//         a lane whose levels go beyond is answered by the single-query
//         engine when it leaves its lane.  lev[v][q] is meaningful only
//         where bit q of seen[v] is set (it is written when the bit is), so
//         an admitted lane clears its seen bits and leaves its level bytes.
//   SSSP  distance lanes: dist[v][q] vertex-major in 16 bits (a row of 64
//         lanes is one 128-byte line), one thread per edge relaxing every
//         lane whose frontier holds the edge's source; improved lanes are
//         ORed into the destination's 64-bit frontier word.  A candidate
//         above 65534 marks the lane as overflowed instead; such a query is
//         answered by the single-query engine when it leaves its lane, so
//         every result is exact on any graph.
//   On a symmetric graph a BFS lane whose next frontier holds more than
//   1/14 of its unexplored edges switches to bottom-up steps (Beamer et al.),
//   and back once its frontier shrinks below V/24 vertices: every vertex that
//   some bottom-up lane has not seen scans its own adjacency for frontier
//   bits of those lanes, X[u] & need, and stops once every needed lane is
//   found.  Bottom-up runs before the top-down kernels of the same step,
//   which skip the bits of bottom-up lanes.
// One engine step advances every admitted query by one frontier round, using
// the same edge-balanced workgroup expansion and grid-cooperative heavy list
// as the single-query engine.  A lane whose query produced no new frontier bit
// in a step is finished; its result is copied out (if requested) and the lane
// is free for the next waiting query at the next step boundary.  Lanes are
// granted up to the device-memory budget (V bytes of per-query state per BFS
// lane, 2 V per SSSP lane).
// Every query's values equal those of the single-query engine: BFS levels are
// level-synchronous per lane, and SSSP lanes converge to the unique distances.
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <type_traits>
#include <unordered_map>

#include "device_common.hpp"
#include "internal.hpp"
#include "traverse.hpp"

namespace asc {

constexpr int kMaxLanes = 64;

struct SeedArgs {
    uint32_t n;
    uint32_t lane[kMaxLanes];
    uint32_t src[kMaxLanes];
};

// What changes from one engine step to the next, read by the step's kernels
// from device memory (written by a copy at the head of the step), so that one
// recorded launch sequence serves every step of the same parity.
struct StepParams {
    uint32_t step;                   // engine step number
    uint32_t pad;
    unsigned long long bu;           // BFS lanes that take a bottom-up step
};

template <class OffT, class EdgeT>
struct BatchArgs {
    const OffT* off;
    const EdgeT* edges;
    unsigned long long* X;           // frontier bits of this step (read and cleared)
    unsigned long long* Y;           // frontier bits of the next step
    unsigned long long* seen;        // BFS
    uint8_t* lev;                    // BFS: lev[v * stride + q], hops + 1, where seen[v] has bit q
    uint16_t* dist;                  // SSSP: dist[v * stride + q], 0xFFFF = unreached
    unsigned long long* ovf;         // lanes whose levels / distances left 8 / 16 bits
    const uint32_t* q_in;
    uint32_t* q_out;
    Slot* slots;                     // pad[0] of slot (s+1)%4: lanes that advanced in step s
    uint32_t* heavy_u;
    unsigned long long* heavy_bits;
    unsigned long long* heavy_base;
    const uint32_t* s0;              // step at which each lane was seeded
    const StepParams* sp;
    uint64_t V;
    uint32_t stride;                 // SSSP row length (lanes)
    uint32_t heavy_thr;
};

// -- BFS lanes ------------------------------------------------------------------

constexpr uint32_t kLev8Max = 0xFEu;   // largest level an 8-bit BFS lane holds (0xFF: unreached)

// Level of v in the lanes of `nw` (each (v, lane) is set by exactly one thread,
// the one whose seen update found the lane new): byte stores into v's row; a
// level beyond 8 bits marks the lane overflowed.
template <class OffT, class EdgeT>
__device__ __forceinline__ void msbfs_set_levels(const BatchArgs<OffT, EdgeT>& a, uint32_t v, unsigned long long nw,
                                                 uint32_t step) {
    uint8_t* row = a.lev + (size_t)v * a.stride;
    unsigned long long over = 0;
    for (unsigned long long m = nw; m; m &= m - 1) {
        const uint32_t q = (uint32_t)__ffsll((long long)m) - 1;
        const uint32_t l = step - a.s0[q] + 2;
        if (l > kLev8Max) over |= 1ull << q;
        row[q] = (uint8_t)(l > kLev8Max ? kLev8Max : l);
    }
    if (over) atomicOr(a.ovf, over);
}

// Top-down visit of kMsUnroll edges per lane (edge j counts when bit j of
// `valid` is set): for each destination v, the lanes of `bits` not yet in
// seen[v] claim v (one atomic OR arbitrates each lane), record its level and
// enter its next-frontier word.  The seen words of all edges are loaded, then
// their seen atomics issued, then their next-frontier atomics, before any
// result is used.  push[j]: v[j] enters the next frontier queue.
constexpr int kMsUnroll = 4;

template <class OffT, class EdgeT>
__device__ __forceinline__ void msbfs_visit_batch(const BatchArgs<OffT, EdgeT>& a, const uint32_t (&v)[kMsUnroll],
                                                  const unsigned long long (&bits)[kMsUnroll], uint32_t valid,
                                                  uint32_t step, unsigned long long& act, bool (&push)[kMsUnroll]) {
    unsigned long long nw[kMsUnroll], old[kMsUnroll];
#pragma unroll
    for (int j = 0; j < kMsUnroll; j++) nw[j] = ((valid >> j) & 1u) ? bits[j] & ~a.seen[v[j]] : 0ull;
#pragma unroll
    for (int j = 0; j < kMsUnroll; j++)
        if (nw[j]) old[j] = atomicOr(a.seen + v[j], nw[j]);
#pragma unroll
    for (int j = 0; j < kMsUnroll; j++) {
        push[j] = false;
        if (!nw[j]) continue;
        nw[j] &= ~old[j];
        if (!nw[j]) continue;
        act |= nw[j];
        msbfs_set_levels(a, v[j], nw[j], step);
    }
#pragma unroll
    for (int j = 0; j < kMsUnroll; j++)
        if (nw[j]) old[j] = atomicOr(a.Y + v[j], nw[j]);
#pragma unroll
    for (int j = 0; j < kMsUnroll; j++) push[j] = nw[j] && old[j] == 0;
}

template <class OffT, class EdgeT>
__global__ __launch_bounds__(kBlock) void k_msbfs_local(BatchArgs<OffT, EdgeT> a) {
    const uint32_t step = a.sp->step;
    const unsigned long long td_mask = ~a.sp->bu;   // lanes expanded top-down
    __shared__ uint32_t s_pref[kBlock];
    __shared__ OffT s_beg[kBlock];
    __shared__ unsigned long long s_bits[kBlock];
    __shared__ uint32_t s_tmp[kWavesPerBlock];
    __shared__ uint32_t s_stage[kWavesPerBlock][kStageCap];
    Slot* cur = a.slots + (step & 3);
    Slot* nxt = a.slots + ((step + 1) & 3);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Slot* z = a.slots + ((step + 2) & 3);
        z->n_res = 0;
        z->heavy = 0;
        z->pad[0] = 0;
    }
    const uint32_t n = cur->n_res;
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
    WaveStage<kStageCap> st{s_stage[w], 0};
    unsigned long long act = 0;
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const uint32_t i = base + tid;
        uint32_t deg = 0;
        OffT beg = 0;
        unsigned long long bits = 0;
        if (i < n) {
            const uint32_t u = a.q_in[i];
            bits = a.X[u] & td_mask;   // bottom-up lanes' bits were consumed already
            a.X[u] = 0;
            if (bits) {
                beg = a.off[u];
                const uint64_t d = (uint64_t)(a.off[u + 1] - beg);
                if (d >= a.heavy_thr) {
                    const unsigned long long p = atomicAdd(&cur->heavy, (1ull << kHeavyShift) | d);
                    const uint32_t idx = (uint32_t)(p >> kHeavyShift);
                    a.heavy_u[idx] = u;
                    a.heavy_bits[idx] = bits;
                    a.heavy_base[idx] = p & kHeavyMask;
                } else {
                    deg = (uint32_t)d;
                }
            }
        }
        uint32_t total;
        const uint32_t pre = block_scan_excl(deg, s_tmp, total);
        s_pref[tid] = pre;
        s_beg[tid] = beg;
        s_bits[tid] = bits;
        __syncthreads();
        const uint32_t cnt = min(n - base, (uint32_t)kBlock);
        for (uint32_t eb = w * kWave * kMsUnroll; eb < total; eb += kBlock * kMsUnroll) {
            uint32_t v[kMsUnroll], valid = 0;
            unsigned long long bb[kMsUnroll];
#pragma unroll
            for (int j = 0; j < kMsUnroll; j++) {
                const uint32_t idx = eb + j * kWave + lane;
                if (idx < total) valid |= 1u << j;
                const uint32_t e = min(idx, total - 1);
                uint32_t k = 0;
#pragma unroll
                for (uint32_t stp = kBlock / 2; stp > 0; stp >>= 1)
                    if (k + stp < cnt && s_pref[k + stp] <= e) k += stp;
                v[j] = edge_dst(a.edges[s_beg[k] + (OffT)(e - s_pref[k])]);
                bb[j] = s_bits[k];
            }
            bool push[kMsUnroll];
            msbfs_visit_batch(a, v, bb, valid, step, act, push);
#pragma unroll
            for (int j = 0; j < kMsUnroll; j++) st.push(push[j], v[j], a.q_out, &nxt->n_res);
        }
        __syncthreads();
    }
    st.flush(a.q_out, &nxt->n_res);
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) act |= __shfl_xor(act, d, kWave);
    if (lane == 0 && act) atomicOr(&nxt->pad[0], act);
}

template <class OffT, class EdgeT>
__global__ __launch_bounds__(kBlock) void k_msbfs_heavy(BatchArgs<OffT, EdgeT> a) {
    const uint32_t step = a.sp->step;
    constexpr int kWin = 8;
    __shared__ unsigned long long s_hb[kWin], s_hbits[kWin];
    __shared__ OffT s_hoff[kWin];
    __shared__ uint32_t s_k0;
    __shared__ uint32_t s_stage[kWavesPerBlock][kStageCap];
    Slot* cur = a.slots + (step & 3);
    Slot* nxt = a.slots + ((step + 1) & 3);
    const unsigned long long h = cur->heavy;
    const uint32_t nh = (uint32_t)(h >> kHeavyShift);
    const uint64_t total = h & kHeavyMask;
    if (nh == 0) return;
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
    WaveStage<kStageCap> st{s_stage[w], 0};
    unsigned long long act = 0;
    for (uint64_t t0 = (uint64_t)blockIdx.x * kTile; t0 < total; t0 += (uint64_t)gridDim.x * kTile) {
        if (tid == 0) {
            uint32_t lo = 0, hi = nh - 1;
            while (lo < hi) {
                uint32_t mid = (lo + hi + 1) >> 1;
                if (a.heavy_base[mid] <= t0) lo = mid; else hi = mid - 1;
            }
            s_k0 = lo;
        }
        __syncthreads();
        const uint32_t k0 = s_k0;
        const uint32_t nwin = min(nh - k0, (uint32_t)kWin);
        if (tid < nwin) {
            s_hb[tid] = a.heavy_base[k0 + tid];
            s_hoff[tid] = a.off[a.heavy_u[k0 + tid]];
            s_hbits[tid] = a.heavy_bits[k0 + tid];
        }
        __syncthreads();
        const uint32_t tlen = (uint32_t)min((uint64_t)kTile, total - t0);
        for (uint32_t r = w * kWave * kMsUnroll; r < tlen; r += kBlock * kMsUnroll) {
            uint32_t v[kMsUnroll], valid = 0;
            unsigned long long bb[kMsUnroll];
#pragma unroll
            for (int jj = 0; jj < kMsUnroll; jj++) {
                const uint32_t rr = r + jj * kWave + lane;
                if (rr < tlen) valid |= 1u << jj;
                const uint64_t e = t0 + min(rr, tlen - 1);
                uint32_t k = 0;
#pragma unroll
                for (uint32_t j = 1; j < (uint32_t)kWin; j++)
                    if (j < nwin && s_hb[j] <= e) k = j;
                v[jj] = edge_dst(a.edges[s_hoff[k] + (OffT)(e - s_hb[k])]);
                bb[jj] = s_hbits[k];
            }
            bool push[kMsUnroll];
            msbfs_visit_batch(a, v, bb, valid, step, act, push);
#pragma unroll
            for (int jj = 0; jj < kMsUnroll; jj++) st.push(push[jj], v[jj], a.q_out, &nxt->n_res);
        }
        __syncthreads();
    }
    st.flush(a.q_out, &nxt->n_res);
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) act |= __shfl_xor(act, d, kWave);
    if (lane == 0 && act) atomicOr(&nxt->pad[0], act);
}

// Neighbour ids loaded before their frontier words are probed (8 and 16 ran
// slower: the scan of a vertex usually ends within its first few neighbours).
constexpr int kBuBatch = 4;

// One bottom-up step for the lanes in the step's bottom-up mask (symmetric
// graph: in-neighbours are the adjacency); returns at once when it is empty.
// Thread per vertex; four neighbour ids are loaded before their frontier
// words are probed.
template <class OffT, class EdgeT>
__global__ __launch_bounds__(kBlock) void k_msbfs_bottom_up(BatchArgs<OffT, EdgeT> a) {
    __shared__ uint32_t s_stage[kWavesPerBlock][kStageCap];
    const uint32_t step = a.sp->step;
    const unsigned long long bu_mask = a.sp->bu;
    if (!bu_mask) return;
    Slot* nxt = a.slots + ((step + 1) & 3);
    WaveStage<kStageCap> st{s_stage[wave_id()], 0};
    unsigned long long act = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * kBlock; base < a.V; base += (uint64_t)gridDim.x * kBlock) {
        const uint64_t v = base + threadIdx.x;
        bool push = false;
        if (v < a.V) {
            const unsigned long long need = bu_mask & ~a.seen[v];
            if (need) {
                const OffT b = a.off[v], e = a.off[v + 1];
                unsigned long long nw = 0;
                for (OffT i = b; i < e && nw != need; i += kBuBatch) {
                    uint32_t u[kBuBatch];
#pragma unroll
                    for (int j = 0; j < kBuBatch; j++) u[j] = edge_dst(a.edges[min(i + (OffT)j, e - 1)]);
#pragma unroll
                    for (int j = 0; j < kBuBatch; j++) nw |= a.X[u[j]] & need;
                }
                if (nw) {
                    const unsigned long long old = atomicOr(a.seen + v, nw);
                    nw &= ~old;
                    if (nw) {
                        act |= nw;
                        msbfs_set_levels(a, (uint32_t)v, nw, step);
                        push = atomicOr(a.Y + v, nw) == 0;
                    }
                }
            }
        }
        st.push(push, (uint32_t)v, a.q_out, &nxt->n_res);
    }
    st.flush(a.q_out, &nxt->n_res);
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) act |= __shfl_xor(act, d, kWave);
    if (lane_id() == 0 && act) atomicOr(&nxt->pad[0], act);
}

// Per lane q: vertices with bit q in Y (cnt[q]) and, for the top-down lanes
// of the step, their summed degrees (cnt[64 + q]), over the queue q of the
// vertices whose Y word is non-zero (count in the next step's slot).  A wave takes 64
// queue entries; for each bit present in the wave a ballot counts the
// vertices, and for the masked bits a wave sum adds the degrees.
template <class OffT>
__global__ __launch_bounds__(kBlock) void k_lane_counts(const OffT* off, const unsigned long long* Y, const uint32_t* q,
                                                        const Slot* slots, const StepParams* sp,
                                                        unsigned long long* cnt) {
    const Slot* slot = slots + ((sp->step + 1) & 3);
    const unsigned long long edge_mask = ~sp->bu;
    __shared__ unsigned long long s_n[kMaxLanes], s_m[kMaxLanes];
    const uint32_t lane = lane_id();
    if (threadIdx.x < kMaxLanes) {
        s_n[threadIdx.x] = 0;
        s_m[threadIdx.x] = 0;
    }
    __syncthreads();
    const uint32_t n = slot->n_res;
    unsigned long long my_n = 0, my_m = 0;
    const uint64_t nw = (uint64_t)gridDim.x * kWavesPerBlock;
    for (uint64_t wv = (uint64_t)blockIdx.x * kWavesPerBlock + wave_id(); wv * kWave < n; wv += nw) {
        const uint64_t i = wv * kWave + lane;
        unsigned long long y = 0, d = 0;
        if (i < n) {
            const uint32_t v = q[i];
            y = Y[v];
            if (y & edge_mask) d = (unsigned long long)(off[v + 1] - off[v]);
        }
        unsigned long long present = y;
#pragma unroll
        for (int k = kWave / 2; k > 0; k >>= 1) present |= __shfl_xor(present, k, kWave);
        for (unsigned long long m = present; m; m &= m - 1) {
            const uint32_t b = (uint32_t)__ffsll((long long)m) - 1;
            const bool on = (y >> b) & 1ull;
            const unsigned long long n_b = __popcll(wave_ballot(on));
            if (lane == b) my_n += n_b;
            if ((edge_mask >> b) & 1ull) {
                const unsigned long long m_b = wave_sum(on ? d : 0ull);
                if (lane == b) my_m += m_b;
            }
        }
    }
    if (my_n) {
        atomicAdd(&s_n[lane], my_n);
        if (my_m) atomicAdd(&s_m[lane], my_m);
    }
    __syncthreads();
    if (threadIdx.x < kMaxLanes && s_n[threadIdx.x]) {
        atomicAdd(cnt + threadIdx.x, s_n[threadIdx.x]);
        if (s_m[threadIdx.x]) atomicAdd(cnt + kMaxLanes + threadIdx.x, s_m[threadIdx.x]);
    }
}

// -- SSSP lanes -----------------------------------------------------------------

constexpr uint32_t kD16Inf = 0xFFFFu;   // unreached in a 16-bit distance lane
// Batches of the lane relax: groups of 8 lanes whose row loads are issued
// together, 4-lane words whose compare-and-swaps are issued together (wider
// batches cost occupancy: 8 / 16 ran slower at scale 26).
constexpr uint32_t kRB = 4;
constexpr uint32_t kRW = 4;

// Packed 16-bit lane arithmetic (v_pk_add_u16 with clamp, v_pk_min_u16): two
// distance lanes per 32-bit register.
typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_add_sat16(uint32_t x, uint32_t y) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_add_sat(__builtin_bit_cast(u16x2_t, x),
                                                                      __builtin_bit_cast(u16x2_t, y)));
}
__device__ __forceinline__ uint32_t pk_min16(uint32_t x, uint32_t y) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2_t, x),
                                                                  __builtin_bit_cast(u16x2_t, y)));
}
__device__ __forceinline__ unsigned long long pk_min16x4(unsigned long long x, unsigned long long y) {
    return (unsigned long long)pk_min16((uint32_t)x, (uint32_t)y) |
           ((unsigned long long)pk_min16((uint32_t)(x >> 32), (uint32_t)(y >> 32)) << 32);
}
// du + w (saturating) for the lanes of a 4-lane word set in `nib`, 0xFFFF in the others.
__device__ __forceinline__ unsigned long long word_candidates(unsigned long long du, uint32_t w2, uint32_t nib) {
    const uint32_t mlo = ((nib & 1u) ? 0xFFFFu : 0u) | ((nib & 2u) ? 0xFFFF0000u : 0u);
    const uint32_t mhi = ((nib & 4u) ? 0xFFFFu : 0u) | ((nib & 8u) ? 0xFFFF0000u : 0u);
    const uint32_t clo = pk_add_sat16((uint32_t)du, w2) | ~mlo;
    const uint32_t chi = pk_add_sat16((uint32_t)(du >> 32), w2) | ~mhi;
    return (unsigned long long)clo | ((unsigned long long)chi << 32);
}
// Lanes (bit j for lane j of the word) whose 16-bit values differ.
__device__ __forceinline__ uint32_t lanes_changed(unsigned long long x, unsigned long long y) {
    const unsigned long long d = x ^ y;
    return ((d & 0xFFFFull) ? 1u : 0u) | ((d & 0xFFFF0000ull) ? 2u : 0u) | ((d & 0xFFFF00000000ull) ? 4u : 0u) |
           ((d & 0xFFFF000000000000ull) ? 8u : 0u);
}

// Relaxes edge (u, v, w) for every lane in `bits` (one thread per edge; the
// lanes of one vertex share its distance row).  Rows are `stride` (a multiple
// of 8) u16 long and 16-byte aligned; phase 1 reads both rows in 16-byte
// groups of 8 lanes with no atomics in between, so the loads of all active
// groups are in flight together; phase 2 lowers the lanes whose candidate
// beat the value read, one compare-and-swap per 64-bit word of four lanes.
// Returns true when v must be queued.
template <class OffT>
__device__ __forceinline__ bool mssp_relax(const BatchArgs<OffT, WEdge>& a, const uint16_t* du_row, uint32_t v,
                                           uint32_t wgt, unsigned long long bits, unsigned long long& act) {
    const uint4* du4 = (const uint4*)du_row;
    const uint4* dv4 = (const uint4*)(a.dist + (size_t)v * a.stride);
    const uint32_t w2 = (wgt >= kD16Inf ? kD16Inf : wgt) * 0x10001u;   // the weight in both halves
    unsigned long long maybe = 0, over = 0;
    // phase 1, kRB groups of eight lanes at a time: the loads of a batch's
    // active groups are issued before any of them is used
    for (uint32_t gb = 0; gb < kMaxLanes / 8; gb += kRB) {
        const unsigned long long bb = kRB == 8 ? bits : (bits >> (8 * gb)) & ((1ull << (8 * kRB)) - 1);
        if (!bb) continue;
        uint4 x[kRB], y[kRB];
#pragma unroll
        for (uint32_t k = 0; k < kRB; k++)
            if ((bb >> (8 * k)) & 0xFFu) {
                x[k] = du4[gb + k];
                y[k] = dv4[gb + k];
            }
#pragma unroll
        for (uint32_t k = 0; k < kRB; k++) {
            const uint32_t byte = (uint32_t)(bb >> (8 * k)) & 0xFFu;
            if (!byte) continue;
            const uint32_t xs[4] = {x[k].x, x[k].y, x[k].z, x[k].w}, ys[4] = {y[k].x, y[k].y, y[k].z, y[k].w};
            // two lanes per 32-bit word: candidate = du + w saturating at
            // 0xFFFF (overflow), improved where min(candidate, dv) != dv
            uint32_t imp8 = 0, ov8 = 0;
#pragma unroll
            for (uint32_t q = 0; q < 4; q++) {
                const uint32_t c = pk_add_sat16(xs[q], w2);
                const uint32_t d = pk_min16(c, ys[q]) ^ ys[q];
                imp8 |= ((d & 0xFFFFu) != 0 ? 1u : 0u) << (2 * q);
                imp8 |= ((d >> 16) != 0 ? 1u : 0u) << (2 * q + 1);
                ov8 |= ((c & 0xFFFFu) == 0xFFFFu ? 1u : 0u) << (2 * q);
                ov8 |= ((c >> 16) == 0xFFFFu ? 1u : 0u) << (2 * q + 1);
            }
            maybe |= (unsigned long long)(imp8 & byte) << (8 * (gb + k));
            over |= (unsigned long long)(ov8 & byte) << (8 * (gb + k));
        }
    }
    if (over) atomicOr(a.ovf, over);
    if (!maybe) return false;
    // phase 2, per 64-bit word of four lanes: one compare-and-swap lowers
    // every lane of the word whose candidate beats it; the compare-and-swaps
    // of kRW words are in flight together, a lost race is retried alone
    const unsigned long long* du64 = (const unsigned long long*)du_row;
    unsigned long long* row = (unsigned long long*)(a.dist + (size_t)v * a.stride);
    unsigned long long imp = 0;
    for (uint32_t wb = 0; wb < kMaxLanes / 4; wb += kRW) {
        const unsigned long long mb = kRW == 16 ? maybe : (maybe >> (4 * wb)) & ((1ull << (4 * kRW)) - 1);
        if (!mb) continue;
        unsigned long long xu[kRW], old[kRW], nw[kRW], prev[kRW];
        uint32_t got[kRW];
#pragma unroll
        for (uint32_t k = 0; k < kRW; k++)
            if ((mb >> (4 * k)) & 0xFu) {
                xu[k] = du64[wb + k];
                old[k] = row[wb + k];
            }
        unsigned long long cand[kRW];
#pragma unroll
        for (uint32_t k = 0; k < kRW; k++) {
            const uint32_t nib = (uint32_t)(mb >> (4 * k)) & 0xFu;
            got[k] = 0;
            if (!nib) continue;
            // candidates of the word's lanes in `nib`, 0xFFFF (no change) elsewhere
            cand[k] = word_candidates(xu[k], w2, nib);
            nw[k] = pk_min16x4(cand[k], old[k]);
            got[k] = lanes_changed(nw[k], old[k]);
        }
#pragma unroll
        for (uint32_t k = 0; k < kRW; k++)
            if (got[k]) prev[k] = atomicCAS(row + wb + k, old[k], nw[k]);
#pragma unroll
        for (uint32_t k = 0; k < kRW; k++) {
            if (!got[k]) continue;
            if (prev[k] == old[k]) {
                imp |= (unsigned long long)got[k] << (4 * (wb + k));
                continue;
            }
            // lost a race: retry this word against the value that won
            unsigned long long o = prev[k];
            for (;;) {
                const unsigned long long n2 = pk_min16x4(cand[k], o);
                const uint32_t g2 = lanes_changed(n2, o);
                if (!g2) break;
                const unsigned long long p2 = atomicCAS(row + wb + k, o, n2);
                if (p2 == o) {
                    imp |= (unsigned long long)g2 << (4 * (wb + k));
                    break;
                }
                o = p2;
            }
        }
    }
    if (!imp) return false;
    act |= imp;
    return atomicOr(a.Y + v, imp) == 0;
}

template <class OffT>
__global__ __launch_bounds__(kBlock) void k_mssp_local(BatchArgs<OffT, WEdge> a) {
    const uint32_t step = a.sp->step;
    __shared__ uint32_t s_pref[kBlock];
    __shared__ OffT s_beg[kBlock];
    __shared__ unsigned long long s_bits[kBlock];
    __shared__ uint32_t s_tmp[kWavesPerBlock];
    __shared__ uint32_t s_stage[kWavesPerBlock][kStageCap];
    __shared__ uint4 s_rows[kBlock][kMaxLanes / 8];   // the tile's source rows (read by every edge of a vertex)
    Slot* cur = a.slots + (step & 3);
    Slot* nxt = a.slots + ((step + 1) & 3);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Slot* z = a.slots + ((step + 2) & 3);
        z->n_res = 0;
        z->heavy = 0;
        z->pad[0] = 0;
    }
    const uint32_t n = cur->n_res;
    const uint32_t tid = threadIdx.x, w = wave_id();
    const uint32_t row_u4 = a.stride / 8;
    WaveStage<kStageCap> st{s_stage[w], 0};
    unsigned long long act = 0;
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const uint32_t i = base + tid;
        uint32_t deg = 0, u = 0;
        OffT beg = 0;
        unsigned long long bits = 0;
        if (i < n) {
            u = a.q_in[i];
            bits = a.X[u];
            a.X[u] = 0;
            beg = a.off[u];
            const uint64_t d = (uint64_t)(a.off[u + 1] - beg);
            if (d >= a.heavy_thr) {
                const unsigned long long p = atomicAdd(&cur->heavy, (1ull << kHeavyShift) | d);
                const uint32_t idx = (uint32_t)(p >> kHeavyShift);
                a.heavy_u[idx] = u;
                a.heavy_bits[idx] = bits;
                a.heavy_base[idx] = p & kHeavyMask;
            } else {
                deg = (uint32_t)d;
                if (deg) {
                    // u's row as of the tile's start: a value lowered later in
                    // the step queues u again, so an older (larger) one is safe
                    const uint4* src = (const uint4*)(a.dist + (size_t)u * a.stride);
                    for (uint32_t g = 0; g < row_u4; g++) s_rows[tid][g] = src[g];
                }
            }
        }
        uint32_t total;
        const uint32_t pre = block_scan_excl(deg, s_tmp, total);
        s_pref[tid] = pre;
        s_beg[tid] = beg;
        s_bits[tid] = bits;
        __syncthreads();
        const uint32_t cnt = min(n - base, (uint32_t)kBlock);
        for (uint32_t eb = w * kWave; eb < total; eb += kBlock) {
            const uint32_t e = eb + lane_id();
            bool push = false;
            uint32_t v = 0;
            if (e < total) {
                uint32_t k = 0;
#pragma unroll
                for (uint32_t stp = kBlock / 2; stp > 0; stp >>= 1)
                    if (k + stp < cnt && s_pref[k + stp] <= e) k += stp;
                const WEdge ed = a.edges[s_beg[k] + (OffT)(e - s_pref[k])];
                v = ed.dst;
                push = mssp_relax(a, (const uint16_t*)s_rows[k], v, ed.w, s_bits[k], act);
            }
            st.push(push, v, a.q_out, &nxt->n_res);
        }
        __syncthreads();
    }
    st.flush(a.q_out, &nxt->n_res);
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) act |= __shfl_xor(act, d, kWave);
    if (lane_id() == 0 && act) atomicOr(&nxt->pad[0], act);
}

template <class OffT>
__global__ __launch_bounds__(kBlock) void k_mssp_heavy(BatchArgs<OffT, WEdge> a) {
    const uint32_t step = a.sp->step;
    constexpr int kWin = 8;
    __shared__ unsigned long long s_hb[kWin], s_hbits[kWin];
    __shared__ OffT s_hoff[kWin];
    __shared__ uint32_t s_hu[kWin];
    __shared__ uint32_t s_k0;
    __shared__ uint32_t s_stage[kWavesPerBlock][kStageCap];
    Slot* cur = a.slots + (step & 3);
    Slot* nxt = a.slots + ((step + 1) & 3);
    const unsigned long long h = cur->heavy;
    const uint32_t nh = (uint32_t)(h >> kHeavyShift);
    const uint64_t total = h & kHeavyMask;
    if (nh == 0) return;
    const uint32_t tid = threadIdx.x, w = wave_id();
    WaveStage<kStageCap> st{s_stage[w], 0};
    unsigned long long act = 0;
    constexpr uint32_t kSTile = kTile;
    for (uint64_t t0 = (uint64_t)blockIdx.x * kSTile; t0 < total; t0 += (uint64_t)gridDim.x * kSTile) {
        if (tid == 0) {
            uint32_t lo = 0, hi = nh - 1;
            while (lo < hi) {
                uint32_t mid = (lo + hi + 1) >> 1;
                if (a.heavy_base[mid] <= t0) lo = mid; else hi = mid - 1;
            }
            s_k0 = lo;
        }
        __syncthreads();
        const uint32_t k0 = s_k0;
        const uint32_t nwin = min(nh - k0, (uint32_t)kWin);
        if (tid < nwin) {
            s_hb[tid] = a.heavy_base[k0 + tid];
            s_hu[tid] = a.heavy_u[k0 + tid];
            s_hoff[tid] = a.off[a.heavy_u[k0 + tid]];
            s_hbits[tid] = a.heavy_bits[k0 + tid];
        }
        __syncthreads();
        const uint32_t tlen = (uint32_t)min((uint64_t)kSTile, total - t0);
        for (uint32_t r = w * kWave; r < tlen; r += kBlock) {
            const uint32_t rr = r + lane_id();
            bool push = false;
            uint32_t v = 0;
            if (rr < tlen) {
                const uint64_t e = t0 + rr;
                uint32_t k = 0;
#pragma unroll
                for (uint32_t j = 1; j < (uint32_t)kWin; j++)
                    if (j < nwin && s_hb[j] <= e) k = j;
                const WEdge ed = a.edges[s_hoff[k] + (OffT)(e - s_hb[k])];
                v = ed.dst;
                push = mssp_relax(a, a.dist + (size_t)s_hu[k] * a.stride, v, ed.w, s_hbits[k], act);
            }
            st.push(push, v, a.q_out, &nxt->n_res);
        }
        __syncthreads();
    }
    st.flush(a.q_out, &nxt->n_res);
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) act |= __shfl_xor(act, d, kWave);
    if (lane_id() == 0 && act) atomicOr(&nxt->pad[0], act);
}

// -- admission ------------------------------------------------------------------

__global__ void k_clear_seen_bits(unsigned long long* seen, uint64_t V, unsigned long long mask) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
        seen[v] &= ~mask;
}

// Lanes of an SSSP distance row set to all ones (0xFFFF: unreached): `m`
// has the lanes' bytes set, rows are nu4 16-byte units (at most 8); each
// thread ORs whole units of one row.
struct RowMask {
    uint4 m[8];
};

__global__ void k_reset_rows(uint4* rows, uint64_t V, uint32_t nu4, RowMask m) {
    const uint64_t n = V * nu4;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 mk = m.m[i % nu4];
        if (!(mk.x | mk.y | mk.z | mk.w)) continue;
        uint4 x = rows[i];
        x.x |= mk.x;
        x.y |= mk.y;
        x.z |= mk.z;
        x.w |= mk.w;
        rows[i] = x;
    }
}

template <class OffT, class EdgeT>
__global__ void k_seed(BatchArgs<OffT, EdgeT> a, SeedArgs sa, uint32_t step, uint32_t* s0, int bfs) {
    const uint32_t i = threadIdx.x;
    if (i >= sa.n) return;
    const uint32_t q = sa.lane[i], src = sa.src[i];
    const unsigned long long bit = 1ull << q;
    s0[q] = step;
    if (bfs) {
        atomicOr(a.seen + src, bit);
        a.lev[(size_t)src * a.stride + q] = 1;
    } else {
        a.dist[(size_t)src * a.stride + q] = 1;
    }
    atomicAnd(a.ovf, ~bit);
    if (atomicOr(a.X + src, bit) == 0) {
        const uint32_t pos = atomicAdd(&a.slots[step & 3].n_res, 1u);
        ((uint32_t*)a.q_in)[pos] = src;
    }
}

// Result column of one BFS lane, unreached (seen bit clear) as V + 1; the
// level bytes of vertices the lane did not reach may hold an earlier query's.
__global__ void k_extract_lev(const uint8_t* lev, const unsigned long long* seen, uint64_t V, uint32_t stride,
                              uint32_t lane, uint32_t* out) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
        out[v] = ((seen[v] >> lane) & 1ull) ? (uint32_t)lev[v * stride + lane] : (uint32_t)(V + 1);
}

__global__ void k_lane_fill(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Result column of one lane, unreached as V + 1.  The lanes relax without the
// reference's cap: there a value starts at V + 1 and only a smaller candidate
// lowers it (GraphMeta.cuh:440-446, gpu_kernels.cuh:499), so a vertex whose
// distance + 1 reaches V + 1 keeps V + 1.  With non-negative weights that
// schedule ends at min(exact value, V + 1), which is applied here.
__global__ void k_extract_lane(const uint16_t* dist, uint64_t V, uint32_t stride, uint32_t lane, uint32_t* out) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t d = dist[v * stride + lane];
        out[v] = d == kD16Inf || (uint64_t)d > V ? (uint32_t)(V + 1) : d;
    }
}

// -- host -----------------------------------------------------------------------

struct SQuery {
    uint64_t ticket = 0;
    int kind = 0;
    uint32_t src = 0;
    uint32_t* out = nullptr;
    double t_sub = 0, t_adm = 0, t_done = 0;
    uint32_t lane = 0, steps = 0;
    bool done = false;
    bool claimed = false;           // a thread waits on this ticket (asc_wait)
};

struct Engine {
    int kind = 0;
    int lanes = 0;
    uint32_t stride = 0;              // SSSP row length: lanes rounded up to 8
    uint64_t V = 0;
    hipStream_t stream = nullptr;
    unsigned long long* bits[2] = {nullptr, nullptr};
    unsigned long long* seen = nullptr;
    uint32_t* lanes_state = nullptr;  // lev (BFS, u32) or dist (SSSP, u16)
    unsigned long long* d_ovf = nullptr;  // SSSP lanes whose distances left 16 bits
    HostBuf<unsigned long long> h_ovf;
    uint32_t* q[2] = {nullptr, nullptr};
    Slot* slots = nullptr;
    uint32_t* heavy_u = nullptr;
    unsigned long long* heavy_bits = nullptr;
    unsigned long long* heavy_base = nullptr;
    uint32_t* s0 = nullptr;
    uint32_t* staging = nullptr;      // SSSP result column
    HostBuf<Slot> h_slot;
    SQuery* lane_q[kMaxLanes] = {};
    unsigned long long busy = 0;
    uint32_t step = 0;
    uint32_t group_step = 0;          // step at which the current group of lanes was admitted
    bool launched = false;
    // BFS direction per lane: bottom-up lanes, next frontier vertices and
    // edges (from k_lane_counts), unexplored edges
    bool bu_ok = false;
    bool launched_bu = false;         // the last step ran a bottom-up kernel
    unsigned long long bu = 0;
    uint64_t nf[kMaxLanes] = {}, prev_nf[kMaxLanes] = {}, mf[kMaxLanes] = {}, mu[kMaxLanes] = {};
    unsigned long long* d_cnt = nullptr;
    HostBuf<unsigned long long> h_cnt;
    // the step's parameters (device copy written at the head of each step)
    StepParams* d_sp = nullptr;
    HostBuf<StepParams> h_sp;
    // recorded step sequences, one per step parity (the frontier buffers
    // alternate), and what they were recorded with
    bool use_graphs = false;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_grid[2] = {0, 0};
    uint32_t graph_heavy[2] = {0, 0};
    uint64_t graph_steps = 0;         // steps run by replaying a recorded sequence
};

}  // namespace asc

struct asc_server {
    asc_graph* g = nullptr;
    asc_server_opts opts{};
    asc::Engine eng[2];
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<asc::SQuery*> pending[2];
    std::unordered_map<uint64_t, asc::SQuery*> tickets;
    uint64_t next_ticket = 1, submitted = 0, completed = 0, steps = 0, outstanding = 0, bu_steps = 0;
    uint64_t fallbacks = 0;           // SSSP queries answered by the single-query engine
    uint64_t graph_steps = 0;         // engine steps replayed from a recorded hipGraph
    double busy_ms = 0;
    bool stop = false;
    int error = ASC_OK;
};

namespace asc {

static void engine_free(Engine& e) {
    auto f = [](void* p) {
        if (p) (void)hipFree(p);
    };
    // nothing of the engine may still run, and no recorded step may outlive
    // the buffers it names
    if (e.stream) (void)hipStreamSynchronize(e.stream);
    for (hipGraphExec_t& x : e.graph) {
        if (x) (void)hipGraphExecDestroy(x);
        x = nullptr;
    }
    f(e.bits[0]); f(e.bits[1]); f(e.seen); f(e.lanes_state); f(e.q[0]); f(e.q[1]); f(e.slots);
    f(e.heavy_u); f(e.heavy_bits); f(e.heavy_base); f(e.s0); f(e.staging); f(e.d_cnt); f(e.d_ovf); f(e.d_sp);
    if (e.stream) (void)hipStreamDestroy(e.stream);
    e.stream = nullptr;
}

static int engine_init(Engine& e, int kind, int lanes, uint64_t V) {
    e.kind = kind;
    e.lanes = lanes;
    e.V = V;
    if (lanes <= 0) return ASC_OK;
    ASC_HIP(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking));
    for (int i = 0; i < 2; i++) {
        ASC_HIP(hipMalloc(&e.bits[i], 8 * V));
        ASC_HIP(hipMemsetAsync(e.bits[i], 0, 8 * V, e.stream));
        ASC_HIP(hipMalloc(&e.q[i], 4 * V));
    }
    if (kind == ASC_QUERY_BFS) {
        ASC_HIP(hipMalloc(&e.seen, 8 * V));
        ASC_HIP(hipMemsetAsync(e.seen, 0, 8 * V, e.stream));
    }
    ASC_HIP(hipMalloc(&e.staging, 4 * V));
    // rows of whole 16-byte units: 16 lanes for the u8 BFS levels, 8 for the u16 SSSP distances
    e.stride = kind == ASC_QUERY_BFS ? (uint32_t)((lanes + 15) / 16 * 16) : (uint32_t)((lanes + 7) / 8 * 8);
    hipError_t r = hipMalloc(&e.lanes_state, (kind == ASC_QUERY_SSSP ? 2 : 1) * V * (size_t)e.stride);
    if (r != hipSuccess) {
        set_last_error("serving state of %d lanes x %llu vertices does not fit", lanes, (unsigned long long)V);
        return ASC_ERR_NOMEM;
    }
    ASC_HIP(hipMalloc(&e.slots, sizeof(Slot) * 4));
    ASC_HIP(hipMemsetAsync(e.slots, 0, sizeof(Slot) * 4, e.stream));
    ASC_HIP(hipMalloc(&e.heavy_u, 4 * (size_t)kHeavyCap));
    ASC_HIP(hipMalloc(&e.heavy_bits, 8 * (size_t)kHeavyCap));
    ASC_HIP(hipMalloc(&e.heavy_base, 8 * (size_t)kHeavyCap));
    ASC_HIP(hipMalloc(&e.s0, 4 * kMaxLanes));
    ASC_TRY(e.h_slot.alloc(4, true));
    ASC_HIP(hipMalloc(&e.d_sp, sizeof(StepParams)));
    ASC_TRY(e.h_sp.alloc(1, true));
    ASC_HIP(hipMalloc(&e.d_ovf, 8));
    ASC_HIP(hipMemsetAsync(e.d_ovf, 0, 8, e.stream));
    ASC_TRY(e.h_ovf.alloc(1, true));
    e.h_ovf[0] = 0;
    if (kind == ASC_QUERY_BFS) {
        ASC_HIP(hipMalloc(&e.d_cnt, 16 * kMaxLanes));
        ASC_TRY(e.h_cnt.alloc(2 * kMaxLanes, true));
    }
    ASC_HIP(hipStreamSynchronize(e.stream));
    return ASC_OK;
}

template <class OffT, class EdgeT>
static BatchArgs<OffT, EdgeT> engine_args(asc_graph* g, Engine& e) {
    BatchArgs<OffT, EdgeT> a{};
    a.off = (const OffT*)g->d_off;
    a.edges = (const EdgeT*)g->d_edges;
    a.X = e.bits[e.step & 1];
    a.Y = e.bits[(e.step + 1) & 1];
    a.seen = e.seen;
    a.lev = e.kind == ASC_QUERY_BFS ? (uint8_t*)e.lanes_state : nullptr;
    a.dist = e.kind == ASC_QUERY_SSSP ? (uint16_t*)e.lanes_state : nullptr;
    a.ovf = e.d_ovf;
    a.q_in = e.q[e.step & 1];
    a.q_out = e.q[(e.step + 1) & 1];
    a.slots = e.slots;
    a.heavy_u = e.heavy_u;
    a.heavy_bits = e.heavy_bits;
    a.heavy_base = e.heavy_base;
    a.s0 = e.s0;
    a.sp = e.d_sp;
    a.V = e.V;
    a.stride = e.stride;
    a.heavy_thr = (uint32_t)effective_heavy_threshold(g);
    return a;
}

constexpr uint64_t kMsAlpha = 14;  // lane goes bottom-up when its frontier edges > unexplored / kMsAlpha
constexpr uint64_t kMsBeta = 24;   // and back when its frontier shrinks below V / kMsBeta vertices

// Chooses the direction of every BFS lane for the coming step from the
// counts of its frontier (fresh lanes start top-down).
static void choose_directions(asc_graph* g, Engine& e, const std::vector<SQuery*>& admitted) {
    unsigned long long fresh = 0;
    for (SQuery* q : admitted) {
        const uint32_t l = q->lane;
        fresh |= 1ull << l;
        const uint64_t d = g->h.off[q->src + 1] - g->h.off[q->src];
        e.nf[l] = 1;
        e.prev_nf[l] = 0;
        e.mf[l] = d;
        e.mu[l] = g->h.E - d;
    }
    e.bu &= ~fresh;
    if (e.bu_ok) {
        for (int l = 0; l < e.lanes; l++) {
            const unsigned long long bit = 1ull << l;
            if (!(e.busy & bit) || (fresh & bit)) continue;
            if (!(e.bu & bit)) {
                if (e.mf[l] > e.mu[l] / kMsAlpha) e.bu |= bit;
            } else if (e.nf[l] < e.V / kMsBeta && e.nf[l] < e.prev_nf[l]) {
                e.bu &= ~bit;
            }
        }
    }
    e.bu &= e.busy;
}

// The launch sequence of one engine step: the step's parameters are copied
// to the device, the step's kernels run, and the slots, the lane counts (BFS
// with bottom-up lanes) and the overflow word (SSSP) are copied back.  The
// same sequence is launched directly or recorded into a hipGraph; it holds no
// memset node (counters are cleared by a kernel).
template <class OffT, class EdgeT>
static int record_step(asc_graph* g, Engine& e, const BatchArgs<OffT, EdgeT>& a, int grid) {
    (void)g;
    ASC_HIP(hipMemcpyAsync(e.d_sp, e.h_sp.p, sizeof(StepParams), hipMemcpyHostToDevice, e.stream));
    if constexpr (std::is_same<EdgeT, WEdge>::value) {
        if (e.kind == ASC_QUERY_SSSP) {
            hipLaunchKernelGGL((k_mssp_local<OffT>), dim3(grid), dim3(kBlock), 0, e.stream, a);
            hipLaunchKernelGGL((k_mssp_heavy<OffT>), dim3(grid), dim3(kBlock), 0, e.stream, a);
        }
    }
    if (e.kind == ASC_QUERY_BFS) {
        if (e.bu_ok)
            hipLaunchKernelGGL((k_msbfs_bottom_up<OffT, EdgeT>), dim3(grid), dim3(kBlock), 0, e.stream, a);
        hipLaunchKernelGGL((k_msbfs_local<OffT, EdgeT>), dim3(grid), dim3(kBlock), 0, e.stream, a);
        hipLaunchKernelGGL((k_msbfs_heavy<OffT, EdgeT>), dim3(grid), dim3(kBlock), 0, e.stream, a);
        if (e.bu_ok) {
            hipLaunchKernelGGL(k_lane_fill, dim3(1), dim3(256), 0, e.stream, (uint32_t*)e.d_cnt,
                               (uint64_t)(4 * kMaxLanes), 0u);
            hipLaunchKernelGGL((k_lane_counts<OffT>), dim3(grid), dim3(kBlock), 0, e.stream, a.off, a.Y, a.q_out,
                               e.slots, e.d_sp, e.d_cnt);
            ASC_HIP(hipMemcpyAsync(e.h_cnt.p, e.d_cnt, 16 * kMaxLanes, hipMemcpyDeviceToHost, e.stream));
        }
    }
    ASC_HIP(hipGetLastError());
    ASC_HIP(hipMemcpyAsync(e.h_slot.p, e.slots, 4 * sizeof(Slot), hipMemcpyDeviceToHost, e.stream));
    if (e.d_ovf) ASC_HIP(hipMemcpyAsync(e.h_ovf.p, e.d_ovf, 8, hipMemcpyDeviceToHost, e.stream));
    return ASC_OK;
}

// Resets and seeds newly admitted lanes, then launches one step.
template <class OffT, class EdgeT>
static int engine_launch(asc_graph* g, Engine& e, const std::vector<SQuery*>& admitted) {
    const int grid = g->ctx->num_cus * (g->blocks_per_cu > 0 ? g->blocks_per_cu : kBlocksPerCU);
    if (e.kind == ASC_QUERY_BFS) choose_directions(g, e, admitted);
    auto a = engine_args<OffT, EdgeT>(g, e);
    if (!admitted.empty()) {
        SeedArgs sa{};
        unsigned long long mask = 0;
        for (SQuery* q : admitted) {
            sa.lane[sa.n] = q->lane;
            sa.src[sa.n] = q->src;
            sa.n++;
            mask |= 1ull << q->lane;
        }
        if (e.kind == ASC_QUERY_BFS) {
            // unreached everywhere; the level bytes are read only where a seen
            // bit is set, so they keep whatever an earlier query left
            hipLaunchKernelGGL(k_clear_seen_bits, dim3(grid), dim3(256), 0, e.stream, e.seen, e.V, mask);
        } else {
            // the admitted lanes' distances: every vertex unreached
            const uint32_t lane_bytes = 2;
            const uint32_t row_bytes = e.stride * lane_bytes;
            RowMask rm{};
            uint8_t* mb = (uint8_t*)rm.m;
            for (unsigned long long m = mask; m; m &= m - 1) {
                const uint32_t q = (uint32_t)__builtin_ctzll(m);
                for (uint32_t b = 0; b < lane_bytes; b++) mb[q * lane_bytes + b] = 0xFF;
            }
            hipLaunchKernelGGL(k_reset_rows, dim3(grid), dim3(256), 0, e.stream, (uint4*)e.lanes_state, e.V,
                               row_bytes / 16, rm);
        }
        hipLaunchKernelGGL((k_seed<OffT, EdgeT>), dim3(1), dim3(kMaxLanes), 0, e.stream, a, sa, e.step, e.s0,
                           e.kind == ASC_QUERY_BFS ? 1 : 0);
        ASC_HIP(hipGetLastError());
    }
    e.h_sp[0].step = e.step;
    e.h_sp[0].pad = 0;
    e.h_sp[0].bu = e.kind == ASC_QUERY_BFS ? e.bu : 0;
    if (e.kind == ASC_QUERY_BFS) e.launched_bu = e.bu != 0;
    if (!e.use_graphs) {
        const int rr = record_step<OffT, EdgeT>(g, e, a, grid);
        if (rr != ASC_OK) return rr;
    } else {
        const int par = (int)(e.step & 1);
        if (!e.graph[par] || e.graph_grid[par] != grid || e.graph_heavy[par] != a.heavy_thr) {
            // record the step's launch sequence once per parity (again when
            // the launch width or the heavy threshold was retuned)
            if (e.graph[par]) ASC_HIP(hipGraphExecDestroy(e.graph[par]));
            e.graph[par] = nullptr;
            hipGraph_t gr = nullptr;
            ASC_HIP(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
            const int rr = record_step<OffT, EdgeT>(g, e, a, grid);
            const hipError_t ec = hipStreamEndCapture(e.stream, &gr);
            if (rr != ASC_OK) {
                if (gr) (void)hipGraphDestroy(gr);
                return rr;
            }
            ASC_HIP(ec);
            const hipError_t ei = hipGraphInstantiate(&e.graph[par], gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            ASC_HIP(ei);
            e.graph_grid[par] = grid;
            e.graph_heavy[par] = a.heavy_thr;
        }
        ASC_HIP(hipGraphLaunch(e.graph[par], e.stream));
        e.graph_steps++;
    }
    e.launched = true;
    return ASC_OK;
}

// After a step: finished lanes copy their results out and are released.
static int engine_finish(asc_server* s, Engine& e, std::vector<SQuery*>& done) {
    if (!e.launched) return ASC_OK;
    ASC_HIP(hipStreamSynchronize(e.stream));
    e.launched = false;
    const unsigned long long act = e.h_slot[(e.step + 1) & 3].pad[0];
    if (e.kind == ASC_QUERY_BFS && e.bu_ok) {
        for (int q = 0; q < e.lanes; q++) {
            e.prev_nf[q] = e.nf[q];
            e.nf[q] = e.h_cnt[q];
            e.mf[q] = e.h_cnt[kMaxLanes + q];
            e.mu[q] = e.mu[q] > e.mf[q] ? e.mu[q] - e.mf[q] : 0;
        }
    }
    const bool zero_based = s->g->ctx->cfg.value_convention == ASC_VALUES_ZERO_BASED;
    const uint32_t unreached = (uint32_t)(e.V + 1);
    for (int q = 0; q < e.lanes; q++) {
        SQuery* qq = e.lane_q[q];
        if (!qq) continue;
        qq->steps++;
        if (act & (1ull << q)) continue;
        const bool overflowed = (e.h_ovf[0] >> q) & 1ull;
        if (qq->out && overflowed) {
            // a level left 8 bits / a distance left 16 bits: the single-query
            // engine answers exactly (value convention applied there)
            ASC_TRY(traverse(s->g, e.kind == ASC_QUERY_SSSP ? ALG_SSSP : ALG_BFS, qq->src, qq->out, nullptr));
            s->fallbacks++;
        } else if (qq->out) {
            const int grid = s->g->ctx->num_cus * 4;
            if (e.kind == ASC_QUERY_BFS)
                hipLaunchKernelGGL(k_extract_lev, dim3(grid), dim3(256), 0, e.stream, (const uint8_t*)e.lanes_state,
                                   (const unsigned long long*)e.seen, e.V, e.stride, (uint32_t)q, e.staging);
            else
                hipLaunchKernelGGL(k_extract_lane, dim3(grid), dim3(256), 0, e.stream,
                                   (const uint16_t*)e.lanes_state, e.V, e.stride, (uint32_t)q, e.staging);
            ASC_HIP(hipMemcpyAsync(qq->out, e.staging, 4 * e.V, hipMemcpyDeviceToHost, e.stream));
            ASC_HIP(hipStreamSynchronize(e.stream));
            if (zero_based) {
                uint32_t* o = qq->out;
                host_pool().parallel_for(e.V, [&](uint64_t b, uint64_t en) {
                    for (uint64_t v = b; v < en; v++) o[v] = o[v] == unreached ? 0xFFFFFFFFu : o[v] - 1;
                });
            }
        }
        qq->t_done = now_ms();
        e.lane_q[q] = nullptr;
        e.busy &= ~(1ull << q);
        done.push_back(qq);
    }
    e.step++;
    return ASC_OK;
}

// Lanes admitted together stay in phase: their frontiers peak in the same
// steps and share the expansions of the vertices they reach together, while
// lanes admitted one by one keep some lane at its peak in every step (the
// union frontier then spans most of the graph each step).  So while an engine
// is busy, freed BFS lanes are refilled only as a group of at least half the
// lanes, and SSSP lanes (whose peaks last longer) only once every lane is
// free; an idle engine admits at once, and queries arriving within the
// first kJoinSteps steps of a group join it (they are still nearly in phase;
// without this a query that finds the engine idle runs alone while the
// queries submitted right after it wait for it: 84 vs 90 SSSP queries/s).  Measured on R-MAT scale 26 at
// saturation against admitting at every step boundary: BFS 979 vs 795
// queries/s, SSSP 59 vs 48.
constexpr uint32_t kJoinSteps = 2;

static bool admit_now(const asc_server* s, const Engine& e) {
    if (!e.busy || s->pending[e.kind].empty()) return true;
    if (e.step - e.group_step < kJoinSteps) return true;   // the group is still in its first steps
    const int free_lanes = e.lanes - __builtin_popcountll(e.busy);
    if (e.kind == ASC_QUERY_SSSP) return false;
    return free_lanes >= std::max(1, e.lanes / 2);
}

template <class OffT, class EdgeT>
static void server_loop(asc_server* s) {
    (void)hipSetDevice(s->g->ctx->device);
    std::vector<SQuery*> admitted[2], done;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] {
                return s->stop || !s->pending[0].empty() || !s->pending[1].empty() || s->eng[0].busy ||
                       s->eng[1].busy;
            });
            if (s->stop && !s->eng[0].busy && !s->eng[1].busy && s->pending[0].empty() && s->pending[1].empty())
                break;
            const double now = now_ms();
            for (int k = 0; k < 2; k++) {
                Engine& e = s->eng[k];
                admitted[k].clear();
                if (!admit_now(s, e)) continue;
                if (!e.busy && !s->pending[k].empty()) e.group_step = e.step;   // a new group starts
                for (int q = 0; q < e.lanes && !s->pending[k].empty(); q++) {
                    if (e.busy & (1ull << q)) continue;
                    SQuery* qq = s->pending[k].front();
                    s->pending[k].pop_front();
                    qq->lane = (uint32_t)q;
                    qq->t_adm = now;
                    e.lane_q[q] = qq;
                    e.busy |= 1ull << q;
                    admitted[k].push_back(qq);
                }
            }
        }
        const double t_step = now_ms();
        int r = ASC_OK;
        for (int k = 0; k < 2 && r == ASC_OK; k++)
            if (s->eng[k].busy) {
                if (k == ASC_QUERY_SSSP) {
                    if constexpr (std::is_same<EdgeT, WEdge>::value)
                        r = engine_launch<OffT, EdgeT>(s->g, s->eng[k], admitted[k]);
                } else {
                    r = engine_launch<OffT, EdgeT>(s->g, s->eng[k], admitted[k]);
                }
            }
        done.clear();
        for (int k = 0; k < 2 && r == ASC_OK; k++) r = engine_finish(s, s->eng[k], done);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->busy_ms += now_ms() - t_step;
            s->steps++;
            if (s->eng[ASC_QUERY_BFS].launched_bu) s->bu_steps++;
            s->graph_steps = s->eng[0].graph_steps + s->eng[1].graph_steps;
            if (r != ASC_OK) {
                // a device error fails every query in flight or waiting
                s->error = r;
                for (int k = 0; k < 2; k++) {
                    for (int q = 0; q < s->eng[k].lanes; q++)
                        if (s->eng[k].lane_q[q]) done.push_back(s->eng[k].lane_q[q]), s->eng[k].lane_q[q] = nullptr;
                    s->eng[k].busy = 0;
                    for (SQuery* qq : s->pending[k]) done.push_back(qq);
                    s->pending[k].clear();
                }
            }
            for (SQuery* qq : done) {
                qq->done = true;
                s->completed++;
                s->outstanding--;
            }
        }
        if (!done.empty()) s->cv_done.notify_all();
    }
}

}  // namespace asc

using namespace asc;

extern "C" {

int asc_server_create(asc_graph* g, const asc_server_opts* opts, asc_server** out) {
    if (!g || !out) return ASC_ERR_INVALID;
    *out = nullptr;
    ASC_TRY(graph_ready(g));
    if (g->h.kind == ASC_KIND_CSC) return ASC_ERR_UNSUPPORTED;
    if (g->static_vertices < g->h.V) {
        set_last_error("serving needs a fully resident graph");
        return ASC_ERR_UNSUPPORTED;
    }
    asc_server_opts o;
    if (opts) o = *opts; else asc_server_opts_init(&o);
    if (o.lanes <= 0) o.lanes = g->serve_lanes > 0 ? g->serve_lanes : kMaxLanes;  // tuned batch width
    if (o.lanes > kMaxLanes) o.lanes = kMaxLanes;
    if (o.queue_capacity <= 0) o.queue_capacity = 4096;
    ASC_HIP(hipSetDevice(g->ctx->device));
    const uint64_t V = g->h.V;
    size_t free_b = 0, total_b = 0;
    ASC_HIP(hipMemGetInfo(&free_b, &total_b));
    // fixed state per engine: two frontier words, seen word, queues, heavy list
    const uint64_t fixed = 8 * V * 3 + 4 * V * 2 + 16 * (uint64_t)kHeavyCap + 4 * V;
    uint64_t budget = o.state_budget_bytes ? o.state_budget_bytes : free_b / 2;
    budget = std::min<uint64_t>(budget, free_b > 2 * fixed ? free_b - 2 * fixed : 0);
    const bool weighted = g->h.kind == ASC_KIND_WCSR;
    // per query: u8 levels (BFS), u16 distances (SSSP)
    const uint64_t per_lane = weighted ? 3 * V : V;
    int lanes = (int)std::min<uint64_t>((uint64_t)o.lanes, budget / per_lane);
    if (lanes < 1) {
        set_last_error("state budget %llu bytes admits no query (%llu bytes per lane)", (unsigned long long)budget,
                       (unsigned long long)per_lane);
        return ASC_ERR_NOMEM;
    }
    asc_server* s = new asc_server();
    s->g = g;
    s->opts = o;
    int r = engine_init(s->eng[ASC_QUERY_BFS], ASC_QUERY_BFS, lanes, V);
    s->eng[ASC_QUERY_BFS].bu_ok = o.bfs_direction == ASC_BFS_AUTO && resident_symmetric(g);
    s->eng[0].use_graphs = s->eng[1].use_graphs = o.step_graphs != 0;
    if (r == ASC_OK) r = engine_init(s->eng[ASC_QUERY_SSSP], ASC_QUERY_SSSP, weighted ? lanes : 0, V);
    if (r != ASC_OK) {
        engine_free(s->eng[0]);
        engine_free(s->eng[1]);
        delete s;
        return r;
    }
    const bool w = weighted, o4 = g->off_bytes == 4;
    g->servers++;
    if (w) s->th = std::thread(o4 ? server_loop<uint32_t, WEdge> : server_loop<uint64_t, WEdge>, s);
    else s->th = std::thread(o4 ? server_loop<uint32_t, uint32_t> : server_loop<uint64_t, uint32_t>, s);
    *out = s;
    return ASC_OK;
}

int asc_server_destroy(asc_server* s) {
    if (!s) return ASC_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    engine_free(s->eng[0]);
    engine_free(s->eng[1]);
    // the ticket map owns every query not yet waited for (pending ones too;
    // the loop drains the queues before it ends)
    for (auto& kv : s->tickets) delete kv.second;
    s->g->servers--;
    delete s;
    return ASC_OK;
}

int asc_submit(asc_server* s, int kind, uint32_t source, uint32_t* out, uint64_t* ticket) {
    if (!s || !ticket || (kind != ASC_QUERY_BFS && kind != ASC_QUERY_SSSP)) return ASC_ERR_INVALID;
    if (source >= s->g->h.V) return ASC_ERR_INVALID;
    if (s->eng[kind].lanes == 0) {
        set_last_error("SSSP queries need a weighted graph");
        return ASC_ERR_UNSUPPORTED;
    }
    SQuery* q = new SQuery();
    q->kind = kind;
    q->src = source;
    q->out = out;
    q->t_sub = now_ms();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->error != ASC_OK) {
            delete q;
            return s->error;
        }
        if (s->outstanding >= (uint64_t)s->opts.queue_capacity) {
            delete q;
            return ASC_ERR_BUSY;
        }
        q->ticket = s->next_ticket++;
        s->tickets[q->ticket] = q;
        s->pending[kind].push_back(q);
        s->submitted++;
        s->outstanding++;
    }
    s->cv_work.notify_one();
    *ticket = q->ticket;
    return ASC_OK;
}

int asc_wait(asc_server* s, uint64_t ticket, asc_qstats* qs) {
    if (!s) return ASC_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return ASC_ERR_INVALID;
    SQuery* q = it->second;
    // one waiter per ticket: a second thread waiting on the same ticket would
    // find the query freed by the first
    if (q->claimed) {
        set_last_error("ticket %llu is already being waited for", (unsigned long long)ticket);
        return ASC_ERR_INVALID;
    }
    q->claimed = true;
    s->cv_done.wait(lk, [&] { return q->done; });
    if (qs) {
        qs->submit_to_admit_ms = q->t_adm - q->t_sub;
        qs->submit_to_done_ms = q->t_done - q->t_sub;
        qs->iterations = q->steps;
        qs->lane = q->lane;
    }
    const int err = q->t_done > 0 ? ASC_OK : s->error;
    s->tickets.erase(ticket);   // `it` may be stale: the map can rehash while this thread waits
    delete q;
    return err;
}

int asc_server_drain(asc_server* s) {
    if (!s) return ASC_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
    return s->error;
}

int asc_server_get_stats(asc_server* s, asc_server_stats* st) {
    if (!s || !st) return ASC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->steps = s->steps;
    st->lanes_bfs = s->eng[ASC_QUERY_BFS].lanes;
    st->lanes_sssp = s->eng[ASC_QUERY_SSSP].lanes;
    st->busy_ms = s->busy_ms;
    st->bottom_up_steps = s->bu_steps;
    st->fallbacks = s->fallbacks;
    st->graph_steps = s->graph_steps;
    return ASC_OK;
}

}  // extern "C"
