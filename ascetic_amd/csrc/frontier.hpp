// Frontier relax kernels shared by the traversal engine (traverse.hip) and the
// resident fast paths (resident.hip): check-then-atomicMin relaxation with
// routing of improved vertices, and the edge-balanced expansion of a resident
// frontier queue (k_expand_local) and of its heavy vertices (k_expand_heavy).
#pragma once

#include "device_common.hpp"
#include "traverse.hpp"

namespace asc {

enum { TO_NONE = 0, TO_RES = 1, TO_PEND = 2, TO_FAR = 3 };

// SSSP: value + weight, saturating at 0xFFFFFFFF (one clamped add), so a
// weight near 2^32 cannot wrap to a small candidate; every saturated candidate
// is above V + 1 and lowers nothing, which leaves min(exact, V + 1) as the
// final value.  BFS passes level + 2 as sv, CC the label.
template <int ALGO>
__device__ __forceinline__ uint32_t candidate(uint32_t sv, uint32_t w) {
    return ALGO == ALG_SSSP ? __builtin_elementwise_add_sat(sv, w) : sv;
}

// Check-then-atomicMin relaxation (gpu_kernels.cuh:390-392, 499-502, 453-456)
// and routing of an improved vertex, in two steps so that a wave can issue
// the probes of several edges before any of their atomics: relax_probe loads
// the destination's word (BFS: visited bits; SSSP: its value filter; CC:
// value),
// relax_commit_batch decides from it and performs the atomics.
template <int ALGO, class OffT, class EdgeT>
__device__ __forceinline__ uint32_t relax_probe(const TravArgs<OffT, EdgeT>& a, uint32_t v) {
    if (ALGO == ALG_BFS) return a.visited[v >> 5];
    if (ALGO == ALG_SSSP && a.val16) {
        // 0xFFFF: nothing known (not lowered yet, or a value beyond 16 bits),
        // so the atomic is tried
        const uint32_t f = a.val16[v];
        return f == 0xFFFFu ? 0xFFFFFFFFu : f;
    }
    return a.val[v];
}

// Edges each lane takes per pass of an edge loop: the loads of kUnroll edges,
// then their probes, are in flight together.
constexpr int kUnroll = 4;

// The commit of a lane's kUnroll edges (edge j counts when bit j of `valid`
// is set).  BFS: a vertex whose visited bit is clear is claimed by an atomic
// OR of that bit and gets its level; SSSP/CC: a candidate below the probed
// value is offered by atomicMin, and a vertex it lowers is routed once per
// step (a far candidate to the far pile, a static vertex to the resident
// frontier, an on-demand one to the pile) by an atomicMax on the matching
// stamp.  The value atomics of all edges are issued before any result is
// used, then the stamp atomics of those that improved, so each lane waits for
// two atomic round trips per pass instead of two per edge.
template <int ALGO, class OffT, class EdgeT>
__device__ __forceinline__ void relax_commit_batch(const TravArgs<OffT, EdgeT>& a, const uint32_t (&v)[kUnroll],
                                                   const uint32_t (&cand)[kUnroll], const uint32_t (&probe)[kUnroll],
                                                   uint32_t valid, uint32_t level, int (&to)[kUnroll]) {
    uint32_t old[kUnroll];
    bool try_[kUnroll];
#pragma unroll
    for (int j = 0; j < kUnroll; j++) {
        to[j] = TO_NONE;
        if (ALGO == ALG_BFS) try_[j] = ((valid >> j) & 1u) && !(probe[j] & (1u << (v[j] & 31)));
        else try_[j] = ((valid >> j) & 1u) && cand[j] < probe[j];
        if (!try_[j]) continue;
        if (ALGO == ALG_BFS) old[j] = atomicOr(a.visited + (v[j] >> 5), 1u << (v[j] & 31));
        else old[j] = atomicMin(a.val + v[j], cand[j]);
    }
    if (ALGO == ALG_BFS) {
#pragma unroll
        for (int j = 0; j < kUnroll; j++) {
            if (!try_[j] || (old[j] & (1u << (v[j] & 31)))) continue;
            a.val[v[j]] = cand[j];
            to[j] = v[j] < a.vs ? TO_RES : TO_PEND;
        }
        return;
    }
    uint32_t* sp[kUnroll];
    uint32_t sv[kUnroll], prev[kUnroll];
    int kind[kUnroll];
#pragma unroll
    for (int j = 0; j < kUnroll; j++) {
        kind[j] = TO_NONE;
        if (!try_[j] || old[j] <= cand[j]) continue;
        if (ALGO == ALG_SSSP && a.val16 && cand[j] < 0xFFFFu) a.val16[v[j]] = (uint16_t)cand[j];
        if (ALGO == ALG_SSSP && cand[j] >= a.theta) {
            sp[j] = a.fstamp + v[j];
            sv[j] = a.fepoch;
            kind[j] = TO_FAR;
        } else if (v[j] < a.vs) {
            sp[j] = a.stamp + v[j];
            sv[j] = level + 1;
            kind[j] = TO_RES;
        } else {
            sp[j] = a.pstamp + v[j];
            sv[j] = a.pepoch;
            kind[j] = TO_PEND;
        }
        prev[j] = atomicMax(sp[j], sv[j]);
    }
#pragma unroll
    for (int j = 0; j < kUnroll; j++)
        if (kind[j] != TO_NONE && prev[j] < sv[j]) to[j] = kind[j];
}

// Summed degrees of the vertices a BFS commit claimed (to == TO_RES).
template <class OffT, class EdgeT>
__device__ __forceinline__ unsigned long long claimed_degrees(const TravArgs<OffT, EdgeT>& a,
                                                              const int (&to)[kUnroll],
                                                              const uint32_t (&v)[kUnroll]) {
    unsigned long long m = 0;
#pragma unroll
    for (int j = 0; j < kUnroll; j++)
        if (to[j] == TO_RES) m += (unsigned long long)(a.off[v[j] + 1] - a.off[v[j]]);
    return m;
}

// One atomic per wave that claimed anything.
__device__ __forceinline__ void add_claimed(unsigned long long* dst, unsigned long long m) {
    m = wave_sum(m);
    if (lane_id() == 0 && m) atomicAdd(dst, m);
}

template <int ALGO>
struct Pushes {
    WaveStage<kStageCap> res, pend, far;

    template <class A>
    __device__ __forceinline__ void push(const A& a, Slot* nxt, int to, uint32_t v) {
        res.push(to == TO_RES, v, a.q_out, &nxt->n_res);
        pend.push(to == TO_PEND, v, a.pend_out, a.pend_count);
        if (ALGO == ALG_SSSP) far.push(to == TO_FAR, v, a.far_out, a.far_count);
    }
    template <class A>
    __device__ __forceinline__ void flush(const A& a, Slot* nxt) {
        res.flush(a.q_out, &nxt->n_res);
        pend.flush(a.pend_out, a.pend_count);
        if (ALGO == ALG_SSSP) far.flush(a.far_out, a.far_count);
    }
};

#define ASC_STAGE_DECL                                              \
    __shared__ uint32_t s_stage[3][kWavesPerBlock][kStageCap];      \
    Pushes<ALGO> pq{{s_stage[0][wave_id()], 0}, {s_stage[1][wave_id()], 0}, {s_stage[2][wave_id()], 0}}

// Expansion of the resident frontier (slot level % 4) by workgroup `bid` of
// `nblk`: 256 frontier vertices at a time, their edges walked edge-balanced;
// vertices with degree >= heavy_thr go to the heavy list instead.  Called by
// k_expand_local and by persistent kernels (with their own block numbering).
template <int ALGO, class OffT, class EdgeT>
__device__ __forceinline__ void expand_local_body(const TravArgs<OffT, EdgeT>& a, uint32_t level, uint32_t bid,
                                                  uint32_t nblk) {
    __shared__ uint32_t s_pref[kBlock];
    __shared__ OffT s_beg[kBlock];
    __shared__ uint32_t s_sv[kBlock];
    __shared__ uint32_t s_tmp[kWavesPerBlock];
    ASC_STAGE_DECL;

    Slot* cur = a.slots + (level & 3);
    Slot* nxt = a.slots + ((level + 1) & 3);
    if (bid == 0 && threadIdx.x == 0) {
        Slot* z = a.slots + ((level + 2) & 3);
        z->n_res = 0;
        z->heavy = 0;
    }
    const uint32_t n = cur->n_res;
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
    unsigned long long my_edges = 0, my_verts = 0, m_claimed = 0;

    for (uint32_t base = bid * kBlock; base < n; base += nblk * kBlock) {
        const uint32_t i = base + tid;
        uint32_t deg = 0, sv = 0;
        OffT beg = 0;
        if (i < n) {
            const uint32_t u = a.q_in[i];
            beg = a.off[u];
            uint64_t d = (uint64_t)(a.off[u + 1] - beg);
            if (ALGO == ALG_BFS) {
                sv = level + 2;
            } else {
                sv = a.val[u];
                if (a.done[u] == sv) d = 0;  // already expanded at this value
                else a.done[u] = sv;
            }
            if (d >= a.heavy_thr) {
                const unsigned long long p = atomicAdd(&cur->heavy, (1ull << kHeavyShift) | d);
                const uint32_t idx = (uint32_t)(p >> kHeavyShift);
                a.heavy_u[idx] = u;
                a.heavy_sv[idx] = sv;
                a.heavy_base[idx] = p & kHeavyMask;
            } else {
                deg = (uint32_t)d;
            }
        }
        uint32_t total;
        const uint32_t pre = block_scan_excl(deg, s_tmp, total);
        s_pref[tid] = pre;
        s_beg[tid] = beg;
        s_sv[tid] = sv;
        __syncthreads();
        const uint32_t cnt = min(n - base, (uint32_t)kBlock);
        for (uint32_t eb = w * kWave * kUnroll; eb < total; eb += kBlock * kUnroll) {
            uint32_t kk[kUnroll], vv[kUnroll], cc[kUnroll], pr[kUnroll];
            EdgeT ed[kUnroll];
            // lanes past the end load the last edge again (no branch around
            // the loads, so all of them are in flight together)
#pragma unroll
            for (int j = 0; j < kUnroll; j++) {
                const uint32_t e = min(eb + j * kWave + lane, total - 1);
                uint32_t k = 0;
#pragma unroll
                for (uint32_t step = kBlock / 2; step > 0; step >>= 1)
                    if (k + step < cnt && s_pref[k + step] <= e) k += step;
                ed[j] = a.edges[s_beg[k] + (OffT)(e - s_pref[k])];
                kk[j] = k;
            }
#pragma unroll
            for (int j = 0; j < kUnroll; j++) {
                vv[j] = edge_dst(ed[j]);
                cc[j] = candidate<ALGO>(s_sv[kk[j]], edge_w(ed[j]));
                pr[j] = relax_probe<ALGO>(a, vv[j]);
            }
            uint32_t valid = 0;
#pragma unroll
            for (int j = 0; j < kUnroll; j++)
                if (eb + j * kWave + lane < total) valid |= 1u << j;
            int to[kUnroll];
            relax_commit_batch<ALGO>(a, vv, cc, pr, valid, level, to);
#pragma unroll
            for (int j = 0; j < kUnroll; j++) pq.push(a, nxt, to[j], vv[j]);
            if (ALGO == ALG_BFS && a.m_next) m_claimed += claimed_degrees(a, to, vv);
        }
        my_edges += total;
        my_verts += cnt;
        __syncthreads();
    }
    pq.flush(a, nxt);
    if (ALGO == ALG_BFS && a.m_next) add_claimed(a.m_next, m_claimed);
    if (tid == 0 && my_verts) {
        atomicAdd(&a.stats->edges_local, my_edges);
        atomicAdd(&a.stats->verts_local, my_verts);
    }
}

template <int ALGO, class OffT, class EdgeT>
__global__ __launch_bounds__(kBlock) void k_expand_local(TravArgs<OffT, EdgeT> a, uint32_t level) {
    expand_local_body<ALGO>(a, level, blockIdx.x, gridDim.x);
}

// Grid-cooperative expansion of the heavy list of slot level % 4 by
// workgroup `bid` of `nblk`.
template <int ALGO, class OffT, class EdgeT>
__device__ __forceinline__ void expand_heavy_body(const TravArgs<OffT, EdgeT>& a, uint32_t level, uint32_t bid,
                                                  uint32_t nblk) {
    constexpr int kWin = 8;  // a tile spans <= kTile / kMinHeavy + 1 heavy entries
    __shared__ unsigned long long s_hb[kWin];
    __shared__ OffT s_hoff[kWin];
    __shared__ uint32_t s_hsv[kWin];
    __shared__ uint32_t s_k0;
    ASC_STAGE_DECL;

    Slot* cur = a.slots + (level & 3);
    Slot* nxt = a.slots + ((level + 1) & 3);
    const unsigned long long h = cur->heavy;
    const uint32_t nh = (uint32_t)(h >> kHeavyShift);
    const uint64_t total = h & kHeavyMask;
    if (nh == 0) return;
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
    unsigned long long my_edges = 0, m_claimed = 0;

    // contiguous runs of tiles per workgroup: only the first tile searches
    // the heavy list; later tiles start at the previous tile's last owner
    const uint64_t n_tiles = (total + kTile - 1) / kTile;
    const uint64_t per = (n_tiles + nblk - 1) / nblk;
    const uint64_t tb = min(n_tiles, (uint64_t)bid * per), te = min(n_tiles, tb + per);
    uint32_t k_last = 0;
    for (uint64_t t = tb; t < te; t++) {
        const uint64_t t0 = t * kTile;
        const uint32_t tlen = (uint32_t)min((uint64_t)kTile, total - t0);
        if (tid == 0) {
            uint32_t k0;
            if (t == tb) {
                uint32_t lo = 0, hi = nh - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (a.heavy_base[mid] <= t0) lo = mid; else hi = mid - 1;
                }
                k0 = lo;
            } else {
                k0 = k_last + (k_last + 1 < nh && a.heavy_base[k_last + 1] <= t0 ? 1u : 0u);
            }
            s_k0 = k0;
        }
        __syncthreads();
        const uint32_t k0 = s_k0;
        const uint32_t nwin = min(nh - k0, (uint32_t)kWin);
        if (tid < nwin) {
            s_hb[tid] = a.heavy_base[k0 + tid];
            s_hoff[tid] = a.off[a.heavy_u[k0 + tid]];
            s_hsv[tid] = a.heavy_sv[k0 + tid];
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t j = 0;
            for (uint32_t i = 1; i < nwin; i++)
                if (s_hb[i] <= t0 + tlen - 1) j = i;
            k_last = k0 + j;
        }
        for (uint32_t r = w * kWave * kUnroll; r < tlen; r += kBlock * kUnroll) {
            uint32_t kk[kUnroll], vv[kUnroll], cc[kUnroll], pr[kUnroll];
            EdgeT ed[kUnroll];
#pragma unroll
            for (int j = 0; j < kUnroll; j++) {
                const uint64_t e = t0 + min(r + j * kWave + lane, tlen - 1);
                uint32_t k = 0;
#pragma unroll
                for (uint32_t i = 1; i < (uint32_t)kWin; i++)
                    if (i < nwin && s_hb[i] <= e) k = i;
                ed[j] = a.edges[s_hoff[k] + (OffT)(e - s_hb[k])];
                kk[j] = k;
            }
#pragma unroll
            for (int j = 0; j < kUnroll; j++) {
                vv[j] = edge_dst(ed[j]);
                cc[j] = candidate<ALGO>(s_hsv[kk[j]], edge_w(ed[j]));
                pr[j] = relax_probe<ALGO>(a, vv[j]);
            }
            uint32_t valid = 0;
#pragma unroll
            for (int j = 0; j < kUnroll; j++)
                if (r + j * kWave + lane < tlen) valid |= 1u << j;
            int to[kUnroll];
            relax_commit_batch<ALGO>(a, vv, cc, pr, valid, level, to);
#pragma unroll
            for (int j = 0; j < kUnroll; j++) pq.push(a, nxt, to[j], vv[j]);
            if (ALGO == ALG_BFS && a.m_next) m_claimed += claimed_degrees(a, to, vv);
        }
        my_edges += tlen;
        __syncthreads();
    }
    pq.flush(a, nxt);
    if (ALGO == ALG_BFS && a.m_next) add_claimed(a.m_next, m_claimed);
    if (tid == 0 && my_edges) atomicAdd(&a.stats->edges_heavy, my_edges);
}

template <int ALGO, class OffT, class EdgeT>
__global__ __launch_bounds__(kBlock) void k_expand_heavy(TravArgs<OffT, EdgeT> a, uint32_t level) {
    expand_heavy_body<ALGO>(a, level, blockIdx.x, gridDim.x);
}

}  // namespace asc
