// Frontier traversal on gfx950: BFS, SSSP and CC (min-label propagation).
//
// Reference path: CalculateOpt.cu:14-545 (bfs_opt / cc_opt / sssp_opt), the
// relax kernels gpu_kernels.cuh:345-532 and the frontier bookkeeping
// gpu_kernels.cu:111-122, gpu_kernels.cuh:399-438 plus Thrust scans.
//
// The reference classifies the active flags into a static and an overload
// list with several O(V) passes and host syncs per iteration, and ships the
// overload adjacency of every active vertex every iteration.  Here the relax
// kernels append improved vertices to device queues through per-wave LDS
// staging, so a step costs O(frontier edges) and one host sync:
//   k_expand_local  one workgroup per 256 frontier vertices; degrees below the
//                   heavy threshold are expanded edge-balanced inside the
//                   workgroup (LDS prefix scan + per-lane binary search), so
//                   consecutive lanes read consecutive edges of a vertex;
//                   heavier vertices go to a grid-wide list.
//   k_expand_heavy  grid-cooperative, edge-balanced over the heavy list.
//   k_expand_part   edge-balanced over one streamed part of on-demand edges
//                   (or over mapped pinned host memory in zero-copy mode).
// Improved vertices are routed to one of three queues:
//   resident frontier  v < static_vertices, expanded next step on the device;
//   on-demand pile     other vertices, deduplicated, streamed in one batch;
//   far pile           SSSP only: tentative distance >= theta (the near/far
//                      band boundary), revisited when theta advances.
// BFS stays level-synchronous: each level expands the resident frontier and
// streams the pile filled by the previous level, so every reached on-demand
// vertex crosses the link once.  SSSP and CC relax the resident frontier to
// exhaustion before streaming the pile, and SSSP advances theta by a bucket
// width only when both are empty, so on-demand vertices are mostly streamed
// once their value is final.  A vertex already expanded at its current value
// (`done`) is skipped.  Final values are unique, so the schedule does not
// change results (the reference's own schedule is chaotic, CalculateOpt.cu:
// 481-523).
#include <algorithm>
#include <cmath>

#include "device_common.hpp"
#include "internal.hpp"
#include "traverse.hpp"
#include "frontier.hpp"

namespace asc {

// ---------------------------------------------------------------------------
// device code
// ---------------------------------------------------------------------------

// Edge-balanced relax over edges [lo, hi) of the concatenated adjacency of
// the streamed list (list[k] owns [prefix[k], prefix[k+1]), no zero-degree
// entries).  SSSP/CC expand each entry with its value when its tile starts,
// BFS with the level.  SRC = PART_GATHERED: edges come from the device part
// buffer holding [lo, hi) of the list's concatenation.  Otherwise they are
// addressed by offset from p.tail_base: SRC = PART_CHUNK, a dense chunk copied
// whole into a device part buffer; SRC = PART_HOST, mapped page-locked host
// memory (the tail from static_edges on).  PART_HOST launches count apart.
enum { PART_GATHERED = 0, PART_CHUNK = 1, PART_HOST = 2 };
template <int ALGO, class OffT, class EdgeT, int SRC, bool PK>
__global__ __launch_bounds__(kBlock) void k_expand_part(TravArgs<OffT, EdgeT> a, uint32_t level, PartArgs<EdgeT> p) {
    constexpr bool ZC = SRC != PART_GATHERED;
    __shared__ unsigned long long s_p[kTile];
    __shared__ uint32_t s_sv[kTile];
    __shared__ unsigned long long s_src[ZC ? kTile : 1];
    __shared__ uint32_t s_k;
    ASC_STAGE_DECL;

    Slot* nxt = a.slots + ((level + 1) & 3);
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = wave_id();
    unsigned long long my_edges = 0, my_verts = 0;

    // each workgroup takes a contiguous run of tiles: only its first tile
    // searches the prefix for its first owner; a later tile's first owner is
    // the previous tile's last owner or the entry after it (no entry is empty)
    const uint64_t n_tiles = (p.hi - p.lo + kTile - 1) / kTile;
    const uint64_t per = (n_tiles + gridDim.x - 1) / gridDim.x;
    const uint64_t tb = min(n_tiles, (uint64_t)blockIdx.x * per), te = min(n_tiles, tb + per);
    uint32_t k_last = 0;
    for (uint64_t t = tb; t < te; t++) {
        const uint64_t t0 = p.lo + t * kTile;
        const uint64_t t1 = min(t0 + (uint64_t)kTile, p.hi);
        if (tid == 0) {
            uint32_t k0;
            if (t == tb) {
                uint32_t lo = 0, hi = p.n - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (p.prefix[mid] <= t0) lo = mid; else hi = mid - 1;
                }
                k0 = lo;
            } else {
                k0 = k_last + (k_last + 1 < p.n && p.prefix[k_last + 1] <= t0 ? 1u : 0u);
            }
            s_k = k0;
        }
        __syncthreads();
        const uint32_t k0 = s_k;
        // owners of the tile: the window entries whose start is <= t1 - 1 (a
        // prefix of the window, at most kTile long), loaded 256 at a time
        const uint32_t avail = min((uint32_t)kTile, p.n - k0);
        uint32_t m = 0;
        for (uint32_t base = 0; base < avail; base += kBlock) {
            const uint32_t j = base + tid;
            bool in = false;
            if (j < avail) {
                const unsigned long long pj = p.prefix[k0 + j];
                in = pj <= t1 - 1;
                if (in) s_p[j] = pj;
            }
            const uint32_t c = (uint32_t)__syncthreads_count(in);
            m += c;
            if (c < kBlock) break;
        }
        k_last = k0 + m - 1;
        for (uint32_t j = tid; j < m; j += kBlock) {
            const uint32_t u = p.list[k0 + j];
            if (ALGO == ALG_BFS) {
                s_sv[j] = level + 2;
            } else {
                // the current value (it may have improved since the pile was
                // filtered, by earlier chunks or resident steps); done[u] ends
                // the round at the largest value any tile expanded u with
                const uint32_t cur = a.val[u];
                s_sv[j] = cur;
                if (a.done[u] < cur) atomicMax(a.done + u, cur);
            }
            if (ZC) s_src[j] = (unsigned long long)a.off[u] - p.tail_base;
        }
        __syncthreads();
        const uint32_t tlen = (uint32_t)(t1 - t0);
        for (uint32_t r = w * kWave * kUnroll; r < tlen; r += kBlock * kUnroll) {
            uint32_t kk[kUnroll], vv[kUnroll], ww[kUnroll], pr[kUnroll];
            // lanes past the end load the last edge again (no branch around
            // the loads, so all of them are in flight together)
#pragma unroll
            for (int j = 0; j < kUnroll; j++) {
                const uint64_t e = t0 + min(r + j * kWave + lane, tlen - 1);
                uint32_t k = 0;
#pragma unroll
                for (uint32_t step = kTile / 2; step > 0; step >>= 1)
                    if (k + step < m && s_p[k + step] <= e) k += step;
                if (PK) {
                    const uint64_t ix = ZC ? s_src[k] + (e - s_p[k]) : e - p.lo;
                    const uint32_t x = ZC ? p.zc32[ix] : p.part32[ix];
                    if (p.pk_split) {
                        vv[j] = x;
                        ww[j] = (uint32_t)(ZC ? p.zc_w8[ix] : p.part_w8[ix]) + p.pk_wmin;
                    } else {
                        vv[j] = x & p.pk_mask;
                        ww[j] = (x >> p.pk_shift) + p.pk_wmin;
                    }
                } else {
                    EdgeT ed;
                    if (ZC) ed = p.zc[s_src[k] + (e - s_p[k])];
                    else ed = p.part[e - p.lo];
                    vv[j] = edge_dst(ed);
                    ww[j] = edge_w(ed);
                }
                kk[j] = k;
            }
            uint32_t cc[kUnroll], valid = 0;
#pragma unroll
            for (int j = 0; j < kUnroll; j++) {
                pr[j] = relax_probe<ALGO>(a, vv[j]);
                cc[j] = candidate<ALGO>(s_sv[kk[j]], ww[j]);
                if (r + j * kWave + lane < tlen) valid |= 1u << j;
            }
            int to[kUnroll];
            relax_commit_batch<ALGO>(a, vv, cc, pr, valid, level, to);
#pragma unroll
            for (int j = 0; j < kUnroll; j++) pq.push(a, nxt, to[j], vv[j]);
        }
        my_edges += tlen;
        my_verts += m;
        __syncthreads();
    }
    pq.flush(a, nxt);
    if (tid == 0 && my_edges) {
        atomicAdd(SRC == PART_HOST ? &a.stats->edges_part_host : &a.stats->edges_part, my_edges);
        atomicAdd(SRC == PART_HOST ? &a.stats->verts_part_host : &a.stats->verts_part, my_verts);
    }
}

// Compacts a pile before it is streamed: drops zero-degree vertices and (SSSP,
// CC) vertices already expanded at their current value.  A kept vertex's
// `done` is set to 0 for the round: each tile of k_expand_part that expands it
// reads its value at that moment and raises `done` to it, so `done` ends the
// round at the largest (weakest) value any of its edges was relaxed with, and
// the next round re-expands it only if its value is lower than that.
// Chunk of on-demand vertex v: the largest i < n_ch with ch_v[i] <= v.
__device__ __forceinline__ uint32_t chunk_of(const uint32_t* ch_v, uint32_t n_ch, uint32_t v) {
    uint32_t lo = 0, hi = n_ch - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (ch_v[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int ALGO, class OffT>
__global__ __launch_bounds__(kBlock) void k_filter_pile(const OffT* off, const uint32_t* pile, uint32_t n,
                                                        uint32_t* val, uint32_t* done, uint32_t* out,
                                                        uint32_t* out_deg, uint32_t* count) {
    const uint32_t lane = lane_id();
    for (uint32_t base = (blockIdx.x * kBlock) & ~(kWave - 1u); base < n; base += gridDim.x * kBlock) {
        const uint32_t i = base + threadIdx.x;
        bool keep = false;
        uint32_t v = 0, d = 0, deg = 0;
        if (i < n) {
            v = pile[i];
            const uint64_t dd = (uint64_t)(off[v + 1] - off[v]);
            deg = dd > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)dd;
            keep = dd > 0;
            if (ALGO != ALG_BFS && keep) {
                d = val[v];
                keep = done[v] != d;
                if (keep) done[v] = 0;   // the part kernels raise it to the values they expand v with
            }
        }
        const uint64_t m = wave_ballot(keep);
        if (m == 0) continue;
        uint32_t at = 0;
        if (lane == 0) at = atomicAdd(count, (uint32_t)__popcll(m));
        at = __shfl(at, 0, kWave);
        if (keep) {
            const uint32_t pos = at + lanes_below(m);
            out[pos] = v;
            out_deg[pos] = deg;
        }
    }
}

// Active edges and vertices of a pile per dense-copy chunk (the vertices
// k_filter_pile keeps: degree > 0 and, for SSSP/CC, not yet expanded at their
// value).
template <int ALGO, class OffT>
__global__ __launch_bounds__(kBlock) void k_pile_chunks(const OffT* off, const uint32_t* pile, uint32_t n,
                                                        const uint32_t* val, const uint32_t* done,
                                                        const uint32_t* ch_v, uint32_t n_ch,
                                                        unsigned long long* ch_edges, uint32_t* ch_count) {
    __shared__ unsigned long long s_h[kMaxChunks];
    __shared__ uint32_t s_c[kMaxChunks];
    for (uint32_t c = threadIdx.x; c < n_ch; c += kBlock) {
        s_h[c] = 0;
        s_c[c] = 0;
    }
    __syncthreads();
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const uint32_t v = pile[i];
        const uint64_t d = (uint64_t)(off[v + 1] - off[v]);
        if (d == 0) continue;
        if (ALGO != ALG_BFS && val[v] == done[v]) continue;
        const uint32_t c = chunk_of(ch_v, n_ch, v);
        atomicAdd(&s_h[c], (unsigned long long)d);
        atomicAdd(&s_c[c], 1u);
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < n_ch; c += kBlock)
        if (s_c[c]) {
            atomicAdd(ch_edges + c, s_h[c]);
            atomicAdd(ch_count + c, s_c[c]);
        }
}

// k_filter_pile for a round with dense chunks: a kept vertex of a dense chunk
// c goes to that chunk's list segment (cursor[c], placed by the host after
// the sparse segment), every other kept vertex to the sparse segment
// (cursor[n_ch], from 0).  Workgroups count their entries per segment in LDS
// and reserve each segment's share with one atomic.
template <int ALGO, class OffT>
__global__ __launch_bounds__(kBlock) void k_filter_pile_chunked(const OffT* off, const uint32_t* pile, uint32_t n,
                                                                uint32_t* val, uint32_t* done, uint32_t* out,
                                                                uint32_t* out_deg,
                                                                const uint32_t* ch_v, uint32_t n_ch,
                                                                const uint8_t* ch_dense, uint32_t* cursor) {
    __shared__ uint32_t s_cnt[kMaxChunks + 1];
    __shared__ uint32_t s_base[kMaxChunks + 1];
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        for (uint32_t c = threadIdx.x; c <= n_ch; c += kBlock) s_cnt[c] = 0;
        __syncthreads();
        const uint32_t i = base + threadIdx.x;
        bool keep = false;
        uint32_t v = 0, d = 0, deg = 0, seg = 0, rank = 0;
        if (i < n) {
            v = pile[i];
            const uint64_t dd = (uint64_t)(off[v + 1] - off[v]);
            deg = dd > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)dd;
            keep = dd > 0;
            if (ALGO != ALG_BFS && keep) {
                d = val[v];
                keep = done[v] != d;
                if (keep) done[v] = 0;   // as in k_filter_pile
            }
            if (keep) {
                const uint32_t c = chunk_of(ch_v, n_ch, v);
                seg = ch_dense[c] ? c : n_ch;
                rank = atomicAdd(&s_cnt[seg], 1u);
            }
        }
        __syncthreads();
        for (uint32_t c = threadIdx.x; c <= n_ch; c += kBlock)
            if (s_cnt[c]) s_base[c] = atomicAdd(cursor + c, s_cnt[c]);
        __syncthreads();
        if (keep) {
            const uint32_t pos = s_base[seg] + rank;
            out[pos] = v;
            out_deg[pos] = deg;
        }
        __syncthreads();
    }
}

// Exclusive u64 prefix of a filtered list's u32 degrees, on the device (the
// part kernels search it): tile sums, one workgroup scanning the tile sums,
// then each tile rescanned from its base.  out[n] = the total.
constexpr uint32_t kScanPer = 8;
constexpr uint32_t kScanTile = kBlock * kScanPer;

__global__ __launch_bounds__(kBlock) void k_scan_tiles(const uint32_t* x, uint32_t n, unsigned long long* tile_sum) {
    __shared__ unsigned long long s_w[kWavesPerBlock];
    const uint64_t b = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
    unsigned long long acc = 0;
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; j++)
        if (b + j < n) acc += x[b + j];
    acc = wave_sum(acc);
    if (lane_id() == 0) s_w[wave_id()] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
#pragma unroll
        for (int i = 0; i < kWavesPerBlock; i++) t += s_w[i];
        tile_sum[blockIdx.x] = t;
    }
}

// One workgroup: exclusive scan of nt tile sums in place; out_total = sum.
__global__ __launch_bounds__(kBlock) void k_scan_tile_sums(unsigned long long* tile_sum, uint32_t nt,
                                                           unsigned long long* out_total) {
    __shared__ unsigned long long s_tmp[kWavesPerBlock];
    const uint32_t per = (nt + kBlock - 1) / kBlock;
    const uint32_t b = threadIdx.x * per, e = min(nt, b + per);
    unsigned long long acc = 0;
    for (uint32_t i = b; i < e; i++) acc += tile_sum[i];
    unsigned long long total;
    unsigned long long base = block_scan_excl(acc, s_tmp, total);
    for (uint32_t i = b; i < e; i++) {
        const unsigned long long v = tile_sum[i];
        tile_sum[i] = base;
        base += v;
    }
    if (threadIdx.x == 0) *out_total = total;
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(const uint32_t* x, uint32_t n,
                                                       const unsigned long long* tile_base, unsigned long long* out) {
    __shared__ unsigned long long s_tmp[kWavesPerBlock];
    const uint64_t b = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
    uint32_t v[kScanPer];
    unsigned long long acc = 0;
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; j++) {
        v[j] = b + j < n ? x[b + j] : 0u;
        acc += v[j];
    }
    unsigned long long total;
    unsigned long long run = tile_base[blockIdx.x] + block_scan_excl(acc, s_tmp, total);
#pragma unroll
    for (uint32_t j = 0; j < kScanPer; j++)
        if (b + j < n) {
            out[b + j] = run;
            run += v[j];
        }
}

// Moves the far pile's entries whose value fell below theta into the current
// frontier (resident) or the on-demand pile; the rest stays far.  A vertex
// enters the far pile once per run (its far stamp stays set): a value that
// falls below theta is expanded in the band and can only fall further, and a
// value that stays above is read again from the one entry at the next split,
// so kept entries are copied without a stamp atomic.
template <class OffT, class EdgeT>
__global__ __launch_bounds__(kBlock) void k_split_far(TravArgs<OffT, EdgeT> a, const uint32_t* far_in, uint32_t n,
                                                      uint32_t level, uint32_t* q_level, Piles* piles) {
    constexpr int ALGO = ALG_SSSP;
    ASC_STAGE_DECL;
    Slot* slot = a.slots + (level & 3);
    uint32_t lmin = 0xFFFFFFFFu;
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const uint32_t i = base + threadIdx.x;
        int to = TO_NONE;
        uint32_t v = 0;
        if (i < n) {
            v = far_in[i];
            const uint32_t d = a.val[v];
            if (d >= a.theta) {
                // a vertex is expanded only below the band of its round and
                // bands only rise, so done[v] < theta here: still pending
                to = TO_FAR;
                lmin = min(lmin, d);
            } else if (a.done[v] != d) {
                if (v < a.vs) to = atomicMax(a.stamp + v, level) < level ? TO_RES : TO_NONE;
                else to = atomicMax(a.pstamp + v, a.pepoch) < a.pepoch ? TO_PEND : TO_NONE;
            }
        }
        pq.res.push(to == TO_RES, v, q_level, &slot->n_res);
        pq.pend.push(to == TO_PEND, v, a.pend_out, a.pend_count);
        pq.far.push(to == TO_FAR, v, a.far_out, a.far_count);
    }
    pq.res.flush(q_level, &slot->n_res);
    pq.pend.flush(a.pend_out, a.pend_count);
    pq.far.flush(a.far_out, a.far_count);
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) lmin = min(lmin, (uint32_t)__shfl_xor(lmin, d, kWave));
    if (lane_id() == 0 && lmin != 0xFFFFFFFFu) atomicMin(&piles->min_far, lmin);
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

// Publishes a level's counters (the next step's slot and the piles) to mapped
// host memory, then the sequence number the host polls for.
__global__ void k_trav_publish(const Slot* slot, const Piles* piles, LevelPub* host, unsigned long long seq) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    host->slot = *slot;
    host->piles = *piles;
    __threadfence_system();
    host->seq = seq;
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Step-0 frontier for BFS / SSSP: the source alone (GraphMeta.cuh:432-446).
__global__ void k_init_source(uint32_t* val, uint32_t* visited, uint32_t* pstamp, uint32_t* q, uint32_t* pend,
                              Slot* slots, Piles* piles, uint32_t src, uint32_t vs) {
    val[src] = 1;
    if (visited) visited[src >> 5] |= 1u << (src & 31);
    if (src < vs) {
        q[0] = src;
        slots[0].n_res = 1;
    } else {
        pend[0] = src;
        piles->n_pend[0] = 1;
        if (pstamp) pstamp[src] = 1;
    }
}

// CC: every vertex starts with its own id and active (GraphMeta.cuh:448-452).
__global__ void k_init_cc(uint32_t* val, uint32_t* pstamp, uint32_t* q, uint32_t* pend, Slot* slots, Piles* piles,
                          uint32_t V, uint32_t vs) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        val[v] = v;
        if (v < vs) {
            q[v] = v;
        } else {
            pend[v - vs] = v;
            pstamp[v] = 1;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        slots[0].n_res = vs;
        piles->n_pend[0] = V - vs;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

void trav_free(TravState* t) {
    if (!t) return;
    auto f = [](void* p) {
        if (p) (void)hipFree(p);
    };
    f(t->val); f(t->val16); f(t->stamp); f(t->pstamp); f(t->fstamp); f(t->done); f(t->visited);
    for (int i = 0; i < 2; i++) {
        f(t->q[i]); f(t->pend[i]); f(t->far[i]);
    }
    for (int i = 0; i < kMaxPartBufs; i++) f(t->part[i]);
    f(t->heavy_u); f(t->heavy_sv); f(t->heavy_base);
    f(t->od_list); f(t->od_prefix); f(t->scan_tmp); f(t->slots); f(t->piles); f(t->stats);
    f(t->bm[0]); f(t->bm[1]); f(t->bm[2]); f(t->pctl); f(t->dcnt); f(t->d_ch_v); f(t->d_ch_edges); f(t->d_ch_count); f(t->d_ch_dense);
    f(t->d_cursor);
    for (int i = 0; i < kMaxPartBufs; i++) {
        if (t->ev_copy[i]) (void)hipEventDestroy(t->ev_copy[i]);
        if (t->ev_kern[i]) (void)hipEventDestroy(t->ev_kern[i]);
    }
    for (int i = 0; i < 2; i++)
        if (t->ev_stage[i]) (void)hipEventDestroy(t->ev_stage[i]);
    if (t->ev_meta) (void)hipEventDestroy(t->ev_meta);
    for (auto e : t->ev_pool) (void)hipEventDestroy(e);
    delete t;
}

// Vertex-aligned chunks of the on-demand range for whole-chunk copies: each
// holds at most one part of edges (a vertex larger than a part gets a chunk
// of its own, never copied whole).  Needs the tail in pinned host memory
// (packed words or the pinned copy) and gather streaming.
static int plan_dense_chunks(asc_graph* g, TravState* t) {
    t->ch_v.clear();
    t->ch_e.clear();
    const asc_config& cfg = g->ctx->cfg;
    if (cfg.stream_mode == ASC_STREAM_ZEROCOPY || cfg.dense_threshold > 1.0) return ASC_OK;
    if (!g->packed && !g->zc_tail.p) return ASC_OK;
    const uint64_t V = g->h.V, vs = g->static_vertices;
    const uint64_t* off = g->h.off.p;
    const uint64_t cap = t->part_edges;
    std::vector<uint32_t> cv;
    std::vector<uint64_t> ce;
    uint64_t v = vs;
    while (v < V) {
        uint64_t v1 = v + 1;
        while (v1 < V && off[v1 + 1] - off[v] <= cap) v1++;
        cv.push_back((uint32_t)v);
        ce.push_back(off[v]);
        v = v1;
        if (cv.size() > kMaxChunks) return ASC_OK;   // too many: gather only
    }
    cv.push_back((uint32_t)V);
    ce.push_back(off[V]);
    const size_t n = cv.size() - 1;
    ASC_HIP(hipMalloc(&t->d_ch_v, 4 * cv.size()));
    ASC_HIP(hipMemcpy(t->d_ch_v, cv.data(), 4 * cv.size(), hipMemcpyHostToDevice));
    ASC_HIP(hipMalloc(&t->d_ch_edges, 8 * n));
    ASC_HIP(hipMalloc(&t->d_ch_count, 4 * n));
    ASC_HIP(hipMalloc(&t->d_ch_dense, n));
    ASC_HIP(hipMalloc(&t->d_cursor, 4 * (n + 1)));
    ASC_TRY(t->h_ch_edges.alloc(n, true));
    t->h_ch_lo.assign(n, 0);
    ASC_TRY(t->h_ch_count.alloc(n, true));
    ASC_TRY(t->h_ch_dense.alloc(n, true));
    ASC_TRY(t->h_cursor.alloc(n + 1, true));
    t->ch_v = std::move(cv);
    t->ch_e = std::move(ce);
    // pinned host-to-device rate for choosing dense chunks
    double h2d = 0, d2h = 0;
    if (asc_measure_link(g->ctx, 64u << 20, 3, &h2d, &d2h) == ASC_OK && h2d > 0) t->link_bps = h2d * 1e9;
    return ASC_OK;
}

int trav_ensure(asc_graph* g) {
    const uint64_t V = g->h.V;
    const size_t EB = g->stream_record_bytes();  // bytes per streamed edge
    TravState* t = g->trav;
    const bool need_parts = g->static_vertices < V;
    const int mode = g->ctx->cfg.stream_mode;
    if (need_parts && (mode == ASC_STREAM_ZEROCOPY || mode == ASC_STREAM_HYBRID) && !g->tail_pinned()) {
        set_last_error("zero-copy and hybrid streaming read the on-demand edges from page-locked memory, which the "
                       "system did not grant for this graph's tail; use gather streaming");
        return ASC_ERR_UNSUPPORTED;
    }
    if (t && t->V == V && (!need_parts || (t->od_list && t->part_edges >= g->part_edges))) return ASC_OK;
    if (t) trav_free(t);
    t = new TravState();
    g->trav = t;
    t->V = V;
    ASC_HIP(hipSetDevice(g->ctx->device));
    for (uint32_t** p : {&t->val, &t->stamp, &t->pstamp, &t->fstamp, &t->done, &t->q[0], &t->q[1], &t->far[0],
                         &t->far[1]})
        ASC_HIP(hipMalloc(p, 4 * V));
    ASC_HIP(hipMalloc(&t->visited, 4 * ((V + 31) / 32)));
    ASC_HIP(hipMalloc(&t->heavy_u, 4 * (size_t)kHeavyCap));
    ASC_HIP(hipMalloc(&t->heavy_sv, 4 * (size_t)kHeavyCap));
    ASC_HIP(hipMalloc(&t->heavy_base, 8 * (size_t)kHeavyCap));
    ASC_HIP(hipMalloc(&t->slots, sizeof(Slot) * 4));
    ASC_HIP(hipMalloc(&t->piles, sizeof(Piles)));
    ASC_HIP(hipMalloc(&t->stats, sizeof(DevStats)));
    ASC_TRY(t->h_slot.alloc(4, true));
    ASC_TRY(t->h_piles.alloc(1, true));
    ASC_TRY(t->h_pub.alloc(1, true));
    t->h_pub[0].seq = 0;
    ASC_HIP(hipHostGetDevicePointer((void**)&t->d_pub, t->h_pub.p, 0));
    ASC_TRY(t->h_stats.alloc(1, true));
    if (need_parts) {
        const uint64_t nod = V - g->static_vertices;
        ASC_HIP(hipMalloc(&t->pend[0], 4 * nod));
        ASC_HIP(hipMalloc(&t->pend[1], 4 * nod));
        // od_list holds the filtered list, then the degrees (nod entries each)
        ASC_HIP(hipMalloc(&t->od_list, 8 * nod));
        ASC_HIP(hipMalloc(&t->od_prefix, 8 * (nod + 1)));
        ASC_HIP(hipMalloc(&t->scan_tmp, 8 * (nod / kScanTile + 2)));
        ASC_TRY(t->h_src.alloc(nod, true));
        ASC_HIP(hipHostGetDevicePointer((void**)&t->d_hsrc, t->h_src.p, 0));
        ASC_TRY(t->h_prefix.alloc(nod + 1, true));
        t->part_edges = g->part_edges;
        const int nb_cfg = g->ctx->cfg.part_buffers;
        t->n_bufs = nb_cfg > 0 ? std::min(std::max(nb_cfg, 2), kMaxPartBufs) : kDefaultPartBufs;
        for (int i = 0; i < t->n_bufs; i++) {
            t->buf_chunk[i] = -1;
            if (g->ctx->cfg.stream_mode != ASC_STREAM_ZEROCOPY)
                ASC_HIP(hipMalloc(&t->part[i], EB * t->part_edges));
            ASC_HIP(hipEventCreateWithFlags(&t->ev_copy[i], hipEventDisableTiming));
            ASC_HIP(hipEventCreateWithFlags(&t->ev_kern[i], hipEventDisableTiming));
            // recorded once so that every copy can wait for its buffer's last reader
            ASC_HIP(hipEventRecord(t->ev_kern[i], g->ctx->s_main));
        }
        for (int i = 0; i < 2; i++) {
            if (g->ctx->cfg.stream_mode == ASC_STREAM_GATHER)
                ASC_TRY(t->stage[i].alloc(EB * t->part_edges, true));
            ASC_HIP(hipEventCreateWithFlags(&t->ev_stage[i], hipEventDisableTiming));
            // recorded once so that every use can wait for the buffer's last copy
            ASC_HIP(hipEventRecord(t->ev_stage[i], g->ctx->s_copy));
        }
        ASC_HIP(hipEventCreateWithFlags(&t->ev_meta, hipEventDisableTiming));
        ASC_TRY(plan_dense_chunks(g, t));
    }
    return ASC_OK;
}

static hipEvent_t pooled_event(TravState* t, size_t i) {
    while (t->ev_pool.size() <= i) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        t->ev_pool.push_back(e);
    }
    return t->ev_pool[i];
}

hipEvent_t next_event(TravState* t) { return pooled_event(t, t->ev_next++); }

int effective_heavy_threshold(const asc_graph* g) {
    uint64_t thr = g->heavy_threshold > 0 ? (uint64_t)g->heavy_threshold : 2048;
    const uint64_t floor_cap = (g->h.E + kHeavyCap - 1) / kHeavyCap;  // list can never overflow
    thr = std::max<uint64_t>(thr, floor_cap);
    thr = std::max<uint64_t>(thr, (uint64_t)kMinHeavy);
    thr = std::min<uint64_t>(thr, uint64_t(1) << 22);
    return (int)thr;
}

// SSSP bucket width: configured, else a multiple of the mean edge weight
// (computed once per graph); a negative setting disables the far pile.
void compute_mean_weight(asc_graph* g) {
    if (g->mean_weight > 0 || g->h.kind != ASC_KIND_WCSR || !g->h.edges.p) return;
    const uint64_t E = g->h.E;
    std::atomic<uint64_t> sum{0};
    host_pool().parallel_for(E, [&](uint64_t b, uint64_t e) {
        uint64_t s = 0;
        for (uint64_t i = b; i < e; i++) s += g->h.edges[2 * i + 1];
        sum += s;
    });
    g->mean_weight = E ? (double)sum.load() / (double)E : 1.0;
}

// The 16-bit SSSP value filter pays when most values fit 16 bits: values grow
// by about the mean weight per hop, and the paths of these graphs stay within
// a few times log2(V) hops.  Where they do not, most entries say "nothing
// known" and every probe tries its atomic (R-MAT scale 24, weights up to
// 2^20: 45.8 ms against 37.6 with the plain value probe).
static bool sssp_filter_pays(asc_graph* g) {
    compute_mean_weight(g);
    if (g->mean_weight <= 0) return true;
    return g->mean_weight * 4.0 * std::log2((double)g->h.V + 1.0) < 65535.0;
}

static uint32_t sssp_delta(asc_graph* g) {
    if (g->sssp_delta > 0) return (uint32_t)g->sssp_delta;
    if (g->sssp_delta < 0) return 0;
    compute_mean_weight(g);
    // streamed graphs take wide bands (each band re-streams the on-demand
    // vertices it improves, so fewer rounds win; streamed vertices are
    // expanded with their value at expansion time, which limits the
    // re-expansions wide bands cause; measured: R-MAT scale 24, weights 1-64,
    // 6 to 30 times the mean weight equal, 2 times 5% slower; scale 28,
    // weights 1-255, 12 times 7% faster than 6); fully resident ones narrow
    // bands (fewer re-expansions; a step costs only a launch and a sync)
    const double f = g->static_vertices == g->h.V ? 0.25 : 12.0;
    return (uint32_t)std::min(std::max(1.0, f * g->mean_weight), 2147483648.0);
}

// Copies the adjacency of list entries overlapping [lo, hi) of the
// concatenated on-demand edge list into `dst` (the reference's
// fillEdgeArrByMultiThread, GraphMeta.cuh:469-514, cut to one part).
// src[k]: first record of the k-th listed vertex in the array gathered from
// (the packed tail, indexed from static_edges on, or the host records).
static void gather_part(asc_graph* g, const unsigned long long* src, const unsigned long long* prefix, uint64_t m,
                        uint64_t lo, uint64_t hi, char* dst, char* dst_w) {
    if (g->packed) {
        gather_edges(*g->ctx->pool, (const char*)g->ptail.p, 4, src, prefix, m, lo, hi, dst);
        if (g->pk_split)   // the weight bytes, at the same positions of their own array
            gather_edges(*g->ctx->pool, (const char*)g->ptail_w.p, 1, src, prefix, m, lo, hi, dst_w);
    } else
        gather_edges(*g->ctx->pool, (const char*)g->h.edges.p, g->h.edge_bytes(), src, prefix, m, lo, hi, dst);
}

// Chooses the chunks copied whole: every chunk with at least half of its
// edges active, then, densest first, further chunks while the estimated
// round time falls, so that the two transfer paths finish together: with S
// the active bytes left to gather and D the bytes copied whole, a round
// takes about max((D + S) / link, S / gather), since gathered parts also
// cross the link.
static void choose_dense_chunks(TravState* t, size_t EB, const std::vector<int>& buf_of) {
    const uint32_t n_ch = (uint32_t)t->ch_v.size() - 1;
    const double L = t->link_bps > 0 ? t->link_bps : 50e9, G = t->gather_bps;
    std::vector<uint32_t> order;
    double S = 0, D = 0;
    for (uint32_t c = 0; c < n_ch; c++) {
        t->h_ch_dense[c] = 0;
        if (t->h_ch_count[c] == 0) continue;
        if (buf_of[c] >= 0) {   // already in a part buffer: read there, no transfer
            t->h_ch_dense[c] = 1;
            continue;
        }
        S += (double)t->h_ch_edges[c] * EB;
        if (t->ch_e[c + 1] - t->ch_e[c] <= t->part_edges) order.push_back(c);
    }
    auto density = [&](uint32_t c) { return (double)t->h_ch_edges[c] / (double)(t->ch_e[c + 1] - t->ch_e[c]); };
    std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return density(x) > density(y); });
    auto cost = [&](double s, double d) { return std::max((d + s) / L, s / G); };
    for (uint32_t c : order) {
        const double k = (double)t->h_ch_edges[c] * EB, C = (double)(t->ch_e[c + 1] - t->ch_e[c]) * EB;
        if (density(c) < 0.5 && cost(S - k, D + C) >= cost(S, D)) break;
        S -= k;
        D += C;
        t->h_ch_dense[c] = 1;
    }
}

// Copies `cnt` (at most one part) streamed records into device part buffer b:
// whole records from `src`, or in the split form the ids from `src` and the
// weight bytes from `src_w`, placed behind the buffer's part_edges ids.
// Orders stream `s` after `ev`, unless the host already sees `ev` complete
// (a part buffer whose last reader or copy ended long ago): the wait would
// only add a dependency between the copy and compute queues.
static int wait_event(hipStream_t s, hipEvent_t ev) {
    const hipError_t q = hipEventQuery(ev);
    if (q == hipSuccess) return ASC_OK;
    if (q != hipErrorNotReady) ASC_HIP(q);
    ASC_HIP(hipStreamWaitEvent(s, ev, 0));
    return ASC_OK;
}

static int copy_to_part(asc_graph* g, TravState* t, int b, const char* src, const char* src_w, uint64_t cnt,
                        hipStream_t s) {
    if (cnt > t->part_edges) {
        set_last_error("part copy of %llu records exceeds the part buffer", (unsigned long long)cnt);
        return ASC_ERR_STATE;
    }
    if (g->pk_split) {
        ASC_HIP(hipMemcpyAsync(t->part[b], src, 4 * cnt, hipMemcpyHostToDevice, s));
        ASC_HIP(hipMemcpyAsync((char*)t->part[b] + 4 * t->part_edges, src_w, cnt, hipMemcpyHostToDevice, s));
    } else {
        ASC_HIP(hipMemcpyAsync(t->part[b], src, cnt * g->stream_record_bytes(), hipMemcpyHostToDevice, s));
    }
    return ASC_OK;
}

// Host addresses of on-demand edge e in the pinned tail: its record (the id
// word in the packed forms) and, in the split form, its weight byte.
static const char* tail_rec(const asc_graph* g, uint64_t e) {
    const uint64_t i = e - g->static_edges;
    return g->packed ? (const char*)(g->ptail.p + i) : (const char*)g->zc_tail.p + i * g->h.edge_bytes();
}
static const char* tail_w8(const asc_graph* g, uint64_t e) {
    return g->pk_split ? (const char*)(g->ptail_w.p + (e - g->static_edges)) : nullptr;
}

// Queues the copy of the on-demand edges [e0, e1) (a whole chunk) into part
// buffer b on stream s: one DMA straight from the tail when it is
// page-locked; from a pageable tail the host threads first copy the records
// into the next pinned staging buffer (the gather's, taken in turn), which
// then goes by DMA.  *fill_ms accumulates the host copy time.
static int copy_chunk_to_part(asc_graph* g, TravState* t, int b, uint64_t e0, uint64_t e1, hipStream_t s,
                              double* fill_ms) {
    if (g->tail_pinned()) return copy_to_part(g, t, b, tail_rec(g, e0), tail_w8(g, e0), e1 - e0, s);
    if (!t->stage[0].p) {
        set_last_error("a pageable on-demand tail needs the gather streaming mode's staging buffers");
        return ASC_ERR_UNSUPPORTED;
    }
    const int sb = (int)(t->stage_seq++ & 1);
    ASC_HIP(hipEventSynchronize(t->ev_stage[sb]));   // its previous copy has left
    const uint64_t n = e1 - e0;
    const size_t rb = g->pk_split ? 4 : g->stream_record_bytes();
    char* dst = t->stage[sb].p;
    char* dst_w = dst + 4 * t->part_edges;
    const char* src = tail_rec(g, e0);
    const char* src_w = tail_w8(g, e0);
    const double t0 = now_ms();
    g->ctx->pool->parallel_for(n, [&](uint64_t lo, uint64_t hi) {
        std::memcpy(dst + lo * rb, src + lo * rb, (hi - lo) * rb);
        if (src_w) std::memcpy(dst_w + lo, src_w + lo, hi - lo);
    });
    if (fill_ms) *fill_ms += now_ms() - t0;
    ASC_TRY(copy_to_part(g, t, b, dst, dst_w, n, s));
    ASC_HIP(hipEventRecord(t->ev_stage[sb], s));
    return ASC_OK;
}

// Streams one pile: filter on the device, list to the host, u64 prefix of
// degrees, then parts (gather + double-buffered copies) or zero-copy reads.
// With dense chunks planned, the pile's active edges are first counted per
// chunk; a chunk whose active edges reach asc_config.dense_threshold of its
// edges is copied whole by DMA from the pinned tail (no host gather) and its
// active vertices, kept in their own list segment, are expanded from it.
// At the start of a run that may stream, the first chunks of the on-demand
// region are copied into the part buffers while the resident part is
// relaxed (the link is idle until the first streaming round).  A round that
// finds a chunk it copies whole already in a part buffer (copied ahead, or
// by an earlier round of the run and not overwritten since) reads it there.
static int prefetch_chunks(asc_graph* g, TravState* t) {
    asc_ctx* ctx = g->ctx;
    for (int b = 0; b < t->n_bufs; b++) {
        t->buf_chunk[b] = -1;
        t->buf_counted[b] = false;
    }
    const uint32_t n_ch = t->ch_v.empty() ? 0 : (uint32_t)t->ch_v.size() - 1;
    if (!n_ch || ctx->cfg.stream_mode == ASC_STREAM_ZEROCOPY || ctx->cfg.dense_threshold > 1.0) return ASC_OK;
    if (!g->packed && !g->zc_tail.p) return ASC_OK;
    for (int b = 0; b < t->n_bufs && (uint32_t)b < n_ch; b++) {
        const uint64_t e0 = t->ch_e[b], e1 = t->ch_e[b + 1];
        if (e1 - e0 > t->part_edges) break;
        ASC_TRY(wait_event(ctx->s_copy, t->ev_kern[b]));
        ASC_TRY(copy_chunk_to_part(g, t, b, e0, e1, ctx->s_copy, nullptr));
        ASC_HIP(hipEventRecord(t->ev_copy[b], ctx->s_copy));
        t->buf_chunk[b] = b;
    }
    return ASC_OK;
}

// First record of each listed vertex's edges in the array the host gathers
// from (offset minus `base`), written to mapped page-locked host memory.
template <class OffT>
__global__ __launch_bounds__(kBlock) void k_list_src(const OffT* off, const uint32_t* list, uint32_t n,
                                                      unsigned long long base, unsigned long long* out) {
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        out[i] = (unsigned long long)off[list[i]] - base;
}

// od_prefix[0..n] = exclusive u64 prefix of the n degrees (device only).
static int scan_degrees(TravState* t, const uint32_t* deg, uint32_t n, hipStream_t s) {
    if (n == 0) return ASC_OK;
    const uint32_t nt = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kBlock), 0, s, deg, n, t->scan_tmp);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(kBlock), 0, s, t->scan_tmp, nt, t->od_prefix + n);
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(kBlock), 0, s, deg, n, (const unsigned long long*)t->scan_tmp,
                       t->od_prefix);
    ASC_HIP(hipGetLastError());
    return ASC_OK;
}

template <int ALGO, class OffT, class EdgeT>
static int stream_pile(asc_graph* g, TravArgs<OffT, EdgeT>& a, uint32_t& level, const uint32_t* pile,
                       uint32_t npile, int grid, RunCounters& rc) {
    asc_ctx* ctx = g->ctx;
    TravState* t = g->trav;
    const bool prof = ctx->cfg.profile != 0;
    hipStream_t s = ctx->s_main;
    const double t_begin = now_ms();
    const uint64_t nod = g->h.V - g->static_vertices;
    uint32_t* list_deg = t->od_list + nod;
    const bool pk = g->packed;
    const size_t EB = g->stream_record_bytes();
    const uint64_t S = g->static_edges;
    const int pile_grid = (int)std::max<uint32_t>(1, std::min<uint32_t>((npile + kBlock - 1) / kBlock, 4096));
    // the host gather's index of the first n list entries (the sparse
    // segment): each vertex's first record in the array gathered from,
    // written by the device into mapped host memory, and the device's u64
    // prefix of their degrees
    auto gather_index = [&](uint32_t n) -> int {
        const int ig = (int)std::max<uint32_t>(1, std::min<uint32_t>((n + kBlock - 1) / kBlock, 4096));
        hipLaunchKernelGGL((k_list_src<OffT>), dim3(ig), dim3(kBlock), 0, s, a.off, (const uint32_t*)t->od_list, n,
                           (unsigned long long)(pk ? S : 0), t->d_hsrc);
        ASC_HIP(hipGetLastError());
        ASC_HIP(hipMemcpyAsync(t->h_prefix.p, t->od_prefix, 8 * ((size_t)n + 1), hipMemcpyDeviceToHost, s));
        return ASC_OK;
    };

    // active edges per chunk -> which chunks are copied whole
    const uint32_t n_ch = t->ch_v.empty() ? 0 : (uint32_t)t->ch_v.size() - 1;
    uint32_t n_dense = 0, n_sparse = 0, m = 0;
    uint64_t sparse_edges = 0;           // active edges outside dense chunks
    bool zc_sparse = false;
    std::vector<int> buf_of(n_ch, -1);   // part buffer holding each chunk whole (-1: none)
    if (n_ch) {
        ASC_HIP(hipMemsetAsync(t->d_ch_edges, 0, 8 * (size_t)n_ch, s));
        ASC_HIP(hipMemsetAsync(t->d_ch_count, 0, 4 * (size_t)n_ch, s));
        hipLaunchKernelGGL((k_pile_chunks<ALGO, OffT>), dim3(pile_grid), dim3(kBlock), 0, s, a.off, pile, npile, a.val,
                           a.done, t->d_ch_v, n_ch, t->d_ch_edges, t->d_ch_count);
        ASC_HIP(hipGetLastError());
        ASC_HIP(hipMemcpyAsync(t->h_ch_edges.p, t->d_ch_edges, 8 * (size_t)n_ch, hipMemcpyDeviceToHost, s));
        ASC_HIP(hipMemcpyAsync(t->h_ch_count.p, t->d_ch_count, 4 * (size_t)n_ch, hipMemcpyDeviceToHost, s));
        ASC_HIP(hipStreamSynchronize(s));
        for (int b = 0; b < t->n_bufs; b++)
            if (t->buf_chunk[b] >= 0 && (uint32_t)t->buf_chunk[b] < n_ch) buf_of[t->buf_chunk[b]] = b;
        if (ctx->cfg.dense_threshold > 0) {
            const double thr = ctx->cfg.dense_threshold;
            for (uint32_t c = 0; c < n_ch; c++) {
                const uint64_t ce = t->ch_e[c + 1] - t->ch_e[c];
                t->h_ch_dense[c] = t->h_ch_count[c] > 0 && ce <= t->part_edges &&
                                   ((double)t->h_ch_edges[c] >= thr * (double)ce || buf_of[c] >= 0);
            }
        } else {
            choose_dense_chunks(t, EB, buf_of);
        }
        for (uint32_t c = 0; c < n_ch; c++) {
            if (t->h_ch_dense[c]) {
                n_dense++;
            } else {
                n_sparse += t->h_ch_count[c];
                sparse_edges += t->h_ch_edges[c];
            }
        }
        // a sparse segment of at most half a part is read by the kernel
        // straight from the page-locked tail: no list download, host prefix,
        // gather or copy (C2's late rounds, 6M edges and below, where those
        // cost more than the reads over the link)
        zc_sparse = ctx->cfg.stream_mode == ASC_STREAM_GATHER && g->tail_pinned() && sparse_edges > 0 &&
                    sparse_edges <= t->part_edges / 2;
        if (ctx->cfg.verbose > 2) {   // per-chunk active edges (D: copied whole)
            printf("round %u chunks:", rc.stream_rounds + 1);
            for (uint32_t c = 0; c < n_ch; c++)
                if (t->h_ch_count[c])
                    printf(" %u:%llu%s", c, (unsigned long long)t->h_ch_edges[c], t->h_ch_dense[c] ? "D" : "");
            printf("\n");
        }
    }
    if (n_dense) {
        // segments: sparse vertices first, then one per dense chunk
        uint32_t at = n_sparse;
        t->h_cursor[n_ch] = 0;
        for (uint32_t c = 0; c < n_ch; c++) {
            t->h_cursor[c] = at;
            if (t->h_ch_dense[c]) at += t->h_ch_count[c];
        }
        m = at;
        ASC_HIP(hipMemcpyAsync(t->d_ch_dense, t->h_ch_dense.p, n_ch, hipMemcpyHostToDevice, s));
        ASC_HIP(hipMemcpyAsync(t->d_cursor, t->h_cursor.p, 4 * ((size_t)n_ch + 1), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL((k_filter_pile_chunked<ALGO, OffT>), dim3(pile_grid), dim3(kBlock), 0, s, a.off, pile,
                           npile, a.val, a.done, t->od_list, list_deg, t->d_ch_v, n_ch,
                           (const uint8_t*)t->d_ch_dense, t->d_cursor);
        ASC_HIP(hipGetLastError());
        ASC_TRY(scan_degrees(t, list_deg, m, s));
        // the host gathers the sparse segment only: its index alone comes back
        if (n_sparse && !zc_sparse && ctx->cfg.stream_mode == ASC_STREAM_GATHER) ASC_TRY(gather_index(n_sparse));
        ASC_HIP(hipStreamSynchronize(s));
    } else {
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, &t->piles->n_list, 0u);
        hipLaunchKernelGGL((k_filter_pile<ALGO, OffT>), dim3(pile_grid), dim3(kBlock), 0, s, a.off, pile, npile,
                           a.val, a.done, t->od_list, list_deg, &t->piles->n_list);
        ASC_HIP(hipGetLastError());
        ASC_HIP(hipMemcpyAsync(t->h_piles.p, t->piles, sizeof(Piles), hipMemcpyDeviceToHost, s));
        ASC_HIP(hipStreamSynchronize(s));
        m = t->h_piles[0].n_list;
        n_sparse = m;
        if (m) {
            ASC_TRY(scan_degrees(t, list_deg, m, s));
            if (ctx->cfg.stream_mode == ASC_STREAM_GATHER) ASC_TRY(gather_index(m));
            ASC_HIP(hipStreamSynchronize(s));
        }
    }
    if (m == 0) {
        rc.pre_ms += now_ms() - t_begin;
        return ASC_OK;
    }
    // edges of the sparse segment: the device prefix's entry at its end (the
    // chunk counts agree with it; the kernel's range must be exact), already
    // on the host with the gather index
    uint64_t Ts = 0;
    if (n_sparse && !zc_sparse && ctx->cfg.stream_mode == ASC_STREAM_GATHER) {
        Ts = t->h_prefix[n_sparse];
    } else if (n_sparse) {   // read by the kernel from mapped memory: the total alone
        unsigned long long tot = 0;
        ASC_HIP(hipMemcpyAsync(&tot, t->od_prefix + n_sparse, 8, hipMemcpyDeviceToHost, s));
        ASC_HIP(hipStreamSynchronize(s));
        Ts = tot;
    }
    t->h_prefix[n_sparse] = Ts;
    // dense segments follow in chunk order; each holds its chunk's active
    // edges, so their bounds in the prefix follow from the chunk counts
    uint64_t T = Ts;
    for (uint32_t c = 0; c < n_ch && n_dense; c++)
        if (t->h_ch_dense[c]) {
            t->h_ch_lo[c] = T;
            T += t->h_ch_edges[c];
        }
    const double t_pre = now_ms();
    rc.pre_ms += t_pre - t_begin;
    rc.node_sum += m;
    rc.stream_rounds++;

    PartArgs<EdgeT> p{};
    p.list = t->od_list;
    p.prefix = t->od_prefix;
    p.n = m;
    p.tail_base = S;
    p.pk_shift = g->pk_shift;
    p.pk_mask = pk && !g->pk_split ? (1u << g->pk_shift) - 1 : 0xFFFFFFFFu;
    p.pk_wmin = g->pk_wmin;
    p.pk_split = g->pk_split ? 1u : 0u;
    const uint64_t PE = t->part_edges;
    // mapped device addresses of the pinned tail (zero-copy and hybrid reads)
    auto map_tail = [&]() -> int {
        void* dptr = nullptr;
        ASC_HIP(hipHostGetDevicePointer(&dptr, pk ? (void*)g->ptail.p : (void*)g->zc_tail.p, 0));
        p.zc = (const EdgeT*)dptr;
        p.zc32 = (const uint32_t*)dptr;
        p.zc_w8 = nullptr;
        if (g->pk_split) {
            void* wptr = nullptr;
            ASC_HIP(hipHostGetDevicePointer(&wptr, (void*)g->ptail_w.p, 0));
            p.zc_w8 = (const uint8_t*)wptr;
        }
        return ASC_OK;
    };
    auto launch_part = [&](int src) {
        const int pgrid = (int)std::max<uint64_t>(1, std::min<uint64_t>(grid, (p.hi - p.lo + kTile - 1) / kTile));
        return timed_launch(t, prof, src == PART_HOST ? KERN_PART_HOST : KERN_PART, s, [&] {
            const dim3 gd(pgrid), bd(kBlock);
            if (src == PART_HOST && pk)
                hipLaunchKernelGGL((k_expand_part<ALGO, OffT, EdgeT, PART_HOST, true>), gd, bd, 0, s, a, level, p);
            else if (src == PART_HOST)
                hipLaunchKernelGGL((k_expand_part<ALGO, OffT, EdgeT, PART_HOST, false>), gd, bd, 0, s, a, level, p);
            else if (src == PART_CHUNK && pk)
                hipLaunchKernelGGL((k_expand_part<ALGO, OffT, EdgeT, PART_CHUNK, true>), gd, bd, 0, s, a, level, p);
            else if (src == PART_CHUNK)
                hipLaunchKernelGGL((k_expand_part<ALGO, OffT, EdgeT, PART_CHUNK, false>), gd, bd, 0, s, a, level, p);
            else if (pk)
                hipLaunchKernelGGL((k_expand_part<ALGO, OffT, EdgeT, PART_GATHERED, true>), gd, bd, 0, s, a, level,
                                   p);
            else
                hipLaunchKernelGGL((k_expand_part<ALGO, OffT, EdgeT, PART_GATHERED, false>), gd, bd, 0, s, a, level,
                                   p);
        });
    };
    // SSSP/CC with stream_overlap: after each part or chunk, one resident step
    // expands what it pushed while the next transfer is in flight.  Each part
    // and each step takes its own level, so a resident vertex improved again
    // after its step is queued again (stamps are per level); BFS keeps one
    // level per round because its values are levels.
    const bool overlap = ALGO != ALG_BFS && ctx->cfg.stream_overlap > 0;
    const int every = std::max(1, ctx->cfg.stream_overlap);
    int since = 0;   // parts or chunks launched since the last resident step
    auto resident_step = [&]() -> int {
        if (!overlap || ++since < every) return ASC_OK;
        since = 0;
        level++;
        a.q_in = t->q[level & 1];
        a.q_out = t->q[(level + 1) & 1];
        ASC_TRY(timed_launch(t, prof, KERN_LOCAL, s, [&] {
            hipLaunchKernelGGL((k_expand_local<ALGO, OffT, EdgeT>), dim3(grid), dim3(kBlock), 0, s, a, level);
        }));
        return timed_launch(t, prof, KERN_HEAVY, s, [&] {
            hipLaunchKernelGGL((k_expand_heavy<ALGO, OffT, EdgeT>), dim3(grid), dim3(kBlock), 0, s, a, level);
        });
    };
    uint64_t dense_edges = 0;
    double fill_round = 0;     // host copy time: staged chunks and the gather
    double gather_round = 0;   // the gather alone (its rate estimate)
    if (ctx->cfg.stream_mode == ASC_STREAM_ZEROCOPY) {
        ASC_TRY(map_tail());
        p.lo = 0;
        p.hi = T;
        ASC_TRY(launch_part(PART_HOST));
        rc.parts += 1;
        rc.streamed_edges += T;
    } else {
        uint32_t nb = 0;   // device part buffers are used round-robin by both loops
        // dense chunks: whole-chunk DMA from the pinned tail, then the chunk's
        // list segment expanded with edges addressed from the chunk start.
        // Chunks already in a part buffer go first (no transfer to wait for),
        // then the others in id order, into buffers in turn.
        std::vector<uint32_t> dorder;
        for (uint32_t c = 0; c < n_ch && n_dense; c++)
            if (t->h_ch_dense[c] && buf_of[c] >= 0) dorder.push_back(c);
        const size_t n_cached = dorder.size();
        for (uint32_t c = 0; c < n_ch && n_dense; c++)
            if (t->h_ch_dense[c] && buf_of[c] < 0) dorder.push_back(c);
        for (size_t oi = 0; oi < dorder.size(); oi++) {
            const uint32_t c = dorder[oi];
            const uint64_t e0 = t->ch_e[c], e1 = t->ch_e[c + 1];
            int b;
            if (oi < n_cached) {
                b = buf_of[c];
                if (t->buf_counted[b]) rc.reused_edges += e1 - e0;   // copied by an earlier round
                else dense_edges += e1 - e0;                          // copied ahead for this round
            } else {
                // the chunks read in place were all launched already, so any
                // buffer may be overwritten once its last reader is done
                b = (int)(nb++ % t->n_bufs);
                ASC_TRY(wait_event(ctx->s_copy, t->ev_kern[b]));   // last reader of part b done
                ASC_TRY(copy_chunk_to_part(g, t, b, e0, e1, ctx->s_copy, &fill_round));
                ASC_HIP(hipEventRecord(t->ev_copy[b], ctx->s_copy));
                dense_edges += e1 - e0;
                if (t->buf_chunk[b] >= 0) buf_of[t->buf_chunk[b]] = -1;
                t->buf_chunk[b] = (int)c;
                buf_of[c] = b;
            }
            t->buf_counted[b] = true;
            ASC_TRY(wait_event(s, t->ev_copy[b]));
            p.zc = (const EdgeT*)t->part[b];
            p.zc32 = (const uint32_t*)t->part[b];
            p.zc_w8 = (const uint8_t*)t->part[b] + 4 * PE;
            p.tail_base = e0;
            p.lo = t->h_ch_lo[c];
            p.hi = t->h_ch_lo[c] + t->h_ch_edges[c];
            if (p.hi > p.lo) ASC_TRY(launch_part(PART_CHUNK));
            ASC_HIP(hipEventRecord(t->ev_kern[b], s));
            ASC_TRY(resident_step());
            if (ctx->cfg.verbose > 2)
                printf("  chunk %u (buffer %d%s) enqueued at %.3f ms\n", c, b, oi < n_cached ? ", held" : "",
                       now_ms() - rc.t_start);
        }
        if (ctx->cfg.stream_mode == ASC_STREAM_HYBRID) {
            // the rest read by the kernel from the mapped pinned tail
            if (Ts > 0) {
                ASC_TRY(map_tail());
                p.tail_base = S;
                p.lo = 0;
                p.hi = Ts;
                ASC_TRY(launch_part(PART_HOST));
                nb++;
            }
            rc.parts += nb;
            rc.streamed_edges += Ts + dense_edges;
            rc.dense_edges += dense_edges;
            rc.overload_ms += now_ms() - t_begin;
            return ASC_OK;
        }
        if (zc_sparse && Ts > 0) {
            // the small sparse segment read from the mapped page-locked tail
            ASC_TRY(map_tail());
            p.tail_base = S;
            p.lo = 0;
            p.hi = Ts;
            ASC_TRY(launch_part(PART_HOST));
            ASC_TRY(resident_step());
            rc.parts += nb + 1;
            rc.streamed_edges += Ts + dense_edges;
            rc.dense_edges += dense_edges;
            rc.fill_ms += fill_round;
            rc.overload_ms += now_ms() - t_begin;
            return ASC_OK;
        }
        // the rest: host gather into pinned staging, double-buffered copies
        const uint64_t np = (Ts + PE - 1) / PE;
        for (uint64_t q = 0; q < np; q++) {
            const int sb = (int)(t->stage_seq++ & 1);
            const int b = (int)(nb++ % t->n_bufs);
            const uint64_t lo = q * PE, hi = std::min(Ts, lo + PE);
            ASC_HIP(hipEventSynchronize(t->ev_stage[sb]));  // staging sb drained
            const double tg = now_ms();
            char* stage_w = t->stage[sb].p + 4 * PE;   // split form: weight bytes behind the ids
            gather_part(g, t->h_src.p, t->h_prefix.p, n_sparse, lo, hi, t->stage[sb].p, stage_w);
            const double dt = now_ms() - tg;
            fill_round += dt;
            gather_round += dt;
            ASC_TRY(wait_event(ctx->s_copy, t->ev_kern[b]));  // device part b free
            t->buf_chunk[b] = -1;                                       // now holds gathered records
            ASC_TRY(copy_to_part(g, t, b, t->stage[sb].p, stage_w, hi - lo, ctx->s_copy));
            ASC_HIP(hipEventRecord(t->ev_stage[sb], ctx->s_copy));
            ASC_HIP(hipEventRecord(t->ev_copy[b], ctx->s_copy));
            ASC_HIP(hipStreamWaitEvent(s, t->ev_copy[b], 0));
            p.part = (const EdgeT*)t->part[b];
            p.part32 = (const uint32_t*)t->part[b];
            p.part_w8 = (const uint8_t*)t->part[b] + 4 * PE;
            p.tail_base = S;
            p.lo = lo;
            p.hi = hi;
            ASC_TRY(launch_part(PART_GATHERED));
            ASC_HIP(hipEventRecord(t->ev_kern[b], s));
            ASC_TRY(resident_step());
            if (ctx->cfg.verbose > 2)
                printf("  part %llu (buffer %d) gathered in %.3f ms, enqueued at %.3f ms\n", (unsigned long long)q, b,
                       dt, now_ms() - rc.t_start);
        }
        rc.fill_ms += fill_round;
        rc.parts += nb;
        rc.streamed_edges += Ts + dense_edges;
        rc.dense_edges += dense_edges;
        if (Ts * EB >= (32u << 20) && gather_round > 0)   // running estimate of the gather rate
            t->gather_bps = 0.5 * t->gather_bps + 0.5 * (double)(Ts * EB) / (gather_round * 1e-3);
    }
    rc.overload_ms += now_ms() - t_begin;
    if (ctx->cfg.verbose > 1)
        printf("stream round %u at %.3f ms: pile %u, %u dense chunks %llu edges, gathered %u vertices %llu edges, "
               "%.3f ms (prepare %.3f)\n",
               rc.stream_rounds, t_begin - rc.t_start, npile, n_dense, (unsigned long long)dense_edges, n_sparse,
               (unsigned long long)t->h_prefix[n_sparse], now_ms() - t_begin, t_pre - t_begin);
    return ASC_OK;
}

template <int ALGO, class OffT, class EdgeT>
static int run_traversal(asc_graph* g, uint32_t src, uint32_t* out, asc_stats* st) {
    asc_ctx* ctx = g->ctx;
    ASC_TRY(trav_ensure(g));
    TravState* t = g->trav;
    const uint64_t V = g->h.V;
    const uint32_t vs = (uint32_t)g->static_vertices;
    hipStream_t s = ctx->s_main;
    ASC_HIP(hipSetDevice(ctx->device));
    const int fill_grid = ctx->num_cus * 4;
    const int grid = ctx->num_cus * (g->blocks_per_cu > 0 ? g->blocks_per_cu : kBlocksPerCU);
    const bool prof = ctx->cfg.profile != 0;
    const uint32_t delta = ALGO == ALG_SSSP ? sssp_delta(g) : 0;
    const bool use16 = ALGO == ALG_SSSP && vs == V && sssp_filter_pays(g);

    // initial state
    ASC_HIP(hipMemsetAsync(t->slots, 0, sizeof(Slot) * 4, s));
    ASC_HIP(hipMemsetAsync(t->piles, 0, sizeof(Piles), s));
    ASC_HIP(hipMemsetAsync(t->stats, 0, sizeof(DevStats), s));
    if (ALGO == ALG_BFS) {
        hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid), dim3(256), 0, s, t->val, V, (uint32_t)(V + 1));
        hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid), dim3(256), 0, s, t->visited, (V + 31) / 32, 0u);
        hipLaunchKernelGGL(k_init_source, dim3(1), dim3(1), 0, s, t->val, t->visited, (uint32_t*)nullptr, t->q[0],
                           t->pend[0], t->slots, t->piles, src, vs);
    } else {
        for (uint32_t* p : {t->stamp, t->pstamp, t->fstamp})
            hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid), dim3(256), 0, s, p, V, 0u);
        hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid), dim3(256), 0, s, t->done, V, kNoValue);
        if (ALGO == ALG_CC) {
            hipLaunchKernelGGL(k_init_cc, dim3(fill_grid), dim3(256), 0, s, t->val, t->pstamp, t->q[0], t->pend[0],
                               t->slots, t->piles, (uint32_t)V, vs);
        } else {
            hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid), dim3(256), 0, s, t->val, V, (uint32_t)(V + 1));
            if (use16) {
                // the value filter serves fully resident graphs (narrow bands:
                // few improvements per vertex); with on-demand edges (wide
                // bands, more stores to it) it measured slower
                if (!t->val16) ASC_HIP(hipMalloc(&t->val16, 4 * ((V + 1) / 2)));
                hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid), dim3(256), 0, s, (uint32_t*)t->val16, (V + 1) / 2,
                                   0xFFFFFFFFu);
            }
            hipLaunchKernelGGL(k_init_source, dim3(1), dim3(1), 0, s, t->val, (uint32_t*)nullptr, t->pstamp, t->q[0],
                               t->pend[0], t->slots, t->piles, src, vs);
        }
    }
    ASC_HIP(hipGetLastError());
    ASC_HIP(hipStreamSynchronize(s));

    TravArgs<OffT, EdgeT> a{};
    a.off = (const OffT*)g->d_off;
    a.edges = (const EdgeT*)g->d_edges;
    a.val = t->val;
    a.val16 = use16 ? t->val16 : nullptr;
    a.stamp = t->stamp;
    a.pstamp = t->pstamp;
    a.fstamp = t->fstamp;
    a.done = t->done;
    a.visited = t->visited;
    a.slots = t->slots;
    a.heavy_u = t->heavy_u;
    a.heavy_sv = t->heavy_sv;
    a.heavy_base = t->heavy_base;
    a.stats = t->stats;
    a.vs = vs;
    a.heavy_thr = (uint32_t)effective_heavy_threshold(g);
    a.theta = (ALGO == ALG_SSSP && delta) ? 1 + delta : 0xFFFFFFFFu;
    a.pepoch = 1;
    a.fepoch = 1;
    uint32_t pcur = 0, fcur = 0;

    uint32_t nr, npend, nfar = 0;
    if (ALGO == ALG_CC) {
        nr = vs;
        npend = (uint32_t)(V - vs);
    } else {
        nr = src < vs ? 1 : 0;
        npend = src < vs ? 0 : 1;
    }
    RunCounters rc{};
    t->timed.clear();
    t->ev_next = 0;
    const double t0 = now_ms();
    rc.t_start = t0;
    if (g->static_vertices < V && t->part[0]) ASC_TRY(prefetch_chunks(g, t));
    uint32_t level = 0;
    bool last_split_empty = false;
    for (;;) {
        enum { M_LEVEL, M_RES, M_STREAM, M_SPLIT } mode;
        if (ALGO == ALG_BFS) {
            if (nr == 0 && npend == 0) break;
            mode = M_LEVEL;
        } else if (nr > 0) {
            mode = M_RES;
        } else if (npend > 0) {
            mode = M_STREAM;
        } else if (nfar > 0) {
            mode = M_SPLIT;
        } else {
            break;
        }
        const uint32_t* stream_list = nullptr;
        uint32_t nstream = 0;
        if (mode == M_LEVEL || mode == M_STREAM) {
            if (npend > 0) {  // the pile filled so far is streamed; new entries go to the other one
                stream_list = t->pend[pcur];
                nstream = npend;
                pcur ^= 1;
                a.pepoch++;
                hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, &t->piles->n_pend[pcur], 0u);
            }
        }
        a.pend_out = t->pend[pcur];
        a.pend_count = &t->piles->n_pend[pcur];
        a.far_out = t->far[fcur];
        a.far_count = &t->piles->n_far[fcur];
        if (mode == M_SPLIT) {
            // advance the band; entries still beyond it move to the other far pile
            const uint64_t th = (uint64_t)(last_split_empty ? std::max(a.theta, t->h_piles[0].min_far) : a.theta) + delta;
            a.theta = (uint32_t)std::min<uint64_t>(th, 0xFFFFFFFFull);
            const uint32_t old = fcur;
            fcur ^= 1;
            a.far_out = t->far[fcur];
            a.far_count = &t->piles->n_far[fcur];
            hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, &t->piles->n_far[fcur], 0u);
            hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, &t->piles->min_far, 0xFFFFFFFFu);
            hipLaunchKernelGGL((k_split_far<OffT, EdgeT>),
                               dim3(std::min<uint32_t>((nfar + kBlock - 1) / kBlock, (uint32_t)grid)), dim3(kBlock), 0,
                               s, a, t->far[old], nfar, level, t->q[level & 1], t->piles);
            ASC_HIP(hipGetLastError());
            rc.splits++;
        } else {
            // one step: expand the resident frontier (possibly empty), then
            // stream the pile
            rc.node_sum += nr;
            a.q_in = t->q[level & 1];
            a.q_out = t->q[(level + 1) & 1];
            hipEvent_t e0 = nullptr;
            if (!prof) {
                e0 = next_event(t);
                if (!e0) return ASC_ERR_DEVICE;
                ASC_HIP(hipEventRecord(e0, s));
            }
            // one workgroup per 256 frontier entries, at most `grid`
            const int lgrid = (int)std::max<uint64_t>(1, std::min<uint64_t>(grid, ((uint64_t)nr + kBlock - 1) / kBlock));
            ASC_TRY(timed_launch(t, prof, KERN_LOCAL, s, [&] {
                hipLaunchKernelGGL((k_expand_local<ALGO, OffT, EdgeT>), dim3(lgrid), dim3(kBlock), 0, s, a, level);
            }));
            if (nr) {
                ASC_TRY(timed_launch(t, prof, KERN_HEAVY, s, [&] {
                    hipLaunchKernelGGL((k_expand_heavy<ALGO, OffT, EdgeT>), dim3(grid), dim3(kBlock), 0, s, a, level);
                }));
            }
            if (!prof) {
                hipEvent_t e1 = next_event(t);
                if (!e1) return ASC_ERR_DEVICE;
                ASC_HIP(hipEventRecord(e1, s));
                t->timed.push_back({-1, e0, e1});
            }
            if (nstream) ASC_TRY((stream_pile<ALGO, OffT, EdgeT>(g, a, level, stream_list, nstream, grid, rc)));
            level++;
        }
        const unsigned long long want = ++t->lvl_seq;
        hipLaunchKernelGGL(k_trav_publish, dim3(1), dim3(1), 0, s, t->slots + (level & 3), t->piles, t->d_pub, want);
        ASC_HIP(hipGetLastError());
        ASC_TRY(wait_host_seq(&t->h_pub.p->seq, want, s));
        {
            const volatile LevelPub* pb = t->h_pub.p;
            nr = pb->slot.n_res;
            Piles pl;
            pl.n_pend[0] = pb->piles.n_pend[0];
            pl.n_pend[1] = pb->piles.n_pend[1];
            pl.n_far[0] = pb->piles.n_far[0];
            pl.n_far[1] = pb->piles.n_far[1];
            pl.n_list = pb->piles.n_list;
            pl.min_far = pb->piles.min_far;
            pl.pad[0] = pl.pad[1] = 0;
            t->h_piles[0] = pl;   // min_far is read when the band advances
        }
        npend = t->h_piles[0].n_pend[pcur];
        nfar = ALGO == ALG_SSSP ? t->h_piles[0].n_far[fcur] : 0;
        last_split_empty = mode == M_SPLIT && nr == 0 && npend == 0;
    }
    const double t1 = now_ms();
    return finish_traversal(g, ALGO, PATH_FRONTIER, out, st, rc, t1 - t0, level);
}

int finish_traversal(asc_graph* g, int algo, int path, uint32_t* out, asc_stats* st, const RunCounters& rc,
                     double total_ms, uint32_t iterations) {
    asc_ctx* ctx = g->ctx;
    TravState* t = g->trav;
    const uint64_t V = g->h.V;
    const bool prof = ctx->cfg.profile != 0;
    double static_ms = 0;
    asc_kernel_stats kern[KERN_COUNT];
    std::memset(kern, 0, sizeof(kern));
    for (auto& x : t->timed) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, x.a, x.b) != hipSuccess) continue;
        if (x.kind < 0) {
            static_ms += ms;
        } else {
            kern[x.kind].ms += ms;
            kern[x.kind].launches++;
        }
    }
    if (prof)
        static_ms = kern[KERN_LOCAL].ms + kern[KERN_HEAVY].ms + kern[KERN_BU].ms + kern[KERN_UF].ms +
                    kern[KERN_PERSIST].ms;
    ASC_HIP(hipMemcpy(t->h_stats.p, t->stats, sizeof(DevStats), hipMemcpyDeviceToHost));
    const DevStats& ds = t->h_stats[0];
    kern[KERN_LOCAL].edges = ds.edges_local;
    kern[KERN_LOCAL].vertices = ds.verts_local;
    kern[KERN_HEAVY].edges = ds.edges_heavy;
    kern[KERN_PART].edges = ds.edges_part;
    kern[KERN_PART].vertices = ds.verts_part;
    kern[KERN_PART_HOST].edges = ds.edges_part_host;
    kern[KERN_PART_HOST].vertices = ds.verts_part_host;
    kern[KERN_BU].edges = ds.edges_bu;
    kern[KERN_BU].vertices = ds.verts_bu;
    kern[KERN_UF].edges = ds.edges_uf;
    kern[KERN_UF].vertices = path == PATH_UNION_FIND ? V : 0;
    if (out) ASC_HIP(hipMemcpy(out, t->val, 4 * V, hipMemcpyDeviceToHost));

    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->total_ms = total_ms;
        st->pre_ms = rc.pre_ms;
        st->static_ms = static_ms;
        st->kernel_ms = static_ms + kern[KERN_PART].ms + kern[KERN_PART_HOST].ms;
        st->overload_ms = rc.overload_ms;
        st->fill_ms = rc.fill_ms;
        st->iterations = iterations;
        st->parts = rc.parts;
        st->node_sum = rc.node_sum;
        st->static_edges_traversed = ds.edges_local + ds.edges_heavy + ds.edges_bu + ds.edges_uf;
        st->edges_traversed = st->static_edges_traversed + ds.edges_part + ds.edges_part_host;
        st->streamed_edges = rc.streamed_edges;
        st->streamed_bytes = rc.streamed_edges * g->stream_record_bytes();
        st->dense_edges = rc.dense_edges;
        st->reused_edges = rc.reused_edges;
        st->vertices_processed = ds.verts_local + ds.verts_bu;
        st->host_cores = hardware_threads();
        st->path = path;
        std::memcpy(st->kern, kern, sizeof(kern));
    }
    if (out) {
        // reached vertices and their degree sum (Graph500 TEPS numerator)
        const uint32_t unreached = (uint32_t)(V + 1);
        std::atomic<uint64_t> reached{0}, dsum{0};
        const uint64_t* off = g->h.off.p;
        const int conv = ctx->cfg.value_convention;
        host_pool().parallel_for(V, [&](uint64_t b, uint64_t e) {
            uint64_t r = 0, d = 0;
            for (uint64_t v = b; v < e; v++) {
                if (algo == ALG_CC || out[v] != unreached) {
                    r++;
                    d += off[v + 1] - off[v];
                }
                if (algo != ALG_CC && conv == ASC_VALUES_ZERO_BASED)
                    out[v] = out[v] == unreached ? 0xFFFFFFFFu : out[v] - 1;
            }
            reached += r;
            dsum += d;
        });
        if (st) {
            st->reached = reached;
            st->reached_degree_sum = dsum;
            // BFS is level-synchronous: each reached vertex is active in
            // exactly one iteration, so the reference's nodeSum (actives summed
            // over iterations, CalculateOpt.cu:183-190) is the reached count,
            // whichever direction each level took
            if (algo == ALG_BFS) st->node_sum = reached;
        }
    }
    if (ctx->cfg.verbose && st) {
        printf("totalProcess time is %.3f ms\npreProcess time is %.3f ms\nstaticProcess time is %.3f ms\n"
               "overloadProcess time is %.3f ms\nfillProcess time is %.3f ms\nnodeSum : %llu\n"
               "move overload size : %llu\n",
               st->total_ms, st->pre_ms, st->static_ms, st->overload_ms, st->fill_ms,
               (unsigned long long)st->node_sum, (unsigned long long)st->streamed_bytes);
    }
    return ASC_OK;
}

int traverse(asc_graph* g, int algo, uint32_t src, uint32_t* out, asc_stats* st) {
    const asc_config& cfg = g->ctx->cfg;
    if (algo == ALG_BFS && cfg.bfs_direction == ASC_BFS_AUTO && resident_symmetric(g))
        return bfs_direction_opt(g, src, out, st);
    if (algo == ALG_CC && cfg.cc_method == ASC_CC_AUTO && resident_symmetric(g)) return cc_union_find(g, out, st);
    const bool w = g->h.kind == ASC_KIND_WCSR;
    const bool o4 = g->off_bytes == 4;
    switch (algo) {
        case ALG_BFS:
            if (w) return o4 ? run_traversal<ALG_BFS, uint32_t, WEdge>(g, src, out, st)
                             : run_traversal<ALG_BFS, uint64_t, WEdge>(g, src, out, st);
            return o4 ? run_traversal<ALG_BFS, uint32_t, uint32_t>(g, src, out, st)
                      : run_traversal<ALG_BFS, uint64_t, uint32_t>(g, src, out, st);
        case ALG_SSSP:
            if (!w) return ASC_ERR_UNSUPPORTED;
            return o4 ? run_traversal<ALG_SSSP, uint32_t, WEdge>(g, src, out, st)
                      : run_traversal<ALG_SSSP, uint64_t, WEdge>(g, src, out, st);
        case ALG_CC:
            if (w) return o4 ? run_traversal<ALG_CC, uint32_t, WEdge>(g, src, out, st)
                             : run_traversal<ALG_CC, uint64_t, WEdge>(g, src, out, st);
            return o4 ? run_traversal<ALG_CC, uint32_t, uint32_t>(g, src, out, st)
                      : run_traversal<ALG_CC, uint64_t, uint32_t>(g, src, out, st);
    }
    return ASC_ERR_INVALID;
}

}  // namespace asc
