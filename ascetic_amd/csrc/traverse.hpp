// Shared declarations of the frontier traversal engine.
#pragma once

#include "device_common.hpp"
#include "internal.hpp"

namespace asc {

enum { ALG_BFS = 0, ALG_SSSP = 1, ALG_CC = 2 };

constexpr int kStageCap = 128;                  // per-wave LDS staging entries, per queue
constexpr int kHeavyShift = 44;                 // heavy list counter: (count << 44) | edges
constexpr unsigned long long kHeavyMask = (1ull << kHeavyShift) - 1;
constexpr uint32_t kHeavyCap = 1u << 20;        // heavy list entries per level
constexpr int kTile = 1024;                     // edges per tile in edge-balanced kernels
constexpr int kMinHeavy = 256;                  // smallest heavy threshold (tile spans <= 5 entries)
constexpr int kBlocksPerCU = 8;
constexpr int kMaxPartBufs = 16;              // device buffers for streamed parts and chunks (round-robin)
constexpr int kDefaultPartBufs = 4;
constexpr uint32_t kNoValue = 0xFFFFFFFFu;      // `done` before a vertex was ever expanded
constexpr uint32_t kMaxChunks = 4096;           // dense on-demand chunks (LDS histogram size)

// Per-level counters, in a ring of four: level L reads slot L%4, appends to
// slot (L+1)%4 and clears slot (L+2)%4.
struct Slot {
    uint32_t n_res;                 // resident frontier entries
    uint32_t pad0;
    unsigned long long heavy;       // heavy list (count << 44 | edges)
    unsigned long long pad[2];
};

// Counters of the on-demand pile and the far pile (double-buffered).
struct Piles {
    uint32_t n_pend[2];
    uint32_t n_far[2];
    uint32_t n_list;                // entries kept by the pile filter
    uint32_t min_far;               // smallest value left in the far pile
    uint32_t pad[2];
};

struct DevStats {
    unsigned long long edges_local;   // k_expand_local
    unsigned long long verts_local;
    unsigned long long edges_heavy;   // k_expand_heavy
    unsigned long long edges_part;    // k_expand_part reading a device part buffer
    unsigned long long verts_part;
    unsigned long long edges_part_host;   // k_expand_part reading mapped host memory
    unsigned long long verts_part_host;
    unsigned long long edges_bu;      // k_bfs_bottom_up: edge records read
    unsigned long long verts_bu;      //   unvisited vertices scanned
    unsigned long long edges_uf;      // union-find link passes: edge records read
};

enum {
    KERN_LOCAL = 0, KERN_HEAVY = 1, KERN_PART = 2, KERN_PR = 3, KERN_BU = 4, KERN_UF = 5, KERN_PERSIST = 6,
    KERN_PART_HOST = 7, KERN_COUNT = 8
};

// asc_stats.path
enum { PATH_FRONTIER = 0, PATH_DIRECTION_OPT = 1, PATH_UNION_FIND = 2 };

// Counters of the resident fast paths (direction-optimizing BFS, union-find CC).
struct DoCounters {
    unsigned long long n_found;       // vertices reached by a bottom-up step
    unsigned long long m_found;       // their summed degrees
    unsigned long long m_frontier;    // summed degrees of a top-down step's output
    unsigned long long pad;
};

template <class OffT, class EdgeT>
struct TravArgs {
    const OffT* off;                // V+1 offsets
    const EdgeT* edges;             // static edge prefix
    uint32_t* val;
    // SSSP on a fully resident graph (else null): 16-bit value filter
    // probed before val.  An entry other than 0xFFFF is a value val[v] held
    // at some moment (stored after the atomicMin that set it, so a late
    // store can leave a larger one): never below val[v], so a candidate not
    // below it cannot lower val[v].  0xFFFF (not lowered yet, or a value
    // beyond 16 bits): the atomic is tried.  Half the bytes of val, so more
    // of the probed vertices stay in cache.
    uint16_t* val16;
    uint32_t* stamp;                // SSSP/CC: level a resident vertex was last queued for
    uint32_t* pstamp;               // SSSP/CC: on-demand pile epoch a vertex was last added to
    uint32_t* fstamp;               // SSSP: set (to fepoch, constant in a run) once a vertex enters the far pile
    uint32_t* done;                 // SSSP/CC: value a vertex was last expanded with
    uint32_t* visited;              // BFS bitmap
    const uint32_t* q_in;
    uint32_t* q_out;
    uint32_t* pend_out;             // on-demand pile being filled
    uint32_t* pend_count;
    uint32_t* far_out;              // far pile being filled
    uint32_t* far_count;
    Slot* slots;
    uint32_t* heavy_u;
    uint32_t* heavy_sv;
    unsigned long long* heavy_base;
    DevStats* stats;
    // BFS, when set: the expansion adds the degrees of the vertices it
    // claims here (the next frontier's edges, for the direction decision)
    unsigned long long* m_next;
    uint32_t vs;                    // vertices [0, vs) are resident
    uint32_t heavy_thr;
    uint32_t theta;                 // SSSP near/far boundary (UINT_MAX: everything near)
    uint32_t pepoch, fepoch;
};

template <class EdgeT>
struct PartArgs {
    const uint32_t* list;           // on-demand vertices (no zero degree)
    const unsigned long long* prefix;  // exclusive edge starts, prefix[n] = total
    uint32_t n;
    unsigned long long lo, hi;      // edge range of this part
    const EdgeT* part;              // device buffer holding edges [lo, hi)
    const EdgeT* zc;                // zero-copy: mapped host edges from tail_base
    const uint32_t* part32;         // packed stream: device buffer of u32 words
    const uint32_t* zc32;           // packed stream, zero-copy
    const uint8_t* part_w8;         // split stream: weight bytes beside part32
    const uint8_t* zc_w8;           // split stream: weight bytes beside zc32
    unsigned long long tail_base;
    uint32_t pk_shift, pk_mask, pk_wmin;  // word = dst | (w - pk_wmin) << pk_shift
    uint32_t pk_split;              // 1: word = dst, weight byte = w - pk_wmin
};

struct RunCounters {
    double pre_ms = 0, overload_ms = 0, fill_ms = 0;
    uint64_t node_sum = 0, streamed_edges = 0, dense_edges = 0;
    uint64_t reused_edges = 0;      // edges of chunks read again from a part buffer (no copy)
    uint32_t parts = 0, stream_rounds = 0, splits = 0, dense_chunks = 0;
    double t_start = 0;             // host clock at the start of the run (per-round log)
};

// A level's counters as the device publishes them to mapped host memory; seq
// is written last, so a host that reads the expected seq finds the rest final.
struct LevelPub {
    Slot slot;
    Piles piles;
    unsigned long long seq;
};

// Waits until the device has written `want` to *seq (mapped host memory),
// i.e. the kernels before the publishing one have ended, by polling instead
// of synchronizing the stream (which adds tens of microseconds).  After 50 ms
// of polling it synchronizes the stream instead, which also reports a failed
// launch.
inline int wait_host_seq(const volatile unsigned long long* seq, unsigned long long want, hipStream_t s) {
    const double t0 = now_ms();
    for (uint32_t i = 0; *seq != want; i++) {
        if ((i & 1023) == 1023 && now_ms() - t0 > 50.0) {
            ASC_HIP(hipStreamSynchronize(s));
            if (*seq != want) {
                set_last_error("the device did not publish its level counters");
                return ASC_ERR_DEVICE;
            }
            break;
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return ASC_OK;
}

struct TravState {
    uint64_t V = 0;
    uint32_t* val = nullptr;
    uint16_t* val16 = nullptr;        // SSSP value filter (TravArgs::val16)
    uint32_t* stamp = nullptr;
    uint32_t* pstamp = nullptr;
    uint32_t* fstamp = nullptr;
    uint32_t* done = nullptr;
    uint32_t* visited = nullptr;
    uint32_t* q[2] = {nullptr, nullptr};
    uint32_t* pend[2] = {nullptr, nullptr};
    uint32_t* far[2] = {nullptr, nullptr};
    uint32_t* heavy_u = nullptr;
    uint32_t* heavy_sv = nullptr;
    unsigned long long* heavy_base = nullptr;
    uint32_t* od_list = nullptr;
    unsigned long long* od_prefix = nullptr;
    unsigned long long* scan_tmp = nullptr;   // tile sums of the device prefix scan
    Slot* slots = nullptr;
    Piles* piles = nullptr;
    DevStats* stats = nullptr;
    void* part[kMaxPartBufs] = {};  // device part buffers, used round-robin (n_bufs of them)
    int n_bufs = kDefaultPartBufs;
    // On-demand chunk whose edges part buffer b holds whole (-1: none), copied
    // ahead at the start of the run or by an earlier round of it; a round
    // that needs the chunk reads it there instead of copying it again.
    // buf_counted[b]: that copy already counts in the run's streamed edges.
    int buf_chunk[kMaxPartBufs] = {};
    bool buf_counted[kMaxPartBufs] = {};
    uint64_t part_edges = 0;
    HostBuf<Slot> h_slot;
    HostBuf<Piles> h_piles;
    HostBuf<LevelPub> h_pub;        // written by the device through d_pub (mapped)
    LevelPub* d_pub = nullptr;
    unsigned long long lvl_seq = 0;
    HostBuf<DevStats> h_stats;
    // the host gather's index of a round's sparse segment: first record of
    // each listed vertex in the gathered array (written by the device through
    // d_hsrc, mapped) and the u64 exclusive prefix of their degrees
    HostBuf<unsigned long long> h_src, h_prefix;
    unsigned long long* d_hsrc = nullptr;
    HostBuf<char> stage[2];
    uint64_t stage_seq = 0;         // uses of the two staging buffers (gather and staged chunk copies)
    // dense on-demand chunks: vertex ranges [ch_v[i], ch_v[i+1]) whose edges
    // [ch_e[i], ch_e[i+1]) fit one part buffer (empty: dense copies off)
    std::vector<uint32_t> ch_v;
    std::vector<uint64_t> ch_e;
    uint32_t* d_ch_v = nullptr;
    unsigned long long* d_ch_edges = nullptr;   // active edges per chunk, this round
    uint32_t* d_ch_count = nullptr;             // active vertices per chunk, this round
    uint8_t* d_ch_dense = nullptr;              // chunk copied whole this round
    uint32_t* d_cursor = nullptr;               // list segment cursors (n_ch + 1)
    HostBuf<unsigned long long> h_ch_edges;
    std::vector<uint64_t> h_ch_lo;  // start of each dense chunk's segment in the list's edge prefix
    HostBuf<uint32_t> h_ch_count, h_cursor;
    HostBuf<uint8_t> h_ch_dense;
    // rates the dense-chunk choice balances (bytes/s): pinned host-to-device
    // copies (measured once) and the host gather (running estimate)
    double link_bps = 0, gather_bps = 30e9;
    // resident fast paths: frontier bitmaps (the persistent BFS keeps three in
    // a ring), counters, sampled labels, the persistent kernel's control words
    uint32_t* bm[3] = {nullptr, nullptr, nullptr};
    uint32_t* pctl = nullptr;
    DoCounters* dcnt = nullptr;
    HostBuf<DoCounters> h_dcnt;     // written by the device through d_hdcnt (mapped): [0] the
    DoCounters* d_hdcnt = nullptr;  // counters, [1].n_found the sequence number of the last publish
    unsigned long long pub_seq = 0;
    HostBuf<uint32_t> h_sample;
    // device part buffer b: copy into it done / last kernel reading it done;
    // staging buffer i: copy out of it done
    hipEvent_t ev_copy[kMaxPartBufs] = {}, ev_kern[kMaxPartBufs] = {}, ev_meta = nullptr;
    hipEvent_t ev_stage[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> ev_pool;
    // event pairs recorded during the current run: (kernel kind, start, end)
    struct Timed {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Timed> timed;
    size_t ev_next = 0;
};

int traverse(asc_graph* g, int algo, uint32_t src, uint32_t* out, asc_stats* st);
int trav_ensure(asc_graph* g);
hipEvent_t next_event(TravState* t);

// Launches through `launch`; when profiling, brackets it with an event pair on
// the launch stream and records the pair under `kind`.
template <class F>
int timed_launch(TravState* t, bool prof, int kind, hipStream_t s, F&& launch) {
    hipEvent_t a = nullptr, b = nullptr;
    if (prof) {
        a = next_event(t);
        b = next_event(t);
        if (!a || !b) return ASC_ERR_DEVICE;
        ASC_HIP(hipEventRecord(a, s));
    }
    launch();
    ASC_HIP(hipGetLastError());
    if (prof) {
        ASC_HIP(hipEventRecord(b, s));
        t->timed.push_back({kind, a, b});
    }
    return ASC_OK;
}

// Shared end of a traversal run: kernel statistics from the recorded events
// and DevStats, the values copied to `out` (value convention applied), and
// the reached count / degree sum.
int finish_traversal(asc_graph* g, int algo, int path, uint32_t* out, asc_stats* st, const RunCounters& rc,
                     double total_ms, uint32_t iterations);

// Fast paths for graphs whose edges are all resident and whose adjacency is
// symmetric: direction-optimizing BFS and union-find CC.
bool resident_symmetric(const asc_graph* g);
int bfs_direction_opt(asc_graph* g, uint32_t src, uint32_t* out, asc_stats* st);
int cc_union_find(asc_graph* g, uint32_t* out, asc_stats* st);

}  // namespace asc
