// Internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ascetic.h"

namespace asc {

// ---- errors ----------------------------------------------------------------
// Thread-local text of the last error, printed by the CLI.
void set_last_error(const char* fmt, ...);
const char* last_error();

#define ASC_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::asc::set_last_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                  __FILE__, __LINE__);                                  \
            return ASC_ERR_DEVICE;                                                      \
        }                                                                               \
    } while (0)

#define ASC_TRY(expr)              \
    do {                           \
        int r_ = (expr);           \
        if (r_ != ASC_OK) return r_; \
    } while (0)

// ---- timing -----------------------------------------------------------------
inline double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// ---- host buffers -----------------------------------------------------------
// Uninitialised host array; pinned (hipHostMalloc) or pageable (malloc).
template <class T>
struct HostBuf {
    T* p = nullptr;
    size_t n = 0;
    bool pinned = false;
    bool borrowed = false;   // the caller's memory: never freed here
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    HostBuf(HostBuf&& o) noexcept { *this = std::move(o); }
    HostBuf& operator=(HostBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n; pinned = o.pinned; borrowed = o.borrowed;
            o.p = nullptr; o.n = 0; o.borrowed = false;
        }
        return *this;
    }
    ~HostBuf() { release(); }
    int alloc(size_t count, bool pin = false) {
        release();
        if (count == 0) return ASC_OK;
        if (pin) {
            if (hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault) != hipSuccess) {
                (void)hipGetLastError();   // reported here; later error checks must not see it
                p = nullptr;
                set_last_error("hipHostMalloc of %zu bytes failed", count * sizeof(T));
                return ASC_ERR_NOMEM;
            }
        } else {
            p = (T*)std::malloc(count * sizeof(T));
            if (!p) {
                set_last_error("malloc of %zu bytes failed", count * sizeof(T));
                return ASC_ERR_NOMEM;
            }
        }
        n = count;
        pinned = pin;
        return ASC_OK;
    }
    // Page-locked when the system grants it (page-locked memory is capped
    // well below the host's RAM: about 70 GB of 300 on the benchmark hosts),
    // else pageable.
    // max_bytes > 0: page-locked only up to that size.
    int alloc_prefer_pinned(size_t count, uint64_t max_bytes) {
        if ((!max_bytes || count * sizeof(T) <= max_bytes) && alloc(count, true) == ASC_OK) return ASC_OK;
        return alloc(count, false);
    }
    // Refers to `count` elements the caller owns and keeps alive.
    void borrow(T* q, size_t count) {
        release();
        p = q;
        n = count;
        pinned = false;
        borrowed = true;
    }
    void release() {
        if (p && !borrowed) {
            if (pinned) (void)hipHostFree(p); else std::free(p);
        }
        p = nullptr;
        n = 0;
        borrowed = false;
    }
    T& operator[](size_t i) { return p[i]; }
    const T& operator[](size_t i) const { return p[i]; }
};

// A temporary device allocation of set-up code, freed when it leaves scope
// (early error returns included).
template <class T>
struct DevTmp {
    T* p = nullptr;
    DevTmp() = default;
    DevTmp(const DevTmp&) = delete;
    DevTmp& operator=(const DevTmp&) = delete;
    ~DevTmp() {
        if (p) (void)hipFree(p);
    }
};

// A timing event destroyed when it leaves scope.
struct ScopedEvent {
    hipEvent_t e = nullptr;
    ScopedEvent() = default;
    ScopedEvent(const ScopedEvent&) = delete;
    ScopedEvent& operator=(const ScopedEvent&) = delete;
    ~ScopedEvent() {
        if (e) (void)hipEventDestroy(e);
    }
};

// Host CSR: offsets with sentinel (V+1), edge words (E, or 2E for {dst,w}).
struct HostCSR {
    uint64_t V = 0, E = 0;
    int kind = ASC_KIND_CSR;
    int offset_bytes = 8;   // width of the offsets in the source (file / view)
    HostBuf<uint64_t> off;
    HostBuf<uint32_t> edges;
    HostBuf<uint32_t> outdeg;
    int edge_words() const { return kind == ASC_KIND_WCSR ? 2 : 1; }
    size_t edge_bytes() const { return 4u * (size_t)edge_words(); }
};

int read_graph_file(const char* path, int format, HostCSR& out);
int write_graph_file(const char* path, int format, const asc_csr_view* v);
int csr_from_view(const asc_csr_view* v, HostCSR& out, bool borrow = false);
int validate_csr(const HostCSR& h);
int generate_graph(const asc_gen_params* p, HostCSR& out);
int read_edge_list(const char* path, const asc_edge_list_opts* o, HostCSR& out);

// ---- thread pool ------------------------------------------------------------
class ThreadPool {
public:
    explicit ThreadPool(int n);
    ~ThreadPool();
    int size() const { return (int)workers_.size() + 1; }
    // Runs fn(t, nthreads) on every thread (the caller is thread 0) and waits.
    void run(const std::function<void(int, int)>& fn);
    // Splits [0, n) into contiguous ranges, one per thread.
    void parallel_for(uint64_t n, const std::function<void(uint64_t, uint64_t)>& fn);

private:
    void loop(int id);
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int, int)>* job_ = nullptr;
    uint64_t gen_ = 0;
    int pending_ = 0;
    bool stop_ = false;
};
int hardware_threads();
// Process-wide pool for host-side preprocessing (generation, loading).
ThreadPool& host_pool();

// Copies edges [lo, hi) of the concatenation of m pieces into dst: piece k
// owns [prefix[k], prefix[k+1]) of the concatenation and starts at record
// src[k] of `edges` (records of EB bytes).
void gather_edges(ThreadPool& pool, const char* edges, size_t EB, const unsigned long long* src,
                  const unsigned long long* prefix, uint64_t m, uint64_t lo, uint64_t hi, char* dst);

}  // namespace asc

// ---- opaque handles ---------------------------------------------------------
struct asc_ctx {
    asc_config cfg{};
    int device = 0;
    int num_cus = 0;
    uint64_t total_mem = 0;
    hipStream_t s_main = nullptr;   // kernels
    hipStream_t s_copy = nullptr;   // host<->device transfers of on-demand parts
    asc::ThreadPool* pool = nullptr;  // gather threads
};

namespace asc {
struct TravState;
struct PrState;
}

struct asc_graph {
    asc_ctx* ctx = nullptr;
    asc::HostCSR h;
    int off_bytes = 8;              // device offset width
    // sizing inputs of this graph's plan (from the context's config at
    // creation; asc_graph_replan changes them for this graph only)
    uint64_t edge_budget_bytes = 0;
    double advise_rate = 0;
    // false after a re-plan whose upload failed part-way: the device copies
    // are gone and every algorithm call returns ASC_ERR_STATE until a
    // successful asc_graph_replan
    bool usable = true;
    // servers created over this graph and not yet destroyed; re-planning or
    // destroying the graph is refused while any exists
    std::atomic<int> servers{0};
    // static / on-demand plan
    uint64_t static_edges = 0;      // edges [0, static_edges) resident on device
    uint64_t static_vertices = 0;   // vertices [0, static_vertices) are static
    uint64_t part_edges = 0;        // capacity of each on-demand part buffer
    double advise_k = 1.0;
    // device copies
    void* d_off = nullptr;          // OffT[V+1]
    void* d_edges = nullptr;        // EdgeT[static_edges]
    uint32_t* d_outdeg = nullptr;   // PR only
    uint64_t device_bytes = 0;
    // pinned mirror of the on-demand edges for zero-copy streaming
    asc::HostBuf<uint32_t> zc_tail;
    // On-demand edges packed one per u32 word, dst | (w - wmin) << pk_shift,
    // when the id and weight bits fit (weighted graphs only); pinned.  The
    // streamed parts then carry 4 instead of 8 bytes per edge.
    asc::HostBuf<uint32_t> ptail;
    // Split form, when the id and weight bits do not fit one word but the
    // weights span at most 256 values: ptail holds the ids and ptail_w the
    // bytes w - wmin (5 bytes per streamed edge).  A device part buffer or
    // staging buffer of PE edges then holds PE ids followed by PE weight bytes.
    asc::HostBuf<uint8_t> ptail_w;
    bool packed = false;
    bool pk_split = false;
    uint32_t pk_shift = 32, pk_wmin = 0;
    // bytes per streamed on-demand edge
    size_t stream_record_bytes() const { return packed ? (pk_split ? 5 : 4) : h.edge_bytes(); }
    // the on-demand records sit in page-locked memory (DMA and zero-copy
    // reads straight from them); pageable packed tails go through staging
    bool tail_pinned() const {
        return packed ? ptail.pinned && (!pk_split || ptail_w.pinned) : zc_tail.p != nullptr;
    }
    // tunables
    int heavy_threshold = 0;
    int pr_group = 0;
    int sssp_delta = 0;
    double mean_weight = 0;         // SSSP: mean edge weight (0 = not computed)
    int serve_lanes = 64;
    bool host_edges_released = false;   // h.edges freed after packing (asc_config.release_host_edges)
    uint64_t fp = 0;                // graph fingerprint (cached before a release)
    int pr_tile_src_bits = 0;       // PageRank tile shape (0 = default)
    int pr_tile_dst_bits = 0;
    int blocks_per_cu = 0;          // workgroups per CU of traversal launches (0 = 8)
    uint32_t max_degree = 0, max_degree_vertex = 0;
    // adjacency symmetric (every u->v has a v->u of equal multiplicity):
    // 1 / 0, or -1 when not checked (checked on upload when fully resident)
    int symmetric = -1;
    asc::TravState* trav = nullptr;
    asc::PrState* pr = nullptr;
};

namespace asc {
int graph_upload(asc_graph* g);       // device copies after a (re)plan
int graph_ready(const asc_graph* g);  // ASC_OK, or ASC_ERR_STATE for an unusable graph
int noalloc_init(int device);          // device address of the no-allocation scratch (device current)
uint64_t fingerprint(const asc_graph* g);   // sampled hash of the graph (tuning tables)
void compute_mean_weight(asc_graph* g);    // mean edge weight, once (SSSP band width)
void graph_free_device(asc_graph* g);
void trav_free(TravState* t);
void pr_free(PrState* p);
int effective_heavy_threshold(const asc_graph* g);
int check_symmetric(asc_graph* g);   // sets g->symmetric from the device copy
int param_size_for(int kind);         // the reference's paramSize per graph kind (sizing rule)
void compute_max_degree(asc_graph* g);
}  // namespace asc
