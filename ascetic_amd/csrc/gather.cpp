// Host gather of on-demand adjacency into a pinned staging buffer.
//
// The reference copies each active overload vertex's edges with 20 threads,
// each taking a contiguous range of vertices (GraphMeta.cuh:469-514).  Here
// the pool splits the part's edge range evenly instead, so a hub is shared by
// all threads.
#include <algorithm>

#include "internal.hpp"

namespace asc {

void gather_edges(ThreadPool& pool, const char* edges, uint64_t base, const uint64_t* off, size_t EB,
                  const uint32_t* list, const unsigned long long* prefix, uint64_t m, uint64_t lo, uint64_t hi,
                  char* dst) {
    // Pieces are small (tens of edges) and scattered, so a thread is bound by
    // DRAM latency unless it keeps several pieces in flight: prefetch the
    // offset of the piece kOff ahead and the first lines of the piece kEdge
    // ahead (whose offset arrived by then).
    constexpr uint64_t kOff = 32, kEdge = 8;
    pool.parallel_for(hi - lo, [&](uint64_t b, uint64_t e) {
        uint64_t a = lo + b;
        const uint64_t end = lo + e;
        uint64_t k = (uint64_t)(std::upper_bound(prefix, prefix + m, (unsigned long long)a) - prefix) - 1;
        while (a < end) {
            if (k + kOff < m) __builtin_prefetch(off + list[k + kOff]);
            if (k + kEdge < m) {
                const char* pf = edges + (off[list[k + kEdge]] - base) * EB;
                __builtin_prefetch(pf);
                __builtin_prefetch(pf + 64);
                __builtin_prefetch(pf + 128);
            }
            const uint64_t kend = std::min<uint64_t>(prefix[k + 1], end);
            const uint64_t from = off[list[k]] + (a - prefix[k]) - base;
            std::memcpy(dst + (a - lo) * EB, edges + from * EB, (kend - a) * EB);
            a = kend;
            k++;
        }
    });
}

}  // namespace asc
