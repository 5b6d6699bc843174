// Host gather of on-demand adjacency into a pinned staging buffer.
//
// The reference copies each active overload vertex's edges with 20 threads,
// each taking a contiguous range of vertices (GraphMeta.cuh:469-514).  Here
// the pool splits the part's edge range evenly instead, so a hub is shared by
// all threads.
#include <algorithm>

#include "internal.hpp"

namespace asc {

void gather_edges(ThreadPool& pool, const char* edges, size_t EB, const unsigned long long* src,
                  const unsigned long long* prefix, uint64_t m, uint64_t lo, uint64_t hi, char* dst) {
    // Pieces are small (tens of edges) and scattered, so a thread is bound by
    // DRAM latency unless it keeps several pieces in flight: the first lines
    // of the piece kAhead ahead are prefetched (its position is read from the
    // sequential src array, so no lookup is on the critical path).
    constexpr uint64_t kAhead = 16;
    pool.parallel_for(hi - lo, [&](uint64_t b, uint64_t e) {
        uint64_t a = lo + b;
        const uint64_t end = lo + e;
        uint64_t k = (uint64_t)(std::upper_bound(prefix, prefix + m, (unsigned long long)a) - prefix) - 1;
        while (a < end) {
            if (k + kAhead < m) {
                const char* pf = edges + src[k + kAhead] * EB;
                __builtin_prefetch(pf);
                __builtin_prefetch(pf + 64);
                __builtin_prefetch(pf + 128);
            }
            const uint64_t kend = std::min<uint64_t>(prefix[k + 1], end);
            const uint64_t from = src[k] + (a - prefix[k]);
            std::memcpy(dst + (a - lo) * EB, edges + from * EB, (kend - a) * EB);
            a = kend;
            k++;
        }
    });
}

}  // namespace asc
