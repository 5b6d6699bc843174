// Static / on-demand planning: which prefix of the edge array stays resident
// on the device, and how the active on-demand edges are cut into parts.
//
// Reference: GraphMeta::getMaxPartitionSize (GraphMeta.cuh:310-372) and
// GraphMeta::caculatePartInfoForEdgeList (GraphMeta.cuh:516-578).  Changes:
//   * all sizes are 64-bit (the reference stores the static size, the part
//     size and the static offsets in 32 bits and caps the prefix at
//     UINT_MAX edges, GraphMeta.cuh:140-142, 349-352);
//   * an explicit byte budget can replace the free-memory rule;
//   * a prefix covering every edge is not floored to a 4096-edge fragment,
//     so a graph that fits is fully resident;
//   * the static set is the vertex prefix [0, vs) whose adjacency ends within
//     the prefix, tested as one comparison `v < vs` on the device instead of
//     a per-vertex flag array;
//   * parts are edge-granular: a vertex larger than a part spans several
//     parts (the reference's vertex-granular split never terminates then).
#include <algorithm>

#include "internal.hpp"

extern "C" int asc_plan_partition(uint64_t V, uint64_t E, const uint64_t* off, uint64_t edge_size,
                                  uint64_t free_bytes, int param_size, double advise_rate,
                                  uint64_t edge_budget_bytes, uint64_t part_bytes, asc_plan* out) {
    if (!off || !out || edge_size == 0 || V == 0) return ASC_ERR_INVALID;
    const uint64_t reduce = (uint64_t)param_size * 4u * V;
    const uint64_t total = free_bytes > reduce ? (free_bytes - reduce) / edge_size : 0;
    double k;
    if (total) {
        k = (double)(float)((10.0 - (double)E / (double)total) / 9.0);
        if (k < 0) k = 0.5;
        if (k > 1) k = 1.0;
    } else {
        k = 0.5;
    }
    if (advise_rate > 0) k = advise_rate;
    uint64_t S;
    if (edge_budget_bytes > 0) {
        S = edge_budget_bytes / edge_size;
    } else {
        S = (uint64_t)(k * (double)total);
    }
    if (S >= E) {
        S = E;
    } else {
        S -= S % 4096;  // fragment granularity (GraphMeta.cuh:143, 353-354)
    }
    // vs = number of leading vertices whose adjacency ends at or before S
    uint64_t lo = 0, hi = V;  // find largest vs with off[vs] <= S among prefix ends
    // off[v+1] <= S  <=>  v < vs ; off is non-decreasing with off[V] = E
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo + 1) / 2;
        if (off[mid] <= S) lo = mid; else hi = mid - 1;
    }
    // lo = largest index i in [0, V] with off[i] <= S; vertices [0, lo) end
    // within S except possibly vertex lo-1 whose end is off[lo] <= S: all fine.
    uint64_t vs = lo;
    if (vs > V) vs = V;
    out->total_device_edges = total;
    out->advise_k = k;
    out->static_edges = off[vs];
    out->static_vertices = vs;
    out->ondemand_edges = E - off[vs];
    const uint64_t pb = part_bytes ? part_bytes : (uint64_t(64) << 20);
    out->part_edges = std::max<uint64_t>(pb / edge_size, 1024);
    return ASC_OK;
}

extern "C" int asc_split_parts(const uint64_t* prefix, uint64_t n, uint64_t part_edges, uint64_t* part_starts,
                               uint64_t cap, uint64_t* n_parts) {
    if (!prefix || !n_parts || part_edges == 0) return ASC_ERR_INVALID;
    const uint64_t total = prefix[n];
    const uint64_t np = total ? (total + part_edges - 1) / part_edges : 0;
    *n_parts = np;
    if (!part_starts) return ASC_OK;
    if (cap < np + 1) return ASC_ERR_RANGE;
    for (uint64_t p = 0; p < np; p++) part_starts[p] = p * part_edges;
    part_starts[np] = total;
    return ASC_OK;
}
