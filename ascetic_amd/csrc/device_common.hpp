// Device primitives for gfx950 (wave64): ballots, lane prefix counts, wave
// and block scans, and LDS-staged queue appends.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace asc {

constexpr int kWave = 64;
constexpr int kBlock = 256;                 // 4 waves per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;

struct WEdge {
    uint32_t dst;
    uint32_t w;
};

__device__ __forceinline__ uint32_t edge_dst(uint32_t e) { return e; }
__device__ __forceinline__ uint32_t edge_dst(const WEdge& e) { return e.dst; }
__device__ __forceinline__ uint32_t edge_w(uint32_t) { return 1u; }
__device__ __forceinline__ uint32_t edge_w(const WEdge& e) { return e.w; }

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (kWave - 1); }
__device__ __forceinline__ uint32_t wave_id() { return threadIdx.x / kWave; }

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

// Number of set bits of `mask` below the calling lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Orders this wave's LDS writes before its later LDS reads of other lanes'
// slots (waitcnt on the LDS counter; no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Inclusive wave scan (Kogge-Stone over cross-lane shuffles).
template <class T>
__device__ __forceinline__ T wave_scan_incl(T x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        T y = __shfl_up(x, d, kWave);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}

template <class T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

// Exclusive block scan of one value per thread; `total` gets the block sum.
// s_tmp holds kWavesPerBlock entries.  Contains one __syncthreads; callers
// must separate consecutive uses of s_tmp by a barrier.
template <class T>
__device__ __forceinline__ T block_scan_excl(T x, T* s_tmp, T& total) {
    const T inc = wave_scan_incl(x);
    if (lane_id() == kWave - 1) s_tmp[wave_id()] = inc;
    __syncthreads();
    T base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; i++) {
        T t = s_tmp[i];
        if ((uint32_t)i < wave_id()) base += t;
        tot += t;
    }
    total = tot;
    return base + inc - x;
}

// Per-wave LDS staging buffer for a global queue.  Lanes that want to append
// write into LDS at their ballot rank; when the buffer would overflow it is
// flushed with a single global atomicAdd and a coalesced copy.  `n` is
// wave-uniform.  All 64 lanes must call these together.
template <int CAP>
struct WaveStage {
    uint32_t* buf;   // LDS, CAP entries
    uint32_t n;

    __device__ __forceinline__ void flush(uint32_t* gq, uint32_t* gcount) {
        if (n == 0) return;
        wave_lds_sync();
        uint32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(gcount, n);
        base = __shfl(base, 0, kWave);
        for (uint32_t j = lane_id(); j < n; j += kWave) gq[base + j] = buf[j];
        wave_lds_sync();
        n = 0;
    }

    __device__ __forceinline__ void push(bool want, uint32_t v, uint32_t* gq, uint32_t* gcount) {
        const uint64_t m = wave_ballot(want);
        if (m == 0) return;
        const uint32_t c = (uint32_t)__popcll(m);
        if (n + c > (uint32_t)CAP) flush(gq, gcount);
        if (want) buf[n + lanes_below(m)] = v;
        n += c;
    }
};

}  // namespace asc
