// The synthetic-graph sampler, shared by the host generator (generate.cpp)
// and the device generator (gen_device.hip) so that both draw the same
// samples bit for bit.
//
// Every sample is a pure function of (seed, sample index): a SplitMix64
// finaliser over a counter.  R-MAT has no noise and no vertex permutation
// (vertex 0 is the largest hub); Erdos-Renyi draws both endpoints uniformly.
// Weights are drawn per unordered pair, so w(u,v) = w(v,u), and are a pure
// function of (seed, pair).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/ascetic.h"

namespace asc {

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t rnd(uint64_t key, uint64_t ctr) { return mix64(key + ctr * 0x9E3779B97F4A7C15ull); }

// High 64 bits of the 128-bit product, from 32-bit halves (exact).
__host__ __device__ inline uint64_t mulhi64(uint64_t a, uint64_t b) {
    const uint64_t a0 = a & 0xFFFFFFFFull, a1 = a >> 32, b0 = b & 0xFFFFFFFFull, b1 = b >> 32;
    const uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
    const uint64_t mid = (p00 >> 32) + (p01 & 0xFFFFFFFFull) + (p10 & 0xFFFFFFFFull);
    return p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

struct Sampler {
    int type;                    // 0 = R-MAT, 1 = Erdos-Renyi
    int scale;
    uint64_t V;
    uint64_t key;
    uint32_t ta, tab, tabc;      // cumulative quadrant thresholds on 32-bit draws

    // One undirected sample.
    __host__ __device__ inline void sample(uint64_t i, uint32_t& u, uint32_t& v) const {
        if (type == 1) {
            u = (uint32_t)mulhi64(rnd(key, 2 * i), V);
            v = (uint32_t)mulhi64(rnd(key, 2 * i + 1), V);
            return;
        }
        uint32_t uu = 0, vv = 0;
        uint64_t word = 0;
        for (int l = 0; l < scale; l++) {
            if ((l & 1) == 0) word = rnd(key, (i << 5) + (uint64_t)(l >> 1));
            const uint32_t r = (uint32_t)(word >> ((l & 1) * 32));
            const uint32_t row = r >= tab, col = (r >= ta && r < tab) || r >= tabc;
            uu = (uu << 1) | row;
            vv = (vv << 1) | col;
        }
        u = uu;
        v = vv;
    }
};

__host__ __device__ inline uint32_t pair_weight(uint64_t wkey, uint32_t a, uint32_t b, uint32_t wmin, uint32_t span) {
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    const uint64_t h = mix64(wkey ^ (((uint64_t)lo << 32) | hi));
    return wmin + (uint32_t)mulhi64(h, span);
}

// Validates the parameters and sets up the sampler; *M = samples to draw.
int make_sampler(const asc_gen_params* p, Sampler& s, uint64_t& M);
// Key of the pair weights.
inline uint64_t weight_key(uint64_t seed) { return mix64(seed ^ 0x3E16A7ull); }

}  // namespace asc
