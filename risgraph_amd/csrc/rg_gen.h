// rg_gen.h -- counter-based synthetic edge streams (host and device).
//
// draw(seed, i, k) = splitmix64(splitmix64(seed) ^ (64*i + k)).
// R-MAT edge i, scale S: for level l = 0..S-1 (most significant bit first)
//   r = draw(seed, i, l) >> 32 selects quadrant a / b / c / d by the integer
//   thresholds floor(a*2^32), floor((a+b)*2^32), floor((a+b+c)*2^32);
//   b sets the dst bit, c the src bit, d both.  Self-loops and duplicates
//   are kept.  Every edge is an independent draw, so index order is already a
//   uniformly random order and no shuffle pass is needed.
// Erdos-Renyi edge i: src = draw(seed,i,0) % V, dst = draw(seed,i,1) % V.
// Weight of stream edge i: 1 + draw(wseed, i, 0) % 1000 (u32 in [1, 1000]).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define RG_HD __host__ __device__ __forceinline__
#else
#define RG_HD inline
#endif

namespace rg {

RG_HD uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

RG_HD uint64_t draw_key(uint64_t seed) { return splitmix64(seed); }

RG_HD uint64_t draw(uint64_t key, uint64_t i, uint64_t k)
{
    return splitmix64(key ^ (i * 64ull + k));
}

struct RmatThresholds {
    uint32_t ta, tab, tabc;
};

inline RmatThresholds rmat_thresholds(double a, double b, double c)
{
    RmatThresholds t;
    t.ta = (uint32_t)(a * 4294967296.0);
    t.tab = (uint32_t)((a + b) * 4294967296.0);
    t.tabc = (uint32_t)((a + b + c) * 4294967296.0);
    return t;
}

RG_HD void rmat_edge(uint64_t key, uint64_t i, int scale, RmatThresholds t,
                     uint64_t *src, uint64_t *dst)
{
    uint64_t s = 0, d = 0;
    for (int l = 0; l < scale; l++) {
        uint32_t r = (uint32_t)(draw(key, i, (uint64_t)l) >> 32);
        uint64_t sb = r >= t.tab;
        uint64_t db = (r >= t.ta && r < t.tab) || r >= t.tabc;
        s = (s << 1) | sb;
        d = (d << 1) | db;
    }
    *src = s;
    *dst = d;
}

RG_HD void er_edge(uint64_t key, uint64_t i, uint64_t V, uint64_t *src, uint64_t *dst)
{
    *src = draw(key, i, 0) % V;
    *dst = draw(key, i, 1) % V;
}

RG_HD uint32_t stream_weight(uint64_t wkey, uint64_t i)
{
    return (uint32_t)(1 + draw(wkey, i, 0) % 1000ull);
}

}  // namespace rg
