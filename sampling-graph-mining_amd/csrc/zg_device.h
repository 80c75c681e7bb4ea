// zg_device.h -- device-side building blocks of the estimator kernels:
// Philox4x32-10 streams keyed by (seed, estimator id), Lemire's bounded draw,
// the HBM graph view, and the sampling primitives of ZGraphInstance
// (src/ZGraphInstance.cpp) re-designed for one estimator per lane.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zg {

// ---------------------------------------------------------------------------
// HBM layout of one graph (all arrays indexed as in the reference's CSR,
// include/Graph.hpp:148-166, at one worker):
//   edges[r] = {src, dst, position of r in src's list, position in dst's list}
//              16 B per edge, r = rank ("edge id")
//   off[v]   = start of v's lists (n+1 entries, u64)
//   adj[i]   = rank << 32 | neighbour     each list ascending by rank
//   sadj[i]  = neighbour << 32 | rank     each list ascending by (neighbour, rank)
// The position index turns the reference's lower_bound on "rank >= e+1" in
// the endpoint lists of a sampled edge e into one load; sadj turns the linear
// conditional_close scan (src/ZGraphInstance.cpp:371-407) into one binary search.
// ---------------------------------------------------------------------------
struct GView {
    const uint4 *__restrict__ edges;
    const uint64_t *__restrict__ off;
    const uint64_t *__restrict__ adj;
    const uint64_t *__restrict__ sadj;
    uint64_t m;
    uint32_t n;
    unsigned long long *reads;   // non-null only in the read-counting launch
};

// counts dependent reads (one 64-byte sector each) in the read-counting launch
__device__ __forceinline__ void count_reads(const GView &g, uint32_t k)
{
    if (g.reads) atomicAdd(g.reads, (unsigned long long)k);
}

enum : uint32_t { TAG_ESTIMATOR = 0, TAG_SHUFFLE = 1, TAG_SUBSAMPLE = 2, TAG_GEN = 3 };

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__host__ __device__ __forceinline__ uint4 philox10(uint4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1,
                       (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0);
    }
    return c;
}

__host__ __device__ __forceinline__ uint64_t key64(uint64_t seed, uint32_t tag, uint64_t id)
{
    uint4 o = philox10(make_uint4(0u, tag, (uint32_t)id, (uint32_t)(id >> 32)), (uint32_t)seed,
                       (uint32_t)(seed >> 32));
    return (uint64_t)o.x | ((uint64_t)o.y << 32);
}

// One estimator's stream: word w comes from counter block (w/4, tag, id).
struct Rng {
    uint32_t k0, k1, tag, idlo, idhi, word, blk;
    uint4 buf;

    __device__ __forceinline__ void init(uint64_t seed, uint32_t t, uint64_t id)
    {
        k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
        tag = t; idlo = (uint32_t)id; idhi = (uint32_t)(id >> 32);
        word = 0; blk = 0xffffffffu;
    }
    __device__ __forceinline__ uint32_t next()
    {
        const uint32_t b = word >> 2;
        if (b != blk) { buf = philox10(make_uint4(b, tag, idlo, idhi), k0, k1); blk = b; }
        const uint32_t i = word & 3u;
        ++word;
        return i == 0 ? buf.x : (i == 1 ? buf.y : (i == 2 ? buf.z : buf.w));
    }
    // U[0, n-1]: Lemire, "Fast random integer generation in an interval",
    // ACM TOMACS 29(1), 2019; 32-bit words for n <= 2^32, 64-bit otherwise.
    __device__ __forceinline__ uint64_t uniform(uint64_t n)
    {
        if (n <= (1ull << 32)) {
            uint64_t mlt = (uint64_t)next() * n;
            uint32_t l = (uint32_t)mlt;
            if ((uint64_t)l < n) {
                const uint64_t t = ((1ull << 32) - n) % n;
                while ((uint64_t)l < t) { mlt = (uint64_t)next() * n; l = (uint32_t)mlt; }
            }
            return mlt >> 32;
        }
        uint64_t x = next();
        x |= (uint64_t)next() << 32;
        uint64_t lo = x * n, hi = __umul64hi(x, n);
        if (lo < n) {
            const uint64_t t = (0 - n) % n;
            while (lo < t) {
                x = next();
                x |= (uint64_t)next() << 32;
                lo = x * n; hi = __umul64hi(x, n);
            }
        }
        return hi;
    }
};

// 128-bit unsigned value with overflow tracking
struct U128 {
    uint64_t lo, hi;
};
__device__ __forceinline__ bool mul_u128(U128 &v, uint64_t f)
{
    const uint64_t l = v.lo * f, lh = __umul64hi(v.lo, f);
    const uint64_t h = v.hi * f, hh = __umul64hi(v.hi, f);
    const uint64_t nh = lh + h;
    v.lo = l; v.hi = nh;
    return hh != 0 || nh < lh;   // true on overflow
}
__device__ __forceinline__ bool add_u128(U128 &a, const U128 &b)
{
    const uint64_t lo = a.lo + b.lo;
    const uint64_t c = lo < a.lo;
    const uint64_t hi = a.hi + b.hi + c;
    const bool ov = hi < a.hi || (hi == a.hi && c);
    a.lo = lo; a.hi = hi;
    return ov;
}

// vertex descriptor: list start and degree
struct VI {
    uint32_t v, deg;
    uint64_t off;
};
__device__ __forceinline__ VI vinfo(const GView &g, uint32_t v)
{
    const uint64_t a = g.off[v], b = g.off[v + 1];
    count_reads(g, ((v & 7u) == 7u) ? 2u : 1u);   // off[v], off[v+1] share a sector unless v%8 == 7
    return VI{v, (uint32_t)(b - a), a};
}

// first position p in [lo, hi] of v's rank-ordered list with rank >= min_valid
// (the lower_bound of core_conditional_sample_edge, src/ZGraphInstance.cpp:147-180)
__device__ __forceinline__ uint32_t lb_rank(const GView &g, const VI &x, uint32_t lo, uint32_t hi,
                                            uint32_t min_valid)
{
    const uint64_t key = (uint64_t)min_valid << 32;
    const uint64_t *__restrict__ a = g.adj + x.off;
    uint32_t probes = 0;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
        ++probes;
    }
    count_reads(g, probes);
    return lo;
}

// conditional_close for one awaited edge {x, y} (src/ZGraphInstance.cpp:371-441):
// true iff some copy of {x, y} has rank >= min_valid; *rank receives the
// smallest such rank.  Binary search in the shorter neighbour-sorted list.
__device__ __forceinline__ bool close_edge(const GView &g, const VI &x, const VI &y, uint32_t min_valid,
                                           uint32_t *rank)
{
    if (x.v == y.v) return false;
    const bool xs = x.deg <= y.deg;
    const uint64_t base = xs ? x.off : y.off;
    uint32_t hi = xs ? x.deg : y.deg;
    const uint32_t other = xs ? y.v : x.v;
    const uint64_t key = ((uint64_t)other << 32) | min_valid;
    const uint64_t *__restrict__ a = g.sadj + base;
    uint32_t lo = 0, probes = 0;
    const uint32_t n = hi;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
        ++probes;
    }
    count_reads(g, probes + (lo < n));
    if (lo >= n) return false;
    const uint64_t e = a[lo];
    if ((uint32_t)(e >> 32) != other) return false;
    if (rank) *rank = (uint32_t)e;
    return true;
}

// sampled edge: rank, owner (src), neighbour (dst), and its position in the
// owner's list when it came from a conditional sample
struct SE {
    uint32_t rank, src, dst, pos;
};

}  // namespace zg
