// zg_device.h -- device-side building blocks of the estimator kernels:
// Philox4x32-10 streams keyed by (seed, estimator id), Lemire's bounded draw,
// the HBM graph view, and the sampling primitives of ZGraphInstance
// (src/ZGraphInstance.cpp) re-designed for one estimator per lane.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace zg {

// ---------------------------------------------------------------------------
// HBM layout of one graph (the reference's CSR, include/Graph.hpp:148-166, at
// one worker):
//   edges[r] = 32-byte record of rank r ("edge id"): its endpoints, and for
//              each endpoint the list start, the degree and the position of r
//              in that list -- one 128-byte line holds everything a sampled
//              edge needs, so no offset lookups follow a sample
//   off[v]   = start of v's lists (n+1 entries, u64; export and generic use)
//   adj[i]   = rank << 32 | neighbour     each list ascending by rank
//   sadj[i]  = neighbour << 32 | rank     each list ascending by (neighbour, rank)
// The stored position turns the reference's lower_bound on "rank >= e+1" in
// the endpoint lists of a sampled edge e into one load.  The linear
// conditional_close scan (src/ZGraphInstance.cpp:371-407) becomes one lookup in
// the edge-set hash below; sadj serves the exact counters and the export.
// List starts are u32, so a graph holds fewer than 2^31 edges.
// ---------------------------------------------------------------------------
// Field order: the first 16 bytes hold the endpoints' degrees and the rank's
// positions -- all that a plan needs to size its first conditional sample
// without reading the sampled entry (the 3-chain plan on a graph without
// parallel edges), so that plan issues one 16-byte load per record; the
// second 16 bytes hold the endpoints and their list starts.
struct alignas(32) ERec {
    uint32_t s_deg, s_pos;   // src's degree, position of this rank in src's list
    uint32_t d_deg, d_pos;   // the same for dst
    uint32_t src, dst;
    uint32_t s_off, d_off;   // list starts
};

// Edge-set hash for the closing checks.  A 128-byte bucket (one L2 line, the
// DRAM read granule) holds eight slots {key = min << 32 | max, rank} behind a
// 16-byte header of eight 16-bit tags: bits 0..14 are a fingerprint of the
// slot's key (0 = empty slot), and bit 15 of tag 0 marks a bucket that some
// key passed over because it was full (linear probing over buckets).  A
// lookup loads the header alone -- one 16-byte load -- and reads a slot only
// where its tag matches; it moves to the next bucket only past a marked one.
// Each load instruction of a lane pays its own address translation, which
// dominates once the table outgrows the translation caches (several GiB), so
// the common lookup is kept to one load.  Four edges per bucket on average:
// about 2 % of the buckets overflow.
constexpr int kBucketSlots = 8;
constexpr uint64_t kEdgesPerBucket = 4;
constexpr uint32_t kBucketPassed = 0x8000u;   // in tag word 0
struct HSlot {
    uint32_t key_lo, key_hi, rank;
};
struct alignas(128) HBucket {
    uint32_t tag[kBucketSlots / 2];   // two 16-bit tags per word, slot s in word s / 2
    HSlot slot[kBucketSlots];
    uint32_t pad[3];
    uint32_t fill;                    // slots claimed while building (may pass 8)
};
static_assert(sizeof(HBucket) == 128, "a hash bucket is one 128-byte line");
static_assert(offsetof(HBucket, slot) == 16, "the tag header is the bucket's first 16 bytes");

struct GView {
    const ERec *__restrict__ edges;
    const uint64_t *__restrict__ off;
    const uint64_t *__restrict__ adj;
    const uint64_t *__restrict__ sadj;
    const HBucket *__restrict__ htab;
    uint64_t m;
    uint32_t n;
    uint32_t hnb;                // number of buckets
    uint32_t simple;             // 1: no pair of vertices is joined twice (no parallel edges)
    uint32_t pad_;
    unsigned long long *reads;   // non-null only in the read-counting launch
};

// murmur3 fmix64: spreads (min, max) pairs over the buckets
__host__ __device__ __forceinline__ uint64_t hmix(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}
__host__ __device__ __forceinline__ uint32_t hbucket_of(uint64_t h, uint32_t nb)
{
    return (uint32_t)(((h >> 32) * (uint64_t)nb) >> 32);
}
// 15-bit nonzero fingerprint from the low bits of the mixed key
__host__ __device__ __forceinline__ uint32_t htag_of(uint64_t h)
{
    const uint32_t t = (uint32_t)h & 0x7fffu;
    return t ? t : 1u;
}

// counts dependent reads in the read-counting launch, one per 128-byte L2 line
// (the DRAM read granule of MI355X: a random 8..32-byte load moves one line)
__device__ __forceinline__ void count_reads(const GView &g, uint32_t k)
{
    if (g.reads) atomicAdd(g.reads, (unsigned long long)k);
}

enum : uint32_t { TAG_ESTIMATOR = 0, TAG_SHUFFLE = 1, TAG_SUBSAMPLE = 2, TAG_GEN = 3,
                  TAG_COMPLETION = 0x1000 /* + j: completion j of a split 4-clique estimator */ };

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
    uint32_t v, deg, off;
};
__device__ __forceinline__ VI vinfo(const GView &g, uint32_t v)
{
    const uint64_t a = g.off[v], b = g.off[v + 1];
    count_reads(g, ((v & 15u) == 15u) ? 2u : 1u);   // off[v], off[v+1] share a line unless v%16 == 15
    return VI{v, (uint32_t)(b - a), (uint32_t)a};
}

// the 32-byte record of rank r (two 16-byte loads of one line)
__device__ __forceinline__ ERec load_edge(const GView &g, uint32_t r)
{
    const uint4 *p = reinterpret_cast<const uint4 *>(g.edges + r);
    const uint4 a = p[0], b = p[1];
    count_reads(g, 1);
    ERec e;
    e.s_deg = a.x; e.s_pos = a.y; e.d_deg = a.z; e.d_pos = a.w;
    e.src = b.x; e.dst = b.y; e.s_off = b.z; e.d_off = b.w;
    return e;
}
// the record's first half only (one 16-byte load): {s_deg, s_pos, d_deg, d_pos}
__device__ __forceinline__ uint4 load_edge_degpos(const GView &g, uint32_t r)
{
    const uint4 a = reinterpret_cast<const uint4 *>(g.edges + r)[0];
    count_reads(g, 1);
    return a;
}
__device__ __forceinline__ VI src_vi(const ERec &e) { return VI{e.src, e.s_deg, e.s_off}; }
__device__ __forceinline__ VI dst_vi(const ERec &e) { return VI{e.dst, e.d_deg, e.d_off}; }
// endpoint v of e: its lists, and the position of e's rank in them
__device__ __forceinline__ VI end_vi(const ERec &e, uint32_t v, uint32_t &pos)
{
    if (e.src == v) { pos = e.s_pos; return src_vi(e); }
    pos = e.d_pos;
    return dst_vi(e);
}

// first position p in [lo, hi] of v's rank-ordered list with rank >= min_valid
// (the lower_bound of core_conditional_sample_edge, src/ZGraphInstance.cpp:147-180).
// Ranks are a uniform random permutation, so an interpolation step on the known
// key bounds [kl, ku) of the range lands within a few entries; a step that
// fails to halve the range is followed by a bisection step (O(log n) worst case).
__device__ __forceinline__ uint32_t lb_rank(const GView &g, const VI &x, uint32_t lo, uint32_t hi,
                                            uint32_t min_valid, uint64_t kl, uint64_t ku)
{
    const uint64_t *__restrict__ a = g.adj + (uint64_t)x.off;
    uint32_t probes = 0;
    uint64_t line = ~0ull;   // last line probed (read counting only)
    bool bisect = false;
    while (hi - lo > 16) {
        const uint32_t n = hi - lo;
        uint32_t p;
        if (bisect || ku <= kl) {
            p = lo + (n >> 1);
        } else {
            const uint64_t t = min_valid > kl ? (uint64_t)min_valid - kl : 0;
            p = lo + (uint32_t)(((uint64_t)n * t) / (ku - kl));
            if (p >= hi) p = hi - 1;
        }
        const uint32_t r = (uint32_t)(a[p] >> 32);
        if (g.reads && ((x.off + p) >> 4) != line) { line = (x.off + p) >> 4; ++probes; }
        if (r < min_valid) { lo = p + 1; kl = (uint64_t)r + 1; }
        else { hi = p; ku = r; }
        bisect = (hi - lo) > (n >> 1);
    }
    const uint64_t key = (uint64_t)min_valid << 32;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
        if (g.reads && ((x.off + mid) >> 4) != line) { line = (x.off + mid) >> 4; ++probes; }
    }
    count_reads(g, probes);
    return lo;
}

// conditional_close for one awaited edge {x, y} (src/ZGraphInstance.cpp:371-441):
// true iff some copy of {x, y} has rank >= min_valid; *rank receives the
// smallest such rank.  One 128-byte hash bucket (one line) in the common case:
// a 16-byte header load, plus one 12-byte slot load per matching tag.
__device__ __forceinline__ bool close_pair(const GView &g, uint32_t x, uint32_t y, uint32_t min_valid,
                                           uint32_t *rank)
{
    if (x == y) return false;
    const uint64_t key = x < y ? ((uint64_t)x << 32) | y : ((uint64_t)y << 32) | x;
    const uint64_t h = hmix(key);
    const uint32_t t = htag_of(h);
    uint32_t b = hbucket_of(h, g.hnb);
    uint32_t best = 0xffffffffu;
    uint32_t buckets = 0;
    for (;;) {
        const uint4 hd = *reinterpret_cast<const uint4 *>(g.htab[b].tag);
        ++buckets;
        const uint32_t w[4] = {hd.x, hd.y, hd.z, hd.w};
        uint32_t match = 0;
#pragma unroll
        for (int s = 0; s < kBucketSlots; ++s)
            match |= (uint32_t)(((w[s >> 1] >> (16 * (s & 1))) & 0x7fffu) == t) << s;
        while (match) {   // a slot in the same line, read only on a tag match
            const int s = __builtin_ctz(match);
            match &= match - 1;
            const HSlot sl = g.htab[b].slot[s];
            if (sl.key_lo == (uint32_t)key && sl.key_hi == (uint32_t)(key >> 32) && sl.rank >= min_valid &&
                sl.rank < best)
                best = sl.rank;
        }
        if (!(hd.x & kBucketPassed)) break;
        b = (b + 1 == g.hnb) ? 0 : b + 1;
    }
    count_reads(g, buckets);
    if (best == 0xffffffffu) return false;
    if (rank) *rank = best;
    return true;
}

// multiplicity of the pair {x, y}, x != y: the stored copies of the edge (0 if
// none), from the same bucket chain as close_pair -- one line in the common case
__device__ __forceinline__ uint32_t close_count(const GView &g, uint32_t x, uint32_t y)
{
    const uint64_t key = x < y ? ((uint64_t)x << 32) | y : ((uint64_t)y << 32) | x;
    const uint64_t h = hmix(key);
    const uint32_t t = htag_of(h);
    uint32_t b = hbucket_of(h, g.hnb);
    uint32_t n = 0, buckets = 0;
    for (;;) {
        const uint4 hd = *reinterpret_cast<const uint4 *>(g.htab[b].tag);
        ++buckets;
        const uint32_t w[4] = {hd.x, hd.y, hd.z, hd.w};
        uint32_t match = 0;
#pragma unroll
        for (int s = 0; s < kBucketSlots; ++s)
            match |= (uint32_t)(((w[s >> 1] >> (16 * (s & 1))) & 0x7fffu) == t) << s;
        while (match) {
            const int s = __builtin_ctz(match);
            match &= match - 1;
            const HSlot sl = g.htab[b].slot[s];
            n += sl.key_lo == (uint32_t)key && sl.key_hi == (uint32_t)(key >> 32);
        }
        if (!(hd.x & kBucketPassed)) break;
        b = (b + 1 == g.hnb) ? 0 : b + 1;
    }
    count_reads(g, buckets);
    return n;
}

__device__ __forceinline__ bool close_edge(const GView &g, const VI &x, const VI &y, uint32_t min_valid,
                                           uint32_t *rank)
{
    return close_pair(g, x.v, y.v, min_valid, rank);
}

// sampled edge: rank, owner (src), neighbour (dst), and its position in the
// owner's list when it came from a conditional sample
struct SE {
    uint32_t rank, src, dst, pos;
};

}  // namespace zg
