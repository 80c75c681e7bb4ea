// zg_exact.hip -- exact subgraph counts on the device, the scalable
// counterpart of naive_implementation/*.cpp (non-induced counts on the
// simple graph underlying the edge list):
//   triangles  naive_implementation/TriangleCounting.cpp:44-72
//   4-cliques  naive_implementation/FourClique.cpp:44-78
//   wedges     naive_implementation/ThreeChain.cpp:44-51 (sum of C(d, 2))
//   3-paths    naive_implementation/FourChain.cpp:46-66 (sum over edges (du-1)(dv-1) - 3T)
//   5-stars    naive_implementation/FiveStar.cpp:51-59 (sum of C(d, 4))
//   5-houses   naive_implementation/FiveHouse.cpp (per ordered edge, k_house)
// Method: orient every undirected pair from lower to higher (degree, id) and
// keep the out-lists sorted by neighbour id (taken from the neighbour-ordered
// copy, parallel edges collapsed).  Triangles: one wavefront per oriented edge
// u->v binary-searches the entries of one out-list in the other, whichever
// takes fewer passes times search depth.  4-cliques:
// per vertex u, bit rows over out(u) (row i: which of out(u) are in
// out(out(u)[i])) and the sum of popcount(row_i & row_j) over the set bits j
// of each row.
#include <cstdlib>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "zg_internal.h"

namespace zg {
namespace {

template <class T>
struct XB {
    T *p = nullptr;
    hipError_t alloc(size_t k) { return hipMalloc(&p, sizeof(T) * (k ? k : 1)); }
    ~XB() { if (p) (void)hipFree(p); }
};
#define XALLOC(b, k)                                                                   \
    do {                                                                               \
        if ((b).alloc(k) != hipSuccess) return set_error(ZG_ENOMEM, "device allocation failed"); \
    } while (0)
#define GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

inline int gridx(uint64_t n)
{
    uint64_t g = (n + 255) / 256, cap = (uint64_t)device_cus() * 64;
    return (int)std::max<uint64_t>(1, std::min(g, cap));
}

// One wavefront per vertex: the lanes read 64 consecutive entries of the
// vertex's neighbour-ordered list at a time (coalesced), so a hub's list is
// shared by 64 lanes instead of walked by one thread.
#define WAVE_STRIDE(v, n)                                                                    \
    for (uint64_t v = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; v < (n);       \
         v += ((uint64_t)gridDim.x * blockDim.x) >> 6)

// neighbour of entry i of a list [b, e), or a value no vertex id takes
__device__ __forceinline__ uint32_t nbr_or_none(const uint64_t *sadj, uint64_t i, uint64_t e)
{
    return i < e ? (uint32_t)(sadj[i] >> 32) : 0xffffffffu;
}

// first entry of a run of equal neighbours (parallel edges collapse to one)
__device__ __forceinline__ bool first_of_run(uint32_t nb, uint32_t &carry)
{
    const uint32_t lane = threadIdx.x & 63;
    uint32_t prev = __shfl_up(nb, 1, 64);
    if (lane == 0) prev = carry;
    carry = __shfl(nb, 63, 64);
    return nb != 0xffffffffu && nb != prev;
}

// simple-graph degree of v: distinct neighbours in its neighbour-ordered list
__global__ void k_simple_deg(const uint64_t *off, const uint64_t *sadj, uint32_t n, uint32_t *sdeg)
{
    WAVE_STRIDE(v, (uint64_t)n)
    {
        const uint64_t b = off[v], e = off[v + 1];
        uint32_t d = 0, carry = 0xffffffffu;
        for (uint64_t base = b; base < e; base += 64) {
            const uint32_t nb = nbr_or_none(sadj, base + (threadIdx.x & 63), e);
            d += (uint32_t)__popcll(__ballot(first_of_run(nb, carry)));
        }
        if ((threadIdx.x & 63) == 0) sdeg[v] = d;
    }
}

__device__ __forceinline__ bool before(const uint32_t *sdeg, uint32_t a, uint32_t b)
{
    const uint32_t da = sdeg[a], db = sdeg[b];
    return da < db || (da == db && a < b);
}

// out-degree (toward higher order) of every vertex
__global__ void k_out_deg(const uint64_t *off, const uint64_t *sadj, const uint32_t *sdeg, uint32_t n,
                          uint64_t *odeg)
{
    WAVE_STRIDE(v, (uint64_t)n)
    {
        const uint64_t b = off[v], e = off[v + 1];
        uint64_t d = 0;
        uint32_t carry = 0xffffffffu;
        for (uint64_t base = b; base < e; base += 64) {
            const uint32_t nb = nbr_or_none(sadj, base + (threadIdx.x & 63), e);
            const bool q = first_of_run(nb, carry) && before(sdeg, (uint32_t)v, nb);
            d += (uint64_t)__popcll(__ballot(q));
        }
        if ((threadIdx.x & 63) == 0) odeg[v] = d;
    }
}

// out-lists in neighbour order (ballot compaction keeps the order)
__global__ void k_out_fill(const uint64_t *off, const uint64_t *sadj, const uint32_t *sdeg, uint32_t n,
                           const uint64_t *ooff, uint32_t *onb, uint32_t *osrc)
{
    const uint32_t lane = threadIdx.x & 63;
    WAVE_STRIDE(v, (uint64_t)n)
    {
        const uint64_t b = off[v], e = off[v + 1];
        uint64_t w = ooff[v];
        uint32_t carry = 0xffffffffu;
        for (uint64_t base = b; base < e; base += 64) {
            const uint32_t nb = nbr_or_none(sadj, base + lane, e);
            const bool q = first_of_run(nb, carry) && before(sdeg, (uint32_t)v, nb);
            const uint64_t mask = __ballot(q);
            if (q) {
                const uint64_t at = w + (uint64_t)__popcll(mask & ((1ull << lane) - 1));
                onb[at] = nb;
                osrc[at] = (uint32_t)v;
            }
            w += (uint64_t)__popcll(mask);
        }
    }
}

__device__ __forceinline__ bool bsearch_u32(const uint32_t *a, uint64_t n, uint32_t x)
{
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo < n && a[lo] == x;
}

// ceil(log2(x)): the depth of a binary search over x - 1 entries
__device__ __forceinline__ uint64_t clog2_u64(uint64_t x) { return x <= 1 ? 0 : 64 - __clzll(x - 1); }

__device__ __forceinline__ unsigned long long block_sum(unsigned long long v)
{
    __shared__ unsigned long long sm[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sm[w];
    return t;
}

// Entries of a[0, na) found in b[0, nb), counted per lane: each lane takes two
// entries of a per pass and runs their binary searches interleaved, so two
// dependent chains of loads are in flight.
__device__ __forceinline__ unsigned long long count_found(const uint32_t *a, uint64_t na, const uint32_t *b,
                                                          uint64_t nb, uint32_t lane)
{
    unsigned long long c = 0;
    for (uint64_t base = 0; base < na; base += 128) {
        const uint64_t i0 = base + lane, i1 = i0 + 64;
        const uint32_t x0 = i0 < na ? a[i0] : 0, x1 = i1 < na ? a[i1] : 0;
        uint64_t lo0 = 0, hi0 = i0 < na ? nb : 0, lo1 = 0, hi1 = i1 < na ? nb : 0;
        while (lo0 < hi0 || lo1 < hi1) {
            const uint64_t m0 = (lo0 + hi0) >> 1, m1 = (lo1 + hi1) >> 1;
            const uint32_t y0 = lo0 < hi0 ? b[m0] : 0, y1 = lo1 < hi1 ? b[m1] : 0;
            if (lo0 < hi0) { if (y0 < x0) lo0 = m0 + 1; else hi0 = m0; }
            if (lo1 < hi1) { if (y1 < x1) lo1 = m1 + 1; else hi1 = m1; }
        }
        c += (i0 < na && lo0 < nb && b[lo0] == x0) ? 1 : 0;
        c += (i1 < na && lo1 < nb && b[lo1] == x1) ? 1 : 0;
    }
    return c;
}

// One wavefront per oriented edge u -> v counts |out(u) & out(v)|, the lanes
// taking the entries of whichever list costs fewer passes of 128 entries
// times search depth in the other: a short list is searched in a long one,
// but 64 entries of a long list searched in a list of 3 cost one pass of
// depth 2 instead of a pass of depth 7.
__global__ __launch_bounds__(256) void k_tri_wave(const uint64_t *ooff, const uint32_t *onb, const uint32_t *osrc,
                                                  uint64_t ne, unsigned long long *partial)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    unsigned long long c = 0;
    for (uint64_t e = wave; e < ne; e += nwaves) {
        const uint32_t u = osrc[e], v = onb[e];
        const uint64_t ub = ooff[u], du = ooff[u + 1] - ub, vb = ooff[v], dv = ooff[v + 1] - vb;
        const bool from_u = ((du + 127) >> 7) * clog2_u64(dv + 1) <= ((dv + 127) >> 7) * clog2_u64(du + 1);
        c += from_u ? count_found(onb + ub, du, onb + vb, dv, lane) : count_found(onb + vb, dv, onb + ub, du, lane);
    }
    const unsigned long long t = block_sum(c);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// 4-cliques by bit rows: the binary-encoded induced subgraphs of Almasri et
// al., "Parallel K-clique Counting on GPUs" (ICS 2022).  For every vertex u
// whose out-list L = out(u) has d >= 3 entries, row i (the row of oriented
// edge u -> L[i]) has bit p set iff L[p] is in out(L[i]), i.e. it is the set
// W = out(u) & out(L[i]) over positions of L.  A 4-clique whose lowest vertex
// is u, next L[i], then L[j], then L[p] is bit p of row_i & row_j, so the
// count at u is the sum over rows i and set bits j of popcount(row_i & row_j).
// Vertex u's rows take d * ceil(d / 64) words from boff[u]; a vertex whose
// out-list is longer than sqrt(2m) cannot exist, so one vertex needs at most
// 2^16 rows of 2^10 words.
constexpr int kK4RowLds = 256;   // words of one row staged in LDS (out-lists up to 16384)

__device__ __forceinline__ uint64_t lower_u32(const uint32_t *a, uint64_t n, uint32_t x)
{
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// words of each vertex's rows (0 below three out-neighbours)
__global__ void k_k4_words(const uint64_t *ooff, uint32_t n, uint64_t *words)
{
    GRID_STRIDE(u, (uint64_t)n + 1)
    {
        const uint64_t d = u < n ? ooff[u + 1] - ooff[u] : 0;
        words[u] = d >= 3 ? d * ((d + 63) >> 6) : 0;
    }
}

// One wavefront per oriented edge of [e0, e1) writes its row, whichever way
// probes less: ballots over 64 positions of L at a time, each lane searching
// its entry of L in out(v); or, for a short out(v), each lane searching one
// entry of out(v) in L and setting its bit in an LDS copy of the row.
__global__ __launch_bounds__(256) void k_k4_rows(const uint64_t *ooff, const uint32_t *onb, const uint32_t *osrc,
                                                 const uint64_t *boff, uint64_t bbase, uint64_t e0, uint64_t e1,
                                                 uint64_t *bits)
{
    __shared__ unsigned long long srows[4 * kK4RowLds];
    const uint32_t lane = threadIdx.x & 63;
    unsigned long long *const srow = srows + (threadIdx.x >> 6) * kK4RowLds;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t e = e0 + wave; e < e1; e += nwaves) {
        const uint32_t u = osrc[e], v = onb[e];
        const uint64_t lb = ooff[u], d = ooff[u + 1] - lb;
        if (d < 3) continue;
        const uint64_t R = (d + 63) >> 6;
        uint64_t *const row = bits + (boff[u] - bbase) + (e - lb) * R;
        const uint32_t *const L = onb + lb;
        const uint64_t vb = ooff[v], dv = ooff[v + 1] - vb;
        const uint32_t *const ov = onb + vb;
        const bool staged = R <= (uint64_t)kK4RowLds && ((dv + 63) >> 6) * clog2_u64(d + 1) < R * clog2_u64(dv + 1);
        if (!staged) {
            // two positions of L per lane and pass, their searches interleaved
            for (uint64_t kb = 0; kb < R; kb += 2) {
                const uint64_t j0 = kb * 64 + lane, j1 = j0 + 64;
                const uint32_t x0 = j0 < d ? L[j0] : 0, x1 = j1 < d ? L[j1] : 0;
                uint64_t lo0 = 0, hi0 = j0 < d ? dv : 0, lo1 = 0, hi1 = j1 < d ? dv : 0;
                while (lo0 < hi0 || lo1 < hi1) {
                    const uint64_t m0 = (lo0 + hi0) >> 1, m1 = (lo1 + hi1) >> 1;
                    const uint32_t y0 = lo0 < hi0 ? ov[m0] : 0, y1 = lo1 < hi1 ? ov[m1] : 0;
                    if (lo0 < hi0) { if (y0 < x0) lo0 = m0 + 1; else hi0 = m0; }
                    if (lo1 < hi1) { if (y1 < x1) lo1 = m1 + 1; else hi1 = m1; }
                }
                const bool h0 = j0 < d && lo0 < dv && ov[lo0] == x0;
                const bool h1 = j1 < d && lo1 < dv && ov[lo1] == x1;
                const uint64_t w0 = __ballot(h0), w1 = __ballot(h1);
                if (lane == 0) {
                    row[kb] = w0;
                    if (kb + 1 < R) row[kb + 1] = w1;
                }
            }
            continue;
        }
        for (uint64_t t = lane; t < R; t += 64) srow[t] = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (uint64_t xb = 0; xb < dv; xb += 64) {
            const uint64_t t = xb + lane;
            if (t < dv) {
                const uint32_t x = ov[t];
                const uint64_t j = lower_u32(L, d, x);
                if (j < d && L[j] == x) atomicOr(&srow[j >> 6], 1ull << (j & 63));
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (uint64_t t = lane; t < R; t += 64) row[t] = srow[t];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();   // the LDS row is rewritten by the next edge
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

// Rows of at most this many words are counted one thread per row, longer
// ones one wavefront per row.
constexpr uint64_t kK4NarrowWords = 8;

// One thread per narrow row of [e0, e1): the sum over its set bits j of
// popcount(row_i & row_j).
__global__ void k_k4_count(const uint64_t *ooff, const uint32_t *osrc, const uint64_t *boff, uint64_t bbase,
                           uint64_t e0, uint64_t e1, const uint64_t *bits, unsigned long long *partial)
{
    unsigned long long c = 0;
    for (uint64_t e = e0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < e1;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = osrc[e];
        const uint64_t lb = ooff[u], d = ooff[u + 1] - lb;
        if (d < 3) continue;
        const uint64_t R = (d + 63) >> 6;
        if (R > kK4NarrowWords) continue;
        const uint64_t *const rb = bits + (boff[u] - bbase);
        const uint64_t *const ri = rb + (e - lb) * R;
        if (R == 1) {
            const uint64_t w0 = ri[0];
            for (uint64_t wb = w0; wb; wb &= wb - 1) c += __popcll(w0 & rb[__builtin_ctzll(wb)]);
            continue;
        }
        for (uint64_t kb = 0; kb < R; ++kb) {
            for (uint64_t wb = ri[kb]; wb; wb &= wb - 1) {
                const uint64_t *const rj = rb + (kb * 64 + __builtin_ctzll(wb)) * R;
                for (uint64_t t = 0; t < R; ++t) c += __popcll(ri[t] & rj[t]);
            }
        }
    }
    const unsigned long long t = block_sum(c);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// One wavefront per wide row of [e0, e1), the lanes sharing its words (held
// in registers up to 64 words).  The wave takes 64 consecutive rows at a time
// and walks the wide ones among them, so a hub's rows spread over many waves.
__global__ __launch_bounds__(256) void k_k4_count_wide(const uint64_t *ooff, const uint32_t *osrc,
                                                       const uint64_t *boff, uint64_t bbase, uint64_t e0,
                                                       uint64_t e1, const uint64_t *bits,
                                                       unsigned long long *partial)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    unsigned long long c = 0;
    for (uint64_t b = e0 + wave * 64; b < e1; b += nwaves * 64) {
        uint64_t d = 0;
        if (b + lane < e1) {
            const uint32_t u = osrc[b + lane];
            d = ooff[u + 1] - ooff[u];
        }
        for (uint64_t m = __ballot(((d + 63) >> 6) > kK4NarrowWords); m; m &= m - 1) {
            const uint64_t e = b + __builtin_ctzll(m);
            const uint32_t u = osrc[e];
            const uint64_t lb = ooff[u], R = (ooff[u + 1] - lb + 63) >> 6;
            const uint64_t *const rb = bits + (boff[u] - bbase);
            const uint64_t *const ri = rb + (e - lb) * R;
            if (R <= 64) {
                // four set bits per pass, so four row loads are in flight
                const uint64_t mine = lane < R ? ri[lane] : 0;
                for (uint64_t kb = 0; kb < R; ++kb) {
                    for (uint64_t wb = ri[kb]; wb;) {
                        uint64_t j[4];
                        bool ok[4];
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            ok[q] = wb != 0;
                            j[q] = kb * 64 + (ok[q] ? (uint64_t)__builtin_ctzll(wb) : 0);
                            wb &= wb - 1;
                        }
                        uint64_t x[4];
#pragma unroll
                        for (int q = 0; q < 4; ++q) x[q] = lane < R ? rb[(ok[q] ? j[q] : j[0]) * R + lane] : 0;
#pragma unroll
                        for (int q = 0; q < 4; ++q) c += ok[q] ? __popcll(mine & x[q]) : 0;
                    }
                }
                continue;
            }
            for (uint64_t kb = 0; kb < R; ++kb) {
                for (uint64_t wb = ri[kb]; wb; wb &= wb - 1) {
                    const uint64_t *const rj = rb + (kb * 64 + __builtin_ctzll(wb)) * R;
                    for (uint64_t t = lane; t < R; t += 64) c += __popcll(ri[t] & rj[t]);
                }
            }
        }
    }
    const unsigned long long t = block_sum(c);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void k_wedges_paths(const uint64_t *off, const uint64_t *sadj, const uint32_t *sdeg, uint32_t n,
                               unsigned long long *wed, unsigned long long *paths)
{
    unsigned long long w = 0, p = 0;
    WAVE_STRIDE(v, (uint64_t)n)
    {
        const unsigned long long d = sdeg[v];
        if ((threadIdx.x & 63) == 0) w += d * (d - (d > 0)) / 2;
        const uint64_t b = off[v], e = off[v + 1];
        uint32_t carry = 0xffffffffu;
        for (uint64_t base = b; base < e; base += 64) {
            const uint32_t nb = nbr_or_none(sadj, base + (threadIdx.x & 63), e);
            if (first_of_run(nb, carry) && (uint32_t)v < nb) p += (d - 1) * (unsigned long long)(sdeg[nb] - 1);
        }
    }
    const unsigned long long tw = block_sum(w);
    __syncthreads();
    const unsigned long long tp = block_sum(p);
    if (threadIdx.x == 0) { wed[blockIdx.x] = tw; paths[blockIdx.x] = tp; }
}

// 5-stars: sum over vertices of C(d, 4) (naive_implementation/FiveStar.cpp:51-59)
__global__ void k_stars4(const uint32_t *sdeg, uint32_t n, unsigned long long *part)
{
    unsigned long long c = 0;
    GRID_STRIDE(v, (uint64_t)n)
    {
        const unsigned __int128 d = sdeg[v];
        if (d >= 4) c += (unsigned long long)(d * (d - 1) * (d - 2) * (d - 3) / 24);
    }
    const unsigned long long t = block_sum(c);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_widen(const uint32_t *a, uint32_t n, uint64_t *b)
{
    GRID_STRIDE(v, (uint64_t)n) b[v] = a[v];
}

// full simple adjacency: distinct neighbours ascending by id, and each entry's owner
__global__ void k_full_fill(const uint64_t *off, const uint64_t *sadj, uint32_t n, const uint64_t *foff,
                            uint32_t *fnb, uint32_t *fsrc)
{
    GRID_STRIDE(v, (uint64_t)n)
    {
        uint64_t w = foff[v];
        uint32_t prev = 0xffffffffu;
        for (uint64_t i = off[v]; i < off[v + 1]; ++i) {
            const uint32_t nb = (uint32_t)(sadj[i] >> 32);
            if (nb != prev) { fnb[w] = nb; fsrc[w] = (uint32_t)v; ++w; }
            prev = nb;
        }
    }
}

__device__ __forceinline__ uint64_t isect_count(const uint32_t *a, uint64_t na, const uint32_t *b, uint64_t nb)
{
    if (na > nb) { const uint32_t *t = a; a = b; b = t; const uint64_t x = na; na = nb; nb = x; }
    uint64_t c = 0;
    for (uint64_t i = 0; i < na; ++i) c += bsearch_u32(b, nb, a[i]);
    return c;
}

// 5-vertex houses (apex x over the edge {a, b} of the square a-c-d-b):
// one wavefront per ordered edge (a, b) counts the embeddings with a, b as
// the roof's base, sum over P = {(c, d): c in N(a)\b, d in N(b)\a, c ~ d}
// of (t_ab - [c in T] - [d in T]) with T = N(a) & N(b), t_ab = |T|:
//   t_ab * (S - (d_a - 1)) - sum_{c in T}(|N(c) & N(b)| - 1) - sum_{d in T}(|N(d) & N(a)| - 1),
//   S = sum_{c in N(a)\b} |N(c) & N(b)|.
// The house count is the embedding total over 2 (its automorphisms).
__global__ __launch_bounds__(256) void k_house(const uint64_t *foff, const uint32_t *fnb, const uint32_t *fsrc,
                                               uint64_t ne, unsigned long long *partial)
{
    const int lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    unsigned long long acc = 0;
    for (uint64_t e = wave; e < ne; e += nwaves) {
        const uint32_t a = fsrc[e], b = fnb[e];
        const uint32_t *Na = fnb + foff[a], *Nb = fnb + foff[b];
        const uint64_t da = foff[a + 1] - foff[a], db = foff[b + 1] - foff[b];
        unsigned long long S = 0, t = 0, X = 0, Y = 0;
        for (uint64_t i = lane; i < da; i += 64) {
            const uint32_t c = Na[i];
            if (c == b) continue;
            const uint32_t *Nc = fnb + foff[c];
            const uint64_t dc = foff[c + 1] - foff[c];
            const uint64_t k = isect_count(Nc, dc, Nb, db);
            S += k;
            if (bsearch_u32(Nb, db, c)) {
                t += 1;
                X += k - 1;
                Y += isect_count(Nc, dc, Na, da) - 1;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            S += __shfl_xor(S, o, 64);
            t += __shfl_xor(t, o, 64);
            X += __shfl_xor(X, o, 64);
            Y += __shfl_xor(Y, o, 64);
        }
        if (lane == 0) acc += t * (S - (da - 1)) - X - Y;
    }
    const unsigned long long tot = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// degree moments: sum d^2 and sum d^3 (as doubles) per block
__global__ void k_deg_moments(const uint64_t *off, uint32_t n, double *p2, double *p3)
{
    __shared__ double sm[2][4];
    double a = 0, b = 0;
    GRID_STRIDE(v, (uint64_t)n)
    {
        const double d = (double)(off[v + 1] - off[v]);
        a += d * d;
        b += d * d * d;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { a += __shfl_xor(a, o, 64); b += __shfl_xor(b, o, 64); }
    if ((threadIdx.x & 63) == 0) { sm[0][threadIdx.x >> 6] = a; sm[1][threadIdx.x >> 6] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0, y = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { x += sm[0][w]; y += sm[1][w]; }
        p2[blockIdx.x] = x;
        p3[blockIdx.x] = y;
    }
}

int host_sum(const unsigned long long *d, int n, hipStream_t st, unsigned long long *out)
{
    std::vector<unsigned long long> h(n);
    ZG_HIP(hipMemcpyAsync(h.data(), d, 8ull * n, hipMemcpyDeviceToHost, st));
    ZG_HIP(hipStreamSynchronize(st));
    unsigned long long s = 0;
    for (auto x : h) s += x;
    *out = s;
    return ZG_OK;
}

}  // namespace
}  // namespace zg

using namespace zg;

extern "C" int zg_exact_counts(zg_graph *g, int what, uint64_t *out)
{
    if (!g || !out) return set_error(ZG_EINVAL, "null argument");
    if (what < 0 || what > 5)
        return set_error(ZG_EINVAL, "what must be 0 (triangles), 1 (4-cliques), 2 (wedges), 3 (3-edge paths), "
                                    "4 (5-stars), 5 (5-houses)");
    const uint32_t n = g->n;
    hipStream_t st = g->stream;
    if (g->m == 0) { *out = 0; return ZG_OK; }
    XB<uint32_t> sdeg;
    XALLOC(sdeg, n);
    hipLaunchKernelGGL(k_simple_deg, dim3(gridx(n)), dim3(256), 0, st, g->off, g->sadj, n, sdeg.p);
    if (what == 4) {
        const int grid = gridx(n);
        XB<unsigned long long> part;
        XALLOC(part, grid);
        hipLaunchKernelGGL(k_stars4, dim3(grid), dim3(256), 0, st, sdeg.p, n, part.p);
        ZG_HIP(hipGetLastError());
        unsigned long long s = 0;
        int rc = host_sum(part.p, grid, st, &s);
        if (rc) return rc;
        *out = s;
        return ZG_OK;
    }
    if (what == 5) {
        XB<uint64_t> fdeg, foff;
        XALLOC(fdeg, (uint64_t)n + 1); XALLOC(foff, (uint64_t)n + 1);
        hipLaunchKernelGGL(k_widen, dim3(gridx(n)), dim3(256), 0, st, sdeg.p, n, fdeg.p);
        ZG_HIP(hipMemsetAsync(fdeg.p + n, 0, 8, st));
        size_t tb = 0;
        ZG_HIP(rocprim::exclusive_scan(nullptr, tb, fdeg.p, foff.p, (uint64_t)0, (size_t)n + 1,
                                       rocprim::plus<uint64_t>(), st));
        XB<char> tmp;
        XALLOC(tmp, tb);
        ZG_HIP(rocprim::exclusive_scan(tmp.p, tb, fdeg.p, foff.p, (uint64_t)0, (size_t)n + 1,
                                       rocprim::plus<uint64_t>(), st));
        uint64_t ne = 0;
        ZG_HIP(hipMemcpyAsync(&ne, foff.p + n, 8, hipMemcpyDeviceToHost, st));
        ZG_HIP(hipStreamSynchronize(st));
        XB<uint32_t> fnb, fsrc;
        XALLOC(fnb, ne); XALLOC(fsrc, ne);
        hipLaunchKernelGGL(k_full_fill, dim3(gridx(n)), dim3(256), 0, st, g->off, g->sadj, n, foff.p, fnb.p, fsrc.p);
        ZG_HIP(hipGetLastError());
        const int grid = (int)std::min<uint64_t>((uint64_t)device_cus() * 8, std::max<uint64_t>(1, (ne + 3) / 4));
        XB<unsigned long long> part;
        XALLOC(part, grid);
        hipLaunchKernelGGL(k_house, dim3(grid), dim3(256), 0, st, foff.p, fnb.p, fsrc.p, ne, part.p);
        ZG_HIP(hipGetLastError());
        unsigned long long s = 0;
        int rc = host_sum(part.p, grid, st, &s);
        if (rc) return rc;
        *out = s / 2;
        return ZG_OK;
    }
    if (what >= 2) {
        const int grid = gridx(n);
        XB<unsigned long long> pw, pp;
        XALLOC(pw, grid); XALLOC(pp, grid);
        hipLaunchKernelGGL(k_wedges_paths, dim3(grid), dim3(256), 0, st, g->off, g->sadj, sdeg.p, n, pw.p, pp.p);
        ZG_HIP(hipGetLastError());
        unsigned long long w = 0, p = 0;
        int rc = host_sum(pw.p, grid, st, &w);
        if (!rc) rc = host_sum(pp.p, grid, st, &p);
        if (rc) return rc;
        if (what == 2) { *out = w; return ZG_OK; }
        uint64_t T = 0;
        rc = zg_exact_counts(g, 0, &T);
        if (rc) return rc;
        *out = p - 3 * T;
        return ZG_OK;
    }
    // oriented out-lists
    XB<uint64_t> odeg, ooff;
    XALLOC(odeg, (uint64_t)n + 1); XALLOC(ooff, (uint64_t)n + 1);
    ZG_HIP(hipMemsetAsync(odeg.p + n, 0, 8, st));
    hipLaunchKernelGGL(k_out_deg, dim3(gridx(n)), dim3(256), 0, st, g->off, g->sadj, sdeg.p, n, odeg.p);
    size_t tb = 0;
    ZG_HIP(rocprim::exclusive_scan(nullptr, tb, odeg.p, ooff.p, (uint64_t)0, (size_t)n + 1, rocprim::plus<uint64_t>(), st));
    XB<char> tmp;
    XALLOC(tmp, tb);
    ZG_HIP(rocprim::exclusive_scan(tmp.p, tb, odeg.p, ooff.p, (uint64_t)0, (size_t)n + 1, rocprim::plus<uint64_t>(), st));
    uint64_t ne = 0;
    ZG_HIP(hipMemcpyAsync(&ne, ooff.p + n, 8, hipMemcpyDeviceToHost, st));
    size_t tr = 0;
    XB<uint64_t> mx;
    XALLOC(mx, 1);
    ZG_HIP(rocprim::reduce(nullptr, tr, odeg.p, mx.p, (uint64_t)0, (size_t)n, rocprim::maximum<uint64_t>(), st));
    XB<char> tmp2;
    XALLOC(tmp2, tr);
    ZG_HIP(rocprim::reduce(tmp2.p, tr, odeg.p, mx.p, (uint64_t)0, (size_t)n, rocprim::maximum<uint64_t>(), st));
    uint64_t maxout = 0;
    ZG_HIP(hipMemcpyAsync(&maxout, mx.p, 8, hipMemcpyDeviceToHost, st));
    ZG_HIP(hipStreamSynchronize(st));
    XB<uint32_t> onb, osrc;
    XALLOC(onb, ne); XALLOC(osrc, ne);
    hipLaunchKernelGGL(k_out_fill, dim3(gridx(n)), dim3(256), 0, st, g->off, g->sadj, sdeg.p, n, ooff.p, onb.p, osrc.p);
    ZG_HIP(hipGetLastError());
    if (what == 0) {
        const int tgrid = (int)std::min<uint64_t>((uint64_t)device_cus() * 16, std::max<uint64_t>(1, (ne + 3) / 4));
        XB<unsigned long long> tpart;
        XALLOC(tpart, tgrid);
        hipLaunchKernelGGL(k_tri_wave, dim3(tgrid), dim3(256), 0, st, ooff.p, onb.p, osrc.p, ne, tpart.p);
        ZG_HIP(hipGetLastError());
        unsigned long long s = 0;
        int rc = host_sum(tpart.p, tgrid, st, &s);
        if (rc) return rc;
        *out = s;
        return ZG_OK;
    }
    // 4-cliques: bit rows in vertex ranges whose rows fit the scratch budget
    XB<uint64_t> words, boff;
    XALLOC(words, (uint64_t)n + 1); XALLOC(boff, (uint64_t)n + 1);
    hipLaunchKernelGGL(k_k4_words, dim3(gridx((uint64_t)n + 1)), dim3(256), 0, st, ooff.p, n, words.p);
    ZG_HIP(hipGetLastError());
    size_t tk = 0;
    ZG_HIP(rocprim::exclusive_scan(nullptr, tk, words.p, boff.p, (uint64_t)0, (size_t)n + 1,
                                   rocprim::plus<uint64_t>(), st));
    XB<char> tmpk;
    XALLOC(tmpk, tk);
    ZG_HIP(rocprim::exclusive_scan(tmpk.p, tk, words.p, boff.p, (uint64_t)0, (size_t)n + 1,
                                   rocprim::plus<uint64_t>(), st));
    auto read_u64 = [&](const uint64_t *p, uint64_t *v) -> int {
        ZG_HIP(hipMemcpyAsync(v, p, 8, hipMemcpyDeviceToHost, st));
        ZG_HIP(hipStreamSynchronize(st));
        return ZG_OK;
    };
    uint64_t total_words = 0;
    int rc = read_u64(boff.p + n, &total_words);
    if (rc) return rc;
    if (total_words == 0) { *out = 0; return ZG_OK; }
    size_t free_b = 0, total_b = 0;
    ZG_HIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t one_vertex = maxout * ((maxout + 63) >> 6);
    uint64_t budget = std::min<uint64_t>(16ull << 30, std::max<uint64_t>(1ull << 30, free_b / 4)) / 8;
    if (const char *ev = getenv("ZGRAPH_K4_SCRATCH_WORDS")) budget = strtoull(ev, nullptr, 10);   // smaller passes
    budget = std::max(budget, one_vertex);
    XB<uint64_t> bits;
    XALLOC(bits, std::min(total_words, budget));
    const int rgrid = (int)std::min<uint64_t>((uint64_t)device_cus() * 16, std::max<uint64_t>(1, (ne + 3) / 4));
    const int cgrid = gridx(ne);
    const int wgrid = (int)std::min<uint64_t>((uint64_t)device_cus() * 16, std::max<uint64_t>(1, (ne + 255) / 256));
    XB<unsigned long long> part, wpart;
    XALLOC(part, cgrid); XALLOC(wpart, wgrid);
    unsigned long long s = 0;
    for (uint32_t u0 = 0; u0 < n;) {
        uint64_t b0 = 0;
        rc = read_u64(boff.p + u0, &b0);
        if (rc) return rc;
        uint32_t u1 = n;
        if (total_words - b0 > budget) {   // largest u1 with boff[u1] - boff[u0] <= budget
            uint32_t lo = u0 + 1, hi = n;
            while (lo < hi) {
                const uint32_t mid = lo + (hi - lo + 1) / 2;
                uint64_t bm = 0;
                rc = read_u64(boff.p + mid, &bm);
                if (rc) return rc;
                if (bm - b0 <= budget) lo = mid; else hi = mid - 1;
            }
            u1 = lo;
        }
        uint64_t e0 = 0, e1 = 0;
        rc = read_u64(ooff.p + u0, &e0);
        if (!rc) rc = read_u64(ooff.p + u1, &e1);
        if (rc) return rc;
        if (e1 > e0) {
            hipLaunchKernelGGL(k_k4_rows, dim3(rgrid), dim3(256), 0, st, ooff.p, onb.p, osrc.p, boff.p, b0, e0, e1,
                               bits.p);
            ZG_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_k4_count, dim3(cgrid), dim3(256), 0, st, ooff.p, osrc.p, boff.p, b0, e0, e1, bits.p,
                               part.p);
            ZG_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_k4_count_wide, dim3(wgrid), dim3(256), 0, st, ooff.p, osrc.p, boff.p, b0, e0, e1,
                               bits.p, wpart.p);
            ZG_HIP(hipGetLastError());
            unsigned long long ps = 0, pw = 0;
            rc = host_sum(part.p, cgrid, st, &ps);
            if (!rc) rc = host_sum(wpart.p, wgrid, st, &pw);
            if (rc) return rc;
            s += ps + pw;
        }
        u0 = u1;
    }
    *out = s;
    return ZG_OK;
}

extern "C" int zg_graph_degree_moments(zg_graph *g, double *sum_d2, double *sum_d3)
{
    if (!g || !sum_d2 || !sum_d3) return set_error(ZG_EINVAL, "null argument");
    const int grid = gridx(g->n);
    XB<double> p2, p3;
    XALLOC(p2, grid); XALLOC(p3, grid);
    hipLaunchKernelGGL(k_deg_moments, dim3(grid), dim3(256), 0, g->stream, g->off, g->n, p2.p, p3.p);
    ZG_HIP(hipGetLastError());
    std::vector<double> h2(grid), h3(grid);
    ZG_HIP(hipMemcpyAsync(h2.data(), p2.p, 8ull * grid, hipMemcpyDeviceToHost, g->stream));
    ZG_HIP(hipMemcpyAsync(h3.data(), p3.p, 8ull * grid, hipMemcpyDeviceToHost, g->stream));
    ZG_HIP(hipStreamSynchronize(g->stream));
    double a = 0, b = 0;
    for (int i = 0; i < grid; ++i) { a += h2[i]; b += h3[i]; }
    *sum_d2 = a;
    *sum_d3 = b;
    return ZG_OK;
}
