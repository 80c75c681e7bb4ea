// zg_exact.hip -- exact subgraph counts on the device, the scalable
// counterpart of naive_implementation/*.cpp (non-induced counts on the
// simple graph underlying the edge list):
//   triangles  naive_implementation/TriangleCounting.cpp:44-72
//   4-cliques  naive_implementation/FourClique.cpp:44-78
//   wedges     naive_implementation/ThreeChain.cpp:44-51 (sum of C(d, 2))
//   3-paths    naive_implementation/FourChain.cpp:46-66 (sum over edges (du-1)(dv-1) - 3T)
//   5-stars    naive_implementation/FiveStar.cpp:51-59 (sum of C(d, 4))
//   5-houses   naive_implementation/FiveHouse.cpp's count as sum_e t_e c4_e - 2 sum_e t_e (t_e - 1):
//              per-edge triangles (k_tri_edges) weighting the 4-cycle listing
//   5-chains   naive_implementation/FiveChain.cpp:45-72 (closed form over degrees,
//              per-vertex triangles and 4-cycles)
//   4-cycles   no reference program (the 5-chain closed form needs them)
// Method: orient every undirected pair from lower to higher (degree, id) and
// keep the out-lists sorted by neighbour id (taken from the neighbour-ordered
// copy, parallel edges collapsed), with in-lists of the tails of each head.
// Triangles by head: out(v) in an LDS cuckoo table, the out-lists of v's
// tails streamed through it (k_head_light / k_head_heavy).  4-cliques: per
// vertex u, bit rows over out(u) (row i: which of out(u) are in
// out(out(u)[i])), built from the same stream, and the sum of
// popcount(row_i & row_j) over the set bits j of each row.  The 5-chain
// count's per-vertex weighted triangle sum: one wavefront per oriented edge
// binary-searching one out-list in the other (k_tri_wave).  4-cycles: at
// their highest vertex, counting the wedges below it per opposite vertex (LDS
// hashes, sorted wedge keys and dense counters for hubs).
#include <cstdlib>
#include <cstring>

#include <type_traits>
#include <utility>

#include <rocprim/rocprim.hpp>

#include "zg_internal.h"

namespace zg {
namespace {

template <class T>
struct XB {
    T *p = nullptr;
    hipError_t alloc(size_t k) { return hipMalloc(&p, sizeof(T) * (k ? k : 1)); }
    void reset()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    ~XB() { reset(); }
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

constexpr uint32_t kNone = 0xffffffffu;   // a value no vertex id takes

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ uint64_t uni64(uint64_t x)
{
    return ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(x >> 32)) << 32) |
           __builtin_amdgcn_readfirstlane((uint32_t)x);
}

// The orientation's total order: (simple degree, id), with degrees capped at
// 65535 so the key array is 2 bytes a vertex.  A vertex of degree d below the
// cap still points only to vertices of degree >= d, so its out-list stays
// within sqrt(2m); one at the cap points to at most 2m / 65535 others, also
// within sqrt(2m) while m <= 2^31.
__device__ __forceinline__ bool before(const uint16_t *okey, uint32_t a, uint32_t b)
{
    const uint32_t da = okey[a], db = okey[b];
    return da < db || (da == db && a < b);
}

__global__ void k_order_key(const uint32_t *sdeg, uint32_t n, uint16_t *okey)
{
    GRID_STRIDE(v, (uint64_t)n) okey[v] = (uint16_t)(sdeg[v] < 65535u ? sdeg[v] : 65535u);
}

// ---- entry walk -----------------------------------------------------------------
// The neighbour-ordered lists taken as one array of 2m entries and walked 64
// entries (a group) per wavefront step, each wavefront taking a tile of kEwTile
// consecutive groups.  The owner of a lane's entry comes from the offsets of
// the next 64 vertices (one coalesced load, mostly cache hits since
// consecutive groups see the same offsets), so a hub's list is spread over as
// many wavefronts as its length asks for and a pass streams the adjacency once,
// where one wavefront per vertex would walk a hub's million entries alone.
constexpr uint64_t kEwTile = 64;   // groups per tile (4096 entries)

// largest v in [0, n] with off[v] <= x (off[0] = 0 <= x): a 64-ary search
__device__ uint32_t ew_find(const uint64_t *off, uint32_t n, uint64_t x)
{
    const uint64_t lane = threadIdx.x & 63;
    uint64_t lo = 0, hi = (uint64_t)n + 1;   // off[lo] <= x, the answer lies in [lo, hi)
    while (hi - lo > 1) {
        const uint64_t step = (hi - lo + 63) >> 6, p = lo + lane * step;
        const uint64_t bal = __ballot(p < hi && off[p] <= x);   // a prefix of the lanes, lane 0 included
        lo += (uint64_t)(63 - __clzll(bal)) * step;
        hi = hi < lo + step ? hi : lo + step;
    }
    return (uint32_t)lo;
}

struct EntryWalk {
    const uint64_t *off, *sadj;
    uint64_t m2, base;   // 2m entries; first entry of the next group
    uint32_t n, v;       // v: a vertex with off[v] <= base
    uint32_t cu, cnb;    // owner and neighbour of entry base - 1 (kNone at a list start)
    // after next(), for entry base + lane of the group just walked:
    uint32_t u, nb, rk;  // owner, neighbour and edge rank
    bool valid, first;   // entry exists; first of its run of equal neighbours in its list

    __device__ EntryWalk(const uint64_t *off_, const uint64_t *sadj_, uint32_t n_, uint64_t m2_, uint64_t b)
        : off(off_), sadj(sadj_), m2(m2_), base(b), n(n_)
    {
        v = ew_find(off, n, b);
        const bool mid = off[v] < b;
        cu = mid ? v : kNone;
        cnb = mid ? (uint32_t)(sadj[b - 1] >> 32) : kNone;
    }

    __device__ void next()
    {
        const uint32_t lane = threadIdx.x & 63;
        const uint64_t i = base + lane;
        valid = i < m2;
        const uint64_t ent = valid ? sadj[i] : ~0ull;
        nb = (uint32_t)(ent >> 32);
        rk = (uint32_t)ent;
        // owner(i) = v + #{k >= 1 : off[v + k] <= i}, 64 offsets a round: r is where
        // vertex w starts relative to the group (0 before it, 64 past it), ascending
        // over the lanes, and a binary search over the lanes counts the r <= lane
        uint32_t cnt = 0;
        for (uint64_t w0 = (uint64_t)v + 1;; w0 += 64) {
            const uint64_t w = w0 + lane;
            const uint64_t o = w <= n ? off[w] : ~0ull;
            const uint32_t r = o <= base ? 0u : (o - base >= 64 ? 64u : (uint32_t)(o - base));
            uint32_t pos = 0;
            for (uint32_t s = 32; s; s >>= 1) pos += __shfl(r, (int)(pos + s - 1), 64) <= lane ? s : 0;
            const uint32_t r63 = __shfl(r, 63, 64);
            pos += (pos == 63 && r63 <= lane) ? 1 : 0;
            cnt += pos;
            if (r63 >= 64 || w0 + 64 > n) break;   // no later vertex starts inside the group
        }
        u = v + cnt;
        uint32_t pu = __shfl_up(u, 1, 64), pn = __shfl_up(nb, 1, 64);
        if (lane == 0) { pu = cu; pn = cnb; }
        first = valid && (u != pu || nb != pn);
        cu = __shfl(u, 63, 64);
        cnb = __shfl(nb, 63, 64);
        v = cu;   // off[cu] <= base + 63
        base += 64;
    }
};

// every wavefront's tiles; g0..g1 are the tile's groups
#define EW_TILES(m2)                                                                                \
    const uint64_t ew_groups = ((m2) + 63) >> 6;                                                    \
    for (uint64_t ew_t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;                     \
         ew_t * kEwTile < ew_groups; ew_t += ((uint64_t)gridDim.x * blockDim.x) >> 6)

// bit j of fmask[g]: entry 64g + j is the first of its run of equal neighbours
// (the parallel copies of an edge collapse onto the first)
__global__ __launch_bounds__(256) void k_ew_first(const uint64_t *off, const uint64_t *sadj, uint32_t n, uint64_t m2,
                                                  uint64_t *fmask)
{
    EW_TILES(m2)
    {
        const uint64_t g0 = ew_t * kEwTile, g1 = g0 + kEwTile < ew_groups ? g0 + kEwTile : ew_groups;
        EntryWalk w(off, sadj, n, m2, g0 << 6);
        for (uint64_t g = g0; g < g1; ++g) {
            w.next();
            const uint64_t bal = __ballot(w.first);
            if ((threadIdx.x & 63) == 0) fmask[g] = bal;
        }
    }
}

// bit j of omask[g]: entry 64g + j is the first of its run and its neighbour
// comes after its owner in the order (an out-list entry); with IN, bit j of
// imask[g] the same for a neighbour before its owner (an in-list entry).
template <bool IN>
__global__ __launch_bounds__(256) void k_ew_orient(const uint64_t *off, const uint64_t *sadj, uint32_t n, uint64_t m2,
                                                   const uint16_t *okey, uint64_t *omask, uint64_t *imask)
{
    EW_TILES(m2)
    {
        const uint64_t g0 = ew_t * kEwTile, g1 = g0 + kEwTile < ew_groups ? g0 + kEwTile : ew_groups;
        EntryWalk w(off, sadj, n, m2, g0 << 6);
        for (uint64_t g = g0; g < g1; ++g) {
            w.next();
            const bool fwd = w.first && before(okey, w.u, w.nb);
            const uint64_t bo = __ballot(fwd);
            const uint64_t bi = IN ? __ballot(w.first && !fwd) : 0;   // no self-loops: first and not forward is backward
            if ((threadIdx.x & 63) == 0) {
                omask[g] = bo;
                if (IN) imask[g] = bi;
            }
        }
    }
}

// The entries flagged in mask, in entry order (lists stay in neighbour order):
// the k-th flagged entry's neighbour to nbr[k] and its owner to own[k] (own may
// be null); pre[g] = flagged entries before group g.  With pos_of_rank, also
// pos_of_rank[edge rank] = k; with edge_of_rank, nbr[k] = edge_of_rank[edge
// rank] instead of the neighbour (an in-list entry becomes the index of its
// twin in the out-lists: both copies of an edge carry its rank, and a run of
// parallel copies keeps its lowest rank first in both lists); with rank_out,
// rank_out[k] = the edge rank.
__global__ __launch_bounds__(256) void k_ew_compact(const uint64_t *off, const uint64_t *sadj, uint32_t n, uint64_t m2,
                                                    const uint64_t *mask, const uint64_t *pre, uint32_t *nbr,
                                                    uint32_t *own, uint32_t *pos_of_rank,
                                                    const uint32_t *edge_of_rank, uint32_t *rank_out)
{
    const uint32_t lane = threadIdx.x & 63;
    EW_TILES(m2)
    {
        const uint64_t g0 = ew_t * kEwTile, g1 = g0 + kEwTile < ew_groups ? g0 + kEwTile : ew_groups;
        EntryWalk w(off, sadj, n, m2, g0 << 6);
        for (uint64_t g = g0; g < g1; ++g) {
            w.next();
            const uint64_t gu = uni64(g), mk = mask[gu];
            if ((mk >> lane) & 1) {
                const uint64_t at = pre[gu] + (uint64_t)__popcll(mk & ((1ull << lane) - 1));
                nbr[at] = edge_of_rank ? edge_of_rank[w.rk] : w.nb;
                if (own) own[at] = w.u;
                if (pos_of_rank) pos_of_rank[w.rk] = (uint32_t)at;
                if (rank_out) rank_out[at] = w.rk;
            }
        }
    }
}

// List offsets of the flagged entries: out[w] = flagged entries before off[w]
// (off[n] = m2 gives the total, as pre[groups] holds it and mask[groups] = 0).
__global__ void k_ew_offsets(const uint64_t *off, uint32_t n, const uint64_t *mask, const uint64_t *pre,
                             uint64_t *out)
{
    GRID_STRIDE(w, (uint64_t)n + 1)
    {
        const uint64_t i = off[w], g = i >> 6;
        out[w] = pre[g] + (uint64_t)__popcll(mask[g] & ((1ull << (i & 63)) - 1));
    }
}

// ---- out-lists in the orientation's order ------------------------------------
// The 4-clique counter takes out-lists of kSortMin < d <= kSortMax entries
// sorted by the orientation's order instead of by id: every entry of out(v)
// comes after v, so in a sorted out(u) the entries that can lie in out(v) for
// v = out(u)[i] are those after position i (half the row stream, rows that are
// zero below their diagonal).  Shorter lists stay by id (their work is small),
// longer ones too (a table holds at most kSortMax keys).  The triangle counter
// keeps every list by id: there the saved stream costs as much as the sort and
// the in-list transposition it needs.
constexpr uint64_t kSortMin = 64, kSortMax = 4096;

__device__ __forceinline__ bool order_sorted(uint64_t d) { return d > kSortMin && d <= kSortMax; }

// vertices whose out-list gets sorted, in any order
__global__ void k_collect_sorted(const uint64_t *ooff, uint32_t n, uint32_t *list, unsigned int *count)
{
    GRID_STRIDE(v, (uint64_t)n)
    {
        if (order_sorted(ooff[v + 1] - ooff[v])) list[atomicAdd(count, 1u)] = (uint32_t)v;
    }
}

// One block per listed vertex: bitonic sort of its out-list in LDS by
// (order key, id), carrying the edge ranks, and pos_of_rank of each entry's
// edge to its new position.
__global__ __launch_bounds__(256) void k_sort_lists(const uint64_t *ooff, const uint32_t *list, uint32_t nlist,
                                                    const uint16_t *okey, uint32_t *onb, const uint32_t *orank,
                                                    uint32_t *pos_of_rank)
{
    __shared__ uint64_t key[kSortMax];
    __shared__ uint32_t val[kSortMax];
    for (uint32_t li = blockIdx.x; li < nlist; li += gridDim.x) {
        const uint32_t u = list[li];
        const uint64_t b = ooff[u], d = ooff[u + 1] - b;
        uint32_t P = 2 * kSortMin;
        while (P < d) P <<= 1;
        for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) {
            if (t < d) {
                const uint32_t x = onb[b + t];
                key[t] = ((uint64_t)okey[x] << 32) | x;
                val[t] = orank[b + t];
            } else {
                key[t] = ~0ull;
                val[t] = 0;
            }
        }
        __syncthreads();
        for (uint32_t k = 2; k <= P; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) {
                    const uint32_t o = t ^ j;
                    if (o > t) {
                        const uint64_t a = key[t], c = key[o];
                        if ((a > c) == ((t & k) == 0)) {
                            key[t] = c;
                            key[o] = a;
                            const uint32_t va = val[t];
                            val[t] = val[o];
                            val[o] = va;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (uint32_t t = threadIdx.x; t < d; t += blockDim.x) {
            onb[b + t] = (uint32_t)key[t];
            pos_of_rank[val[t]] = (uint32_t)(b + t);
        }
        __syncthreads();
    }
}

struct Popc {
    __device__ __host__ uint64_t operator()(uint64_t x) const { return (uint64_t)__builtin_popcountll(x); }
};

// simple degree from list offsets (first-of-run counts, or the lists themselves
// on a graph without parallel edges)
__global__ void k_deg_from_offsets(const uint64_t *o, uint32_t n, uint32_t *sdeg)
{
    GRID_STRIDE(v, (uint64_t)n) sdeg[v] = (uint32_t)(o[v + 1] - o[v]);
}

// largest gap o[v + 1] - o[v] (the longest list)
__global__ void k_max_gap(const uint64_t *o, uint32_t n, unsigned long long *mx)
{
    unsigned long long b = 0;
    GRID_STRIDE(v, (uint64_t)n)
    {
        const unsigned long long d = o[v + 1] - o[v];
        b = d > b ? d : b;
    }
    for (int s = 32; s > 0; s >>= 1) {
        const unsigned long long x = __shfl_xor(b, s, 64);
        b = x > b ? x : b;
    }
    if ((threadIdx.x & 63) == 0 && b) atomicMax(mx, b);
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

// (lo, hi) += (blo, bhi) in 128 bits
__device__ __forceinline__ void add128(unsigned long long &lo, unsigned long long &hi, unsigned long long blo,
                                       unsigned long long bhi)
{
    const unsigned long long l = lo + blo;
    hi += bhi + (l < lo ? 1 : 0);
    lo = l;
}

__device__ __forceinline__ unsigned long long block_sum(unsigned long long v)
{
    __shared__ unsigned long long sm[16];   // blocks of up to 1024 threads
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sm[w];
    return t;
}

// (lo, hi) summed over the workgroup in 128 bits; thread 0 writes the pair to
// part[2 * blockIdx.x], part[2 * blockIdx.x + 1]
__device__ __forceinline__ void block_sum128_store(unsigned long long lo, unsigned long long hi,
                                                   unsigned long long *part)
{
    __shared__ unsigned long long sl[4], sh[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long ol = __shfl_xor(lo, off, 64), oh = __shfl_xor(hi, off, 64);
        add128(lo, hi, ol, oh);
    }
    if ((threadIdx.x & 63) == 0) { sl[threadIdx.x >> 6] = lo; sh[threadIdx.x >> 6] = hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long L = 0, H = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) add128(L, H, sl[w], sh[w]);
        part[2 * blockIdx.x] = L;
        part[2 * blockIdx.x + 1] = H;
    }
}

// Entries of a[0, na) found in b[0, nb), counted per lane: each lane takes two
// entries of a per pass and runs their binary searches interleaved, so two
// dependent chains of loads are in flight.  With WEIGHT, *wsum also gets
// sdeg[x] - 1 for every entry x found.
template <bool WEIGHT = false>
__device__ __forceinline__ unsigned long long count_found(const uint32_t *a, uint64_t na, const uint32_t *b,
                                                          uint64_t nb, uint32_t lane,
                                                          const uint32_t *sdeg = nullptr,
                                                          unsigned long long *wsum = nullptr)
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
        const bool f0 = i0 < na && lo0 < nb && b[lo0] == x0;
        const bool f1 = i1 < na && lo1 < nb && b[lo1] == x1;
        c += (f0 ? 1 : 0) + (f1 ? 1 : 0);
        if (WEIGHT) {
            if (f0) *wsum += sdeg[x0] - 1;
            if (f1) *wsum += sdeg[x1] - 1;
        }
    }
    return c;
}

// One wavefront per oriented edge u -> v counts |out(u) & out(v)|, the lanes
// taking the entries of whichever list costs fewer passes of 128 entries
// times search depth in the other: a short list is searched in a long one,
// but 64 entries of a long list searched in a list of 3 cost one pass of
// depth 2 instead of a pass of depth 7.
// With WEIGHT the kernel also sums, over the triangles it finds, the
// simple degrees minus one of their three vertices (sum_v (d_v - 1) t_v, with
// t_v the triangles at v) into wpartial.
template <bool WEIGHT>
__global__ __launch_bounds__(256) void k_tri_wave(const uint64_t *ooff, const uint32_t *onb, const uint32_t *osrc,
                                                  uint64_t ne, const uint32_t *sdeg, unsigned long long *partial,
                                                  unsigned long long *wpartial)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    unsigned long long c = 0, wsum = 0;
    for (uint64_t e = wave; e < ne; e += nwaves) {
        const uint32_t u = osrc[e], v = onb[e];
        const uint64_t ub = ooff[u], du = ooff[u + 1] - ub, vb = ooff[v], dv = ooff[v + 1] - vb;
        const bool from_u = ((du + 127) >> 7) * clog2_u64(dv + 1) <= ((dv + 127) >> 7) * clog2_u64(du + 1);
        const unsigned long long f =
            from_u ? count_found<WEIGHT>(onb + ub, du, onb + vb, dv, lane, sdeg, &wsum)
                   : count_found<WEIGHT>(onb + vb, dv, onb + ub, du, lane, sdeg, &wsum);
        c += f;
        if (WEIGHT) wsum += f * ((unsigned long long)sdeg[u] + sdeg[v] - 2);
    }
    const unsigned long long t = block_sum(c);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
    if (WEIGHT) {
        __syncthreads();
        const unsigned long long tw = block_sum(wsum);
        if (threadIdx.x == 0) wpartial[blockIdx.x] = tw;
    }
}

// Triangles by head: every triangle a -> b -> c (a -> c) is found once, at
// its middle vertex b, as an entry c of out(a) that is also in out(b).  A work
// item is a head v and up to kTriChunk entries of its in-list: out(v) goes
// into an LDS hash table, then the out-lists of the item's tails u stream
// through it, 64 tails at a time.  The lanes take 64 consecutive entries of
// the tails' concatenated lists per load (a wavefront-uniform cursor hands
// them out, so a long list is read in full lines and a short one costs one
// scalar step), kTriUnroll loads in flight per lane.  The stream reads sum over
// tails of dout(u)^2 list entries, each list once per edge out of its tail.
// Tables are at most a quarter full: heads with up to kTriLight
// out-neighbours take one wavefront and 2048 slots, larger ones a block of
// 16 wavefronts sharing 16384 slots; past 4096 out-neighbours (or should a
// table's build fail) a binary search of out(v) replaces the table.
constexpr uint32_t kTriChunk = 1024;            // in-list entries per wavefront work item
constexpr uint32_t kTriChunkHeavy = 8192;       // in-list entries per block work item
constexpr uint32_t kTriHeavyThreads = 1024;
constexpr int kTriWaveBits = 11, kTriBlockBits = 14;
constexpr uint64_t kTriLight = 1u << (kTriWaveBits - 2);   // out-degree limit of a wavefront item
constexpr int kTriUnroll = 8, kTriUnrollHeavy = 8, kK4Unroll = 4;   // loads per lane per step (two steps in flight)

// work items of head v: ceil(din / kTriChunk) when v has out-neighbours in
// the class (HEAVY: more than kTriLight of them)
template <bool HEAVY>
__global__ void k_tri_items_count(const uint64_t *ooff, const uint64_t *iin, uint32_t n, uint64_t *cnt)
{
    GRID_STRIDE(v, (uint64_t)n + 1)
    {
        uint64_t c = 0;
        if (v < n) {
            const uint64_t dv = ooff[v + 1] - ooff[v];
            const uint64_t chunk = HEAVY ? kTriChunkHeavy : kTriChunk;
            if (dv > 0 && (dv > kTriLight) == HEAVY) c = (iin[v + 1] - iin[v] + chunk - 1) / chunk;
        }
        cnt[v] = c;
    }
}

__global__ void k_tri_items_fill(const uint64_t *ipre, uint32_t n, uint64_t *items)
{
    GRID_STRIDE(v, (uint64_t)n)
    {
        for (uint64_t k = ipre[v]; k < ipre[v + 1]; ++k) items[k] = (v << 32) | (k - ipre[v]);
    }
}

// The head's table: 2^bits slots, a quarter full at most, cuckoo hashing
// (Pagh and Rodler): a key sits in slot h1 or slot h2 of its own, so a lookup
// is two LDS reads and no probe sequence.  Insertion evicts with atomic
// exchanges, each evicted key moving to its other slot (Alcantara et al.'s
// GPU scheme); a key still homeless after kCuckooMoves moves fails the build.
constexpr int kCuckooMoves = 256;

// Both slots from one multiplicative hash p = x * mult (mult odd, bits <= 16):
// h1 is its top `bits` bits and h2 = h1 plus an odd step taken from the next
// `bits` bits, so h2 != h1 in a table of a power-of-two size.  A build fails
// with probability of order 1 / 2^bits however empty the table (a small
// component of the slot graph with more keys than slots), so a failed build
// is redone with the next multiplier, up to kCuckooTries of them.
constexpr int kCuckooTries = 4;

__device__ __forceinline__ uint32_t cuckoo_mult(int attempt) { return 0x9E3779B1u + 0x6A09E668u * (uint32_t)attempt; }

__device__ __forceinline__ void cuckoo_slots(uint32_t x, uint32_t bits, uint32_t mult, uint32_t &h1, uint32_t &h2)
{
    const uint32_t p = x * mult, hm = (1u << bits) - 1;
    h1 = p >> (32 - bits);
    h2 = (h1 + (((p >> (32 - 2 * bits)) & hm) | 1u)) & hm;
}

// insert list[t], list[t + lanes], ...; false if one of them found no slot
__device__ __forceinline__ bool tri_insert(uint32_t *T, uint32_t bits, uint32_t mult, const uint32_t *list, uint64_t dv,
                                           uint32_t t, uint32_t lanes)
{
    bool placed = true;
    for (uint64_t k = t; k < dv; k += lanes) {
        uint32_t cur = list[k], pos, other;
        cuckoo_slots(cur, bits, mult, pos, other);
        int moves = 0;
        for (; moves < kCuckooMoves; ++moves) {
            cur = atomicExch(&T[pos], cur);
            if (cur == kNone) break;
            uint32_t p1, p2;
            cuckoo_slots(cur, bits, mult, p1, p2);
            pos = pos == p1 ? p2 : p1;
        }
        if (moves == kCuckooMoves) placed = false;
    }
    return placed;
}

__device__ __forceinline__ uint64_t readlane64(uint64_t x, uint32_t l)
{
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(x >> 32), (int)l) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, (int)l);
}

// Wavefront-uniform cursor over the concatenated out-lists of 64 tails (lane
// s holds tail s's list: d entries at b): tail cs, position cp in its list of
// cd entries at cb; fill() hands out the next U windows of 64 entries.
struct TriCursor {
    uint64_t nz, b, cb;   // nz: tails with a non-empty list
    uint32_t d, cs, cp, cd;

    __device__ TriCursor(uint64_t b_, uint32_t d_) : b(b_), d(d_)
    {
        nz = __ballot(d > 0);
        cs = nz ? (uint32_t)__builtin_ctzll(nz) : 64;
        cp = 0;
        cd = cs < 64 ? (uint32_t)__builtin_amdgcn_readlane((int)d, (int)cs) : 0;
        cb = cs < 64 ? readlane64(b, cs) : 0;
    }
    __device__ bool more() const { return cs < 64; }

    // bit k of the result: the lane's entry of window k exists (x[k] is loaded
    // either way, so no select waits for the loads)
    template <int U>
    __device__ __forceinline__ uint32_t fill(const uint32_t *onb, uint32_t (&x)[U])
    {
        const uint32_t lane = threadIdx.x & 63;
        if (cs < 64 && cd - cp > 64u * U) {   // all U windows inside the current list, short of its end
            const uint32_t *p = onb + cb + cp + lane;
#pragma unroll
            for (int k = 0; k < U; ++k) x[k] = p[64 * k];
            cp += 64 * U;
            return (1u << U) - 1;
        }
        uint32_t okbits = 0;
#pragma unroll
        for (int k = 0; k < U; ++k) {
            // lane l takes entry `filled + l` of the window
            uint64_t at = 0;
            bool ok = false;
            uint32_t filled = 0;
            if (cs < 64 && cd - cp > 64) {   // inside one list, short of its end
                at = cb + cp + lane;
                ok = true;
                cp += 64;
                filled = 64;
            }
            while (cs < 64 && filled < 64) {
                const uint32_t take = cd - cp < 64 - filled ? cd - cp : 64 - filled;
                if (lane >= filled && lane < filled + take) {
                    at = cb + cp + (lane - filled);
                    ok = true;
                }
                filled += take;
                cp += take;
                if (cp == cd) {
                    const uint64_t rest = cs < 63 ? nz >> (cs + 1) : 0;
                    cs = rest ? cs + 1 + (uint32_t)__builtin_ctzll(rest) : 64;
                    cp = 0;
                    if (cs < 64) {
                        cd = (uint32_t)__builtin_amdgcn_readlane((int)d, (int)cs);
                        cb = readlane64(b, cs);
                    }
                }
            }
            x[k] = onb[ok ? at : 0];   // every lane loads, so the loads are not branched over
            okbits |= (ok ? 1u : 0u) << k;
        }
        return okbits;
    }
};

// bit k: x[k] (with bit k of ok set) is in the table T (2^bits slots) or,
// when T is null, in the list vl of dv entries (binary search when it is by
// id, a scan when it is in the orientation's order)
template <int U>
__device__ __forceinline__ uint32_t tri_found(const uint32_t (&x)[U], uint32_t ok, const uint32_t *T, uint32_t bits,
                                              uint32_t mult, const uint32_t *vl, uint64_t dv)
{
    uint32_t f = 0;
    if (T) {
        // two reads per entry; a lane with nothing to look up reads slot 0,
        // which all such lanes share (a broadcast, no bank conflict)
        uint32_t y1[U], y2[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const bool v = (ok >> k) & 1;
            uint32_t h1, h2;
            cuckoo_slots(x[k], bits, mult, h1, h2);
            y1[k] = T[v ? h1 : 0];
            y2[k] = T[v ? h2 : 0];
        }
#pragma unroll
        for (int k = 0; k < U; ++k) f |= ((((ok >> k) & 1) && (y1[k] == x[k] || y2[k] == x[k])) ? 1u : 0u) << k;
    } else if (!order_sorted(dv)) {
#pragma unroll
        for (int k = 0; k < U; ++k)
            if (((ok >> k) & 1) && bsearch_u32(vl, dv, x[k])) f |= 1u << k;
    } else {   // only should a table build fail
        for (uint64_t t = 0; t < dv; ++t) {
            const uint32_t y = vl[t];
#pragma unroll
            for (int k = 0; k < U; ++k)
                if (((ok >> k) & 1) && x[k] == y) f |= 1u << k;
        }
    }
    return f;
}

// Tails sb + lane (< ie) of the item: entries of their out-lists found in the
// head's table.  Software-pipelined: the next U windows' loads are in flight
// while the current ones probe the table.
template <int U>
__device__ __forceinline__ unsigned long long tri_stream(const uint64_t *ooff, const uint32_t *onb,
                                                         const uint32_t *inb, uint64_t sb,
                                                         uint64_t ie, const uint32_t *T, uint32_t bits,
                                                         uint32_t mult, const uint32_t *vl, uint64_t dv)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t si = sb + lane;
    uint64_t b = 0;
    uint32_t d = 0;
    if (si < ie) {
        const uint32_t u = inb[si];
        b = ooff[u];
        d = (uint32_t)(ooff[u + 1] - b);
    }
    TriCursor cur(b, d);
    if (!cur.more()) return 0;
    unsigned long long c = 0;
    uint32_t xa[U], xb[U];
    uint32_t oka = cur.fill<U>(onb, xa), okb;
    for (;;) {   // two buffers in turn, no copies (a copy would wait for the loads)
        bool more = cur.more();
        okb = cur.fill<U>(onb, xb);
        c += __popc(tri_found<U>(xa, oka, T, bits, mult, vl, dv));
        if (!more) break;
        more = cur.more();
        oka = cur.fill<U>(onb, xa);
        c += __popc(tri_found<U>(xb, okb, T, bits, mult, vl, dv));
        if (!more) break;
    }
    return c;
}

// ---- 4-clique rows from the same stream ----------------------------------------
// Row i of vertex u (the row of oriented edge u -> v, v = out(u)[i]) is the set
// of positions p of out(u) with out(u)[p] in out(v): the entries of out(u)
// found in v's table.  So the rows come from the triangle stream: head v's
// table, tails u from v's in-list (each entry the index e of edge u -> v in
// the out-lists, so i = e - ooff[u]), and the hits of every 64-entry word of
// out(u) stored as one row word.  Windows hold whole words of the lists (a
// word that does not fit the rest of a window opens the next one), so each
// row word is stored once, by the lane at the word's first position.
struct RowCursor {
    uint64_t nz, b, cb;   // as TriCursor
    uint32_t d, rb, cs, cp, cd, crb;   // rb: lane s's row (its first word in the row buffer)

    __device__ RowCursor(uint64_t b_, uint32_t d_, uint32_t rb_) : b(b_), d(d_), rb(rb_)
    {
        nz = __ballot(d > 0);
        cs = nz ? (uint32_t)__builtin_ctzll(nz) : 64;
        cp = 0;
        cd = cs < 64 ? (uint32_t)__builtin_amdgcn_readlane((int)d, (int)cs) : 0;
        cb = cs < 64 ? readlane64(b, cs) : 0;
        crb = cs < 64 ? (uint32_t)__builtin_amdgcn_readlane((int)rb, (int)cs) : 0;
    }
    __device__ bool more() const { return cs < 64; }

    // window k: x[k] loaded, wa[k] the row word of the lane's position; bit k
    // of ok: the lane has a position, of st: it is a word's first position
    template <int U>
    __device__ __forceinline__ void fill(const uint32_t *onb, uint32_t (&x)[U], uint32_t (&wa)[U], uint32_t &ok,
                                         uint32_t &st)
    {
        const uint32_t lane = threadIdx.x & 63;
        ok = 0;
        st = 0;
#pragma unroll
        for (int k = 0; k < U; ++k) {
            uint64_t at = 0;
            uint32_t w = 0;
            bool has = false, first = false;
            uint32_t filled = 0;
            while (cs < 64) {
                const uint32_t take = cd - cp < 64 ? cd - cp : 64;   // the rest of one word (cp is a multiple of 64)
                if (take > 64 - filled) break;
                if (lane >= filled && lane < filled + take) {
                    at = cb + cp + (lane - filled);
                    w = crb + (cp >> 6);
                    has = true;
                    first = lane == filled;
                }
                filled += take;
                cp += take;
                if (cp == cd) {
                    const uint64_t rest = cs < 63 ? nz >> (cs + 1) : 0;
                    cs = rest ? cs + 1 + (uint32_t)__builtin_ctzll(rest) : 64;
                    cp = 0;
                    if (cs < 64) {
                        cd = (uint32_t)__builtin_amdgcn_readlane((int)d, (int)cs);
                        cb = readlane64(b, cs);
                        crb = (uint32_t)__builtin_amdgcn_readlane((int)rb, (int)cs);
                    }
                }
                if (filled == 64) break;
            }
            x[k] = onb[has ? at : 0];
            wa[k] = w;
            ok |= (has ? 1u : 0u) << k;
            st |= (first ? 1u : 0u) << k;
        }
    }
};

// store each word's hits: the word's first lane takes the hit bits from its
// own lane up to the next word's first lane
template <int U>
__device__ __forceinline__ void k4_store(uint64_t *rows, const uint32_t (&wa)[U], uint32_t found, uint32_t st)
{
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < U; ++k) {
        const uint64_t hit = __ballot((found >> k) & 1), starts = __ballot((st >> k) & 1);
        if ((st >> k) & 1) {
            const uint64_t above = lane == 63 ? 0 : starts & (~0ull << (lane + 1));
            const uint32_t len = (above ? (uint32_t)__builtin_ctzll(above) : 64u) - lane;
            const uint64_t word = (hit >> lane) & (len == 64 ? ~0ull : ((1ull << len) - 1));
            if (word) rows[wa[k]] = word;
        }
    }
}

// Arguments of the head-table kernels: the triangle count (MODE 0) or the
// 4-clique rows of tails in [u0, u1) (MODE 1, in-lists holding edge indices).
struct HeadArgs {
    const uint64_t *ooff;
    const uint32_t *onb;
    const uint64_t *iin;
    const uint32_t *inb;
    const uint64_t *items;
    uint64_t nitems;
    const uint32_t *osrc;
    const uint64_t *boff;
    uint64_t bbase;
    uint32_t u0, u1;
    uint64_t *rows;
};

template <int U>
__device__ __forceinline__ void k4_stream(const HeadArgs &A, uint64_t sb, uint64_t ie, const uint32_t *T,
                                          uint32_t bits, uint32_t mult, const uint32_t *vl, uint64_t dv)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t si = sb + lane;
    uint64_t b = 0;
    uint32_t d = 0, rb = 0;
    if (si < ie) {
        const uint32_t e = A.inb[si], u = A.osrc[e];
        if (u >= A.u0 && u < A.u1) {
            const uint64_t lb = A.ooff[u], D = A.ooff[u + 1] - lb;
            if (D >= 3) {
                // a sorted out(u): from the word holding position i + 1 (the
                // words before it stay zero)
                const uint64_t s0 = order_sorted(D) ? (e - lb + 1) & ~63ull : 0;
                b = lb + s0;
                d = (uint32_t)(D - s0);
                rb = (uint32_t)(A.boff[u] - A.bbase + (e - lb) * ((D + 63) >> 6) + (s0 >> 6));
            }
        }
    }
    RowCursor cur(b, d, rb);
    if (!cur.more()) return;
    uint32_t xa[U], xb[U], wa[U], wb[U], oka, sta, okb, stb;
    cur.fill<U>(A.onb, xa, wa, oka, sta);
    for (;;) {   // two buffers in turn, as in tri_stream
        bool more = cur.more();
        cur.fill<U>(A.onb, xb, wb, okb, stb);
        k4_store<U>(A.rows, wa, tri_found<U>(xa, oka, T, bits, mult, vl, dv), sta);
        if (!more) break;
        more = cur.more();
        cur.fill<U>(A.onb, xa, wa, oka, sta);
        k4_store<U>(A.rows, wb, tri_found<U>(xb, okb, T, bits, mult, vl, dv), stb);
        if (!more) break;
    }
}

// one wavefront per work item (heads with up to kTriLight out-neighbours)
template <int MODE>
__global__ __launch_bounds__(256) void k_head_light(HeadArgs A, unsigned long long *partial)
{
    __shared__ uint32_t tab[4][1u << kTriWaveBits];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t *const T = tab[wv];
    unsigned long long c = 0;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t it = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; it < A.nitems; it += nwaves) {
        const uint64_t item = uni64(A.items[uni64(it)]);
        const uint32_t v = (uint32_t)(item >> 32);
        const uint64_t vb = A.ooff[v], dv = A.ooff[v + 1] - vb;
        const uint64_t ib = A.iin[v] + (item & 0xffffffffull) * kTriChunk, iend = A.iin[v + 1];
        const uint64_t ie = ib + kTriChunk < iend ? ib + kTriChunk : iend;
        const uint32_t bits = kTriWaveBits;
        bool built = false;
        uint32_t mult = 0;
        for (int attempt = 0; attempt < kCuckooTries && !built; ++attempt) {
            mult = cuckoo_mult(attempt);
            for (uint32_t t = lane; t < (1u << bits); t += 64) T[t] = kNone;
            wave_sync();
            built = __ballot(!tri_insert(T, bits, mult, A.onb + vb, dv, lane, 64)) == 0;
            wave_sync();
        }
        for (uint64_t sb = ib; sb < ie; sb += 64) {
            if (MODE == 0)
                c += tri_stream<kTriUnroll>(A.ooff, A.onb, A.inb, sb, ie, built ? T : nullptr, bits, mult,
                                            A.onb + vb, dv);
            else
                k4_stream<kK4Unroll>(A, sb, ie, built ? T : nullptr, bits, mult, A.onb + vb, dv);
        }
        wave_sync();   // before the table is cleared
    }
    if (MODE == 0) {
        const unsigned long long t = block_sum(c);
        if (threadIdx.x == 0) partial[blockIdx.x] = t;
    }
}

// One block of 16 wavefronts per work item of a head with more than kTriLight
// out-neighbours, sharing one table; the wavefronts claim the item's batches
// of 64 tails from an LDS counter.
template <int MODE>
__global__ __launch_bounds__(kTriHeavyThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_head_heavy(
    HeadArgs A, unsigned long long *partial)
{
    __shared__ uint32_t T[1u << kTriBlockBits];
    __shared__ uint32_t next_batch, failed;
    const uint32_t lane = threadIdx.x & 63;
    unsigned long long c = 0;
    for (uint64_t it = blockIdx.x; it < A.nitems; it += gridDim.x) {
        const uint64_t item = A.items[it];
        const uint32_t v = (uint32_t)(item >> 32);
        const uint64_t vb = A.ooff[v], dv = A.ooff[v + 1] - vb;
        const uint64_t ib = A.iin[v] + (item & 0xffffffffull) * kTriChunkHeavy, iend = A.iin[v + 1];
        const uint64_t ie = ib + kTriChunkHeavy < iend ? ib + kTriChunkHeavy : iend;
        const bool fits = 4 * dv <= (1u << kTriBlockBits);
        const uint32_t bits = kTriBlockBits;
        bool hashed = false;
        uint32_t mult = 0;
        for (int attempt = 0; fits && attempt < kCuckooTries && !hashed; ++attempt) {
            mult = cuckoo_mult(attempt);
            if (threadIdx.x == 0) failed = 0;
            for (uint32_t t = threadIdx.x; t < (1u << bits); t += blockDim.x) T[t] = kNone;
            __syncthreads();
            if (!tri_insert(T, bits, mult, A.onb + vb, dv, threadIdx.x, blockDim.x)) failed = 1;
            __syncthreads();
            hashed = !failed;
            __syncthreads();   // every thread has read the flag before the next attempt clears it
        }
        if (threadIdx.x == 0) next_batch = 0;
        __syncthreads();
        const uint32_t batches = (uint32_t)((ie - ib + 63) >> 6);
        for (;;) {
            uint32_t bi = 0;
            if (lane == 0) bi = atomicAdd(&next_batch, 1u);
            bi = __builtin_amdgcn_readfirstlane(bi);
            if (bi >= batches) break;
            if (MODE == 0)
                c += tri_stream<kTriUnrollHeavy>(A.ooff, A.onb, A.inb, ib + 64ull * bi, ie, hashed ? T : nullptr,
                                                 bits, mult, A.onb + vb, dv);
            else
                k4_stream<kK4Unroll>(A, ib + 64ull * bi, ie, hashed ? T : nullptr, bits, mult, A.onb + vb, dv);
        }
        __syncthreads();   // before the table and the counter are reset
    }
    if (MODE == 0) {
        const unsigned long long t = block_sum(c);
        if (threadIdx.x == 0) partial[blockIdx.x] = t;
    }
}

// 4-cliques by bit rows: the binary-encoded induced subgraphs of Almasri et
// al., "Parallel K-clique Counting on GPUs" (ICS 2022).  For every vertex u
// whose out-list L = out(u) has d >= 3 entries, row i (the row of oriented
// edge u -> L[i]) has bit p set iff L[p] is in out(L[i]), i.e. it is the set
// W = out(u) & out(L[i]) over positions of L.  A 4-clique whose lowest vertex
// is u, next L[i], then L[j], then L[p] is bit p of row_i & row_j, so the
// count at u is the sum over rows i and set bits j of popcount(row_i & row_j).
// Vertex u's rows take d * ceil(d / 64) words from boff[u]; out-lists stay
// within sqrt(2m), so one vertex needs at most 2^16 rows of 2^10 words.  The
// rows themselves come from the head-table stream (k4_stream).
// words of each vertex's rows (0 below three out-neighbours)
__global__ void k_k4_words(const uint64_t *ooff, uint32_t n, uint64_t *words)
{
    GRID_STRIDE(u, (uint64_t)n + 1)
    {
        const uint64_t d = u < n ? ooff[u + 1] - ooff[u] : 0;
        words[u] = d >= 3 ? d * ((d + 63) >> 6) : 0;
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
        const bool tri = order_sorted(d);   // row j zero below word (j + 1) / 64
        for (uint64_t kb = 0; kb < R; ++kb) {
            for (uint64_t wb = ri[kb]; wb; wb &= wb - 1) {
                const uint64_t j = kb * 64 + __builtin_ctzll(wb);
                const uint64_t *const rj = rb + j * R;
                for (uint64_t t = tri ? (j + 1) >> 6 : 0; t < R; ++t) c += __popcll(ri[t] & rj[t]);
            }
        }
    }
    const unsigned long long t = block_sum(c);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// One wavefront per wide row i of [e0, e1).  Lanes form groups of G lanes (G =
// R rounded up to a power of two), lane w of a group holding word w of row i;
// the set bits j of each word of row i are compacted into LDS, and each group
// takes one j per step (64 / G of them at once, kK4CountUnroll steps of loads
// in flight) and adds popcount(row_i[w] & row_j[w]).  A sorted vertex's row j
// is zero below word (j + 1) / 64, so those words are not read.  The wave
// takes 64 consecutive rows at a time and walks the wide ones among them, so a
// hub's rows spread over many waves.
constexpr int kK4CountUnroll = 4;

__global__ __launch_bounds__(256) void k_k4_count_wide(const uint64_t *ooff, const uint32_t *osrc,
                                                       const uint64_t *boff, uint64_t bbase, uint64_t e0,
                                                       uint64_t e1, const uint64_t *bits,
                                                       unsigned long long *partial)
{
    __shared__ uint32_t jls[4][64];
    uint32_t *const jl = jls[threadIdx.x >> 6];
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
            const uint64_t lb = ooff[u], D = ooff[u + 1] - lb, R = (D + 63) >> 6;
            const uint64_t *const rb = bits + (boff[u] - bbase);
            const uint64_t *const ri = rb + (e - lb) * R;
            if (R <= 64) {
                const bool tri = order_sorted(D);
                const uint32_t G = R <= 16 ? 16 : (R <= 32 ? 32 : 64), per = 64 / G;
                const uint32_t w = lane & (G - 1), grp = lane / G;
                const uint64_t mine = w < R ? ri[w] : 0;
                for (uint64_t kb = tri ? (e - lb + 1) >> 6 : 0; kb < R; ++kb) {
                    const uint64_t word = ri[kb];   // uniform
                    if (!word) continue;
                    if ((word >> lane) & 1) jl[__popcll(word & ((1ull << lane) - 1))] = (uint32_t)(kb * 64 + lane);
                    wave_sync();
                    const uint32_t cnt = (uint32_t)__popcll(word);
                    for (uint32_t q0 = 0; q0 < cnt; q0 += per * kK4CountUnroll) {
                        uint64_t x[kK4CountUnroll];
#pragma unroll
                        for (int k = 0; k < kK4CountUnroll; ++k) {
                            const uint32_t q = q0 + per * k + grp;
                            const uint32_t j = q < cnt ? jl[q] : 0;
                            const bool v = q < cnt && w < R && (!tri || w >= ((j + 1) >> 6));
                            x[k] = v ? rb[(uint64_t)j * R + w] : 0;
                        }
#pragma unroll
                        for (int k = 0; k < kK4CountUnroll; ++k) c += __popcll(mine & x[k]);
                    }
                    wave_sync();   // before jl is rewritten
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

// wedges: sum over vertices of C(d, 2), one per-block partial
__global__ void k_wedges(const uint32_t *sdeg, uint32_t n, unsigned long long *partial)
{
    unsigned long long w = 0;
    GRID_STRIDE(v, (uint64_t)n)
    {
        const unsigned long long d = sdeg[v];
        w += d * (d - (d > 0)) / 2;
    }
    const unsigned long long t = block_sum(w);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// sum over simple edges {u, v} of (d_u - 1)(d_v - 1), each edge taken at its
// lower id's entry
__global__ __launch_bounds__(256) void k_ew_paths(const uint64_t *off, const uint64_t *sadj, uint32_t n, uint64_t m2,
                                                  const uint32_t *sdeg, unsigned long long *partial)
{
    unsigned long long p = 0;
    EW_TILES(m2)
    {
        const uint64_t g0 = ew_t * kEwTile, g1 = g0 + kEwTile < ew_groups ? g0 + kEwTile : ew_groups;
        EntryWalk w(off, sadj, n, m2, g0 << 6);
        for (uint64_t g = g0; g < g1; ++g) {
            w.next();
            if (w.first && w.u < w.nb) p += (unsigned long long)(sdeg[w.u] - 1) * (sdeg[w.nb] - 1);
        }
    }
    const unsigned long long t = block_sum(p);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// 5-stars: sum over vertices of C(d, 4) (naive_implementation/FiveStar.cpp:51-59),
// in 128 bits (one hub of 150 000 neighbours already passes 2^64); per-wave
// (lo, hi) partials
__global__ __launch_bounds__(256) void k_stars4(const uint32_t *sdeg, uint32_t n, unsigned long long *part)
{
    unsigned long long lo = 0, hi = 0;
    GRID_STRIDE(v, (uint64_t)n)
    {
        const unsigned __int128 d = sdeg[v];
        if (d >= 4) {
            const unsigned __int128 c = d * (d - 1) * (d - 2) * (d - 3) / 24;
            add128(lo, hi, (unsigned long long)c, (unsigned long long)(c >> 64));
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long ol = __shfl_xor(lo, off, 64), oh = __shfl_xor(hi, off, 64);
        add128(lo, hi, ol, oh);
    }
    if ((threadIdx.x & 63) == 0) {
        const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
        part[2 * wave] = lo;
        part[2 * wave + 1] = hi;
    }
}

__device__ __forceinline__ uint64_t isect_count(const uint32_t *a, uint64_t na, const uint32_t *b, uint64_t nb)
{
    if (na > nb) { const uint32_t *t = a; a = b; b = t; const uint64_t x = na; na = nb; nb = x; }
    uint64_t c = 0;
    for (uint64_t i = 0; i < na; ++i) c += bsearch_u32(b, nb, a[i]);
    return c;
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

// ---- 4-cycles and 5-vertex chains -------------------------------------------
// The full simple adjacency with every entry as an order key
// (simple degree << 32 | id) and each list sorted by key, so that the
// neighbours ranked below a vertex form a prefix of its list.
__global__ void k_keyify(const uint32_t *fnb, const uint32_t *sdeg, uint64_t ne, uint64_t *key)
{
    GRID_STRIDE(i, ne)
    {
        const uint32_t x = fnb[i];
        key[i] = ((uint64_t)sdeg[x] << 32) | x;
    }
}

__device__ __forceinline__ uint64_t lower_u64(const uint64_t *a, uint64_t n, uint64_t x)
{
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// 4-cycles at their highest vertex u (the degree ordering of Chiba and
// Nishizeki, "Arboricity and subgraph listing algorithms", SIAM J. Comput.
// 14(1), 1985): for every neighbour v below u and every neighbour w of v
// below u, cnt[w] += 1; the cycles whose top is u and whose vertex opposite u
// is w number C(cnt[w], 2).  The work is at most the sum over edges of the
// smaller degree.  cnt lives in a per-wavefront LDS hash; a vertex that fills
// it past three quarters goes to a list counted by sorting its wedges
// (k_c4s_fill, radix sort, run-length encoding, k_c4s_pairs), or, past the
// sort batch's size, with a dense counter (k_c4d_clear, k_c4d_insert,
// k_c4d_sum).
constexpr int kC4Bits = 10;
constexpr uint32_t kC4Slots = 1u << kC4Bits;
constexpr uint32_t kC4Empty = 0xffffffffu;

__device__ __forceinline__ uint32_t c4_hash(uint32_t w, uint32_t bits) { return (w * 0x9E3779B1u) >> (32 - bits); }

// W (houses): each wedge u - v - w also adds t(u, v) + t(v, w) (tf: the
// triangles of each full-list entry's edge) to w's weight, and the top vertex
// u with opposite w adds (cnt[w] - 1) * weight[w] -- the sum over its
// C(cnt[w], 2) cycles of their four edges' triangle counts -- instead of
// C(cnt[w], 2).
template <bool W>
__global__ __launch_bounds__(256) void k_c4_lds(const uint64_t *foff, const uint64_t *fkey, const uint32_t *sdeg,
                                                uint32_t n, unsigned long long *partial, uint32_t *ovf_list,
                                                unsigned int *ovf_count, const uint32_t *tf)
{
    __shared__ uint32_t keys[4 * kC4Slots], vals[4 * kC4Slots];
    __shared__ unsigned long long wts[W ? 4 * kC4Slots : 1];
    __shared__ uint32_t fill[4], full[4];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t *const K = keys + wv * kC4Slots;
    uint32_t *const V = vals + wv * kC4Slots;
    unsigned long long *const WT = W ? wts + wv * kC4Slots : wts;
    for (uint32_t t = lane; t < kC4Slots; t += 64) {
        K[t] = kC4Empty;
        V[t] = 0;
        if (W) WT[t] = 0;
    }
    if (lane == 0) { fill[wv] = 0; full[wv] = 0; }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    unsigned long long clo = 0, chi = 0;
    WAVE_STRIDE(u, (uint64_t)n)
    {
        const uint64_t ub = foff[u], ku = ((uint64_t)sdeg[u] << 32) | u;
        const uint64_t pu = lower_u64(fkey + ub, foff[u + 1] - ub, ku);   // neighbours below u
        if (pu < 2) continue;
        for (uint64_t jb = 0; jb < pu; jb += 64) {
            if (jb + lane < pu) {
                const uint32_t v = (uint32_t)fkey[ub + jb + lane];
                const uint32_t tuv = W ? tf[ub + jb + lane] : 0;
                const uint64_t vb = foff[v];
                const uint64_t qv = lower_u64(fkey + vb, foff[v + 1] - vb, ku);   // neighbours of v below u
                for (uint64_t k = 0; k < qv; ++k) {
                    if (*(volatile uint32_t *)&full[wv]) break;
                    const uint32_t w = (uint32_t)fkey[vb + k];
                    uint32_t h = c4_hash(w, kC4Bits);
                    bool done = false;
                    for (uint32_t probe = 0; probe < kC4Slots && !done; ++probe) {
                        const uint32_t old = atomicCAS(&K[h], kC4Empty, w);
                        if (old == kC4Empty || old == w) {
                            atomicAdd(&V[h], 1u);
                            if (W) atomicAdd(&WT[h], (unsigned long long)tuv + tf[vb + k]);
                            if (old == kC4Empty && atomicAdd(&fill[wv], 1u) + 1 >= kC4Slots * 3 / 4)
                                atomicOr(&full[wv], 1u);
                            done = true;
                        } else {
                            h = (h + 1) & (kC4Slots - 1);
                        }
                    }
                    if (!done) atomicOr(&full[wv], 1u);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            if (full[wv]) break;
        }
        if (!full[wv]) {
            for (uint32_t t = lane; t < kC4Slots; t += 64) {
                const unsigned long long x = V[t];
                if (W) {
                    const unsigned __int128 q = (unsigned __int128)(x - (x > 0)) * WT[t];
                    add128(clo, chi, (unsigned long long)q, (unsigned long long)(q >> 64));
                } else {
                    add128(clo, chi, x * (x - (x > 0)) / 2, 0);
                }
            }
        } else if (lane == 0) {
            ovf_list[atomicAdd(ovf_count, 1u)] = (uint32_t)u;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (uint32_t t = lane; t < kC4Slots; t += 64) {
            K[t] = kC4Empty;
            V[t] = 0;
            if (W) WT[t] = 0;
        }
        if (lane == 0) { fill[wv] = 0; full[wv] = 0; }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    block_sum128_store(clo, chi, partial);
}

// For each listed vertex: the number of its neighbours below it, and its
// wedges u - v - w with v and w below u (each v's share found by the same
// search the counting kernels make).
__global__ void k_c4_wedges(const uint64_t *foff, const uint64_t *fkey, const uint32_t *sdeg, const uint32_t *list,
                            uint32_t nlist, uint64_t *below, uint64_t *wedges)
{
    const uint32_t lane = threadIdx.x & 63;
    WAVE_STRIDE(i, (uint64_t)nlist)
    {
        const uint32_t u = list[i];
        const uint64_t ub = foff[u], ku = ((uint64_t)sdeg[u] << 32) | u;
        const uint64_t pu = lower_u64(fkey + ub, foff[u + 1] - ub, ku);
        unsigned long long w = 0;
        for (uint64_t j = lane; j < pu; j += 64) {
            const uint32_t v = (uint32_t)fkey[ub + j];
            const uint64_t vb = foff[v];
            w += lower_u64(fkey + vb, foff[v + 1] - vb, ku);
        }
        for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
        if (lane == 0) {
            below[i] = pu;
            wedges[i] = w;
        }
    }
}

// Sorted counting of a batch of listed vertices: one wavefront per (vertex
// u = list[i], neighbour v below u) writes the keys (i << wbits | w) of v's
// neighbours w below u at places reserved from u's cursor; after a radix sort
// and a run-length encoding, every run of equal keys is one cnt[w] and adds
// C(run, 2).
template <class Key>
__global__ __launch_bounds__(256) void k_c4s_fill(const uint64_t *foff, const uint64_t *fkey, const uint32_t *sdeg,
                                                  const uint32_t *list, const uint64_t *ipre, uint32_t nb,
                                                  uint32_t wbits, unsigned long long *cursor, Key *keys)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t items = ipre[nb];
    for (uint64_t t = wave; t < items; t += nwaves) {
        uint32_t lo = 0, hi = nb;   // ipre[lo] <= t < ipre[hi]
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (ipre[mid] <= t) lo = mid; else hi = mid;
        }
        const uint32_t u = list[lo];
        const uint64_t ub = foff[u], ku = ((uint64_t)sdeg[u] << 32) | u;
        const uint32_t v = (uint32_t)fkey[ub + (t - ipre[lo])];
        const uint64_t vb = foff[v];
        const uint64_t qv = lower_u64(fkey + vb, foff[v + 1] - vb, ku);
        unsigned long long base = 0;
        if (lane == 0 && qv) base = atomicAdd(&cursor[lo], (unsigned long long)qv);
        base = __shfl(base, 0, 64);
        for (uint64_t k = lane; k < qv; k += 64) keys[base + k] = ((Key)lo << wbits) | (Key)(uint32_t)fkey[vb + k];
    }
}

// Weighted sorted counting (5-houses): every wedge u - v - w of a batch's
// vertex u writes the key (i << wbits | w) with the value {1, t(u, v) + t(v, w)};
// after sorting the pairs and reducing them by key, a run {c, s} adds
// (c - 1) * s.
struct CW {
    unsigned long long c, s;
};
struct CWPlus {
    __host__ __device__ CW operator()(const CW &a, const CW &b) const { return CW{a.c + b.c, a.s + b.s}; }
};
// A packed value holds the count above bit sb and the weight below it; sums of
// packed values stay exact while every run's weight stays below 2^sb and its
// count below 2^(64 - sb), which the host checks from the maximum degree (a
// run counts at most max_degree wedges, each weighing at most 2 max_degree).
__device__ __forceinline__ void cw_make(CW &v, unsigned long long w, uint32_t) { v = CW{1ull, w}; }
__device__ __forceinline__ void cw_make(unsigned long long &v, unsigned long long w, uint32_t sb)
{
    v = (1ull << sb) + w;
}
__device__ __forceinline__ CW cw_get(const CW &v, uint32_t) { return v; }
__device__ __forceinline__ CW cw_get(unsigned long long v, uint32_t sb)
{
    return CW{v >> sb, v & ((1ull << sb) - 1)};
}

template <class Key, class Val>
__global__ __launch_bounds__(256) void k_c4s_fill_w(const uint64_t *foff, const uint64_t *fkey, const uint32_t *sdeg,
                                                    const uint32_t *list, const uint64_t *ipre, uint32_t nb,
                                                    uint32_t wbits, unsigned long long *cursor, const uint32_t *tf,
                                                    uint32_t sb, Key *keys, Val *vals)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t items = ipre[nb];
    for (uint64_t t = wave; t < items; t += nwaves) {
        uint32_t lo = 0, hi = nb;   // ipre[lo] <= t < ipre[hi]
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (ipre[mid] <= t) lo = mid; else hi = mid;
        }
        const uint32_t u = list[lo];
        const uint64_t ub = foff[u], ku = ((uint64_t)sdeg[u] << 32) | u;
        const uint64_t j = ub + (t - ipre[lo]);
        const uint32_t v = (uint32_t)fkey[j];
        const unsigned long long tuv = tf[j];
        const uint64_t vb = foff[v];
        const uint64_t qv = lower_u64(fkey + vb, foff[v + 1] - vb, ku);
        unsigned long long base = 0;
        if (lane == 0 && qv) base = atomicAdd(&cursor[lo], (unsigned long long)qv);
        base = __shfl(base, 0, 64);
        for (uint64_t k = lane; k < qv; k += 64) {
            keys[base + k] = ((Key)lo << wbits) | (Key)(uint32_t)fkey[vb + k];
            cw_make(vals[base + k], tuv + tf[vb + k], sb);
        }
    }
}

template <class Val>
__global__ void k_c4s_pairs_w(const Val *runs, const unsigned int *nruns, uint32_t sb, unsigned long long *partial)
{
    const uint64_t n = *nruns;
    unsigned long long lo = 0, hi = 0;
    GRID_STRIDE(t, n)
    {
        const CW r = cw_get(runs[t], sb);
        const unsigned __int128 q = (unsigned __int128)(r.c - (r.c > 0)) * r.s;
        add128(lo, hi, (unsigned long long)q, (unsigned long long)(q >> 64));
    }
    block_sum128_store(lo, hi, partial);
}

__global__ void k_c4s_pairs(const unsigned int *counts, const unsigned int *nruns, unsigned long long *partial)
{
    const uint64_t runs = *nruns;
    unsigned long long lo = 0, hi = 0;
    GRID_STRIDE(t, runs)
    {
        const unsigned long long x = counts[t];
        add128(lo, hi, x * (x - 1) / 2, 0);
    }
    block_sum128_store(lo, hi, partial);
}

__global__ void k_c4d_clear(uint32_t *cnt, uint64_t total)
{
    GRID_STRIDE(t, total) cnt[t] = 0;
}

// The listed vertices (hubs) of one batch, each with a dense counter over all
// n vertices (cnt + i * n): one wavefront per (vertex u, neighbour v below u),
// the lanes adding 1 at every neighbour of v below u (atomics whose results
// are not read).  ipre is the exclusive prefix of the vertices' neighbour
// counts below them.
template <bool W>
__global__ __launch_bounds__(256) void k_c4d_insert(const uint64_t *foff, const uint64_t *fkey, const uint32_t *sdeg,
                                                    const uint32_t *list, const uint64_t *ipre, uint32_t nb,
                                                    uint32_t n, uint32_t *cnt, const uint32_t *tf,
                                                    unsigned long long *wsum)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t items = ipre[nb];
    for (uint64_t t = wave; t < items; t += nwaves) {
        uint32_t lo = 0, hi = nb;   // ipre[lo] <= t < ipre[hi]
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (ipre[mid] <= t) lo = mid; else hi = mid;
        }
        const uint32_t u = list[lo];
        const uint64_t ub = foff[u], ku = ((uint64_t)sdeg[u] << 32) | u;
        const uint32_t v = (uint32_t)fkey[ub + (t - ipre[lo])];
        const uint32_t tuv = W ? tf[ub + (t - ipre[lo])] : 0;
        const uint64_t vb = foff[v];
        const uint64_t qv = lower_u64(fkey + vb, foff[v + 1] - vb, ku);
        uint32_t *const cu = cnt + (uint64_t)lo * n;
        unsigned long long *const wu = W ? wsum + (uint64_t)lo * n : wsum;
        for (uint64_t k = lane; k < qv; k += 64) {
            const uint32_t w = (uint32_t)fkey[vb + k];
            atomicAdd(&cu[w], 1u);
            if (W) atomicAdd(&wu[w], (unsigned long long)tuv + tf[vb + k]);
        }
    }
}

template <bool W>
__global__ void k_c4d_sum(const uint32_t *cnt, uint64_t total, const unsigned long long *wsum,
                          unsigned long long *partial)
{
    unsigned long long lo = 0, hi = 0;
    GRID_STRIDE(t, total)
    {
        const unsigned long long x = cnt[t];
        if (W) {
            const unsigned __int128 q = (unsigned __int128)(x - (x > 0)) * wsum[t];
            add128(lo, hi, (unsigned long long)q, (unsigned long long)(q >> 64));
        } else {
            add128(lo, hi, x * (x - (x > 0)) / 2, 0);
        }
    }
    block_sum128_store(lo, hi, partial);
}

__global__ void k_clear_u64(unsigned long long *p, uint64_t total)
{
    GRID_STRIDE(t, total) p[t] = 0;
}

// 5-houses: triangles per edge.  One wavefront per oriented edge u -> v finds
// the common out-neighbours x of u and v (each triangle once, at the edge
// between its two lowest vertices) and adds 1 to the counts of its three
// edges u -> v, u -> x and v -> x (te: one count per out-list entry).
__global__ __launch_bounds__(256) void k_tri_edges(const uint64_t *ooff, const uint32_t *onb, const uint32_t *osrc,
                                                   uint64_t ne, uint32_t *te)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t e = wave; e < ne; e += nwaves) {
        const uint32_t u = osrc[e], v = onb[e];
        const uint64_t ub = ooff[u], du = ooff[u + 1] - ub, vb = ooff[v], dv = ooff[v + 1] - vb;
        // search the entries of the list that costs fewer passes times depth in the other
        const bool from_u = ((du + 63) >> 6) * clog2_u64(dv + 1) <= ((dv + 63) >> 6) * clog2_u64(du + 1);
        const uint64_t abase = from_u ? ub : vb, na = from_u ? du : dv;
        const uint64_t bbase = from_u ? vb : ub, nb = from_u ? dv : du;
        unsigned long long c = 0;
        for (uint64_t i = lane; i < na; i += 64) {
            const uint32_t x = onb[abase + i];
            uint64_t lo = 0, hi = nb;
            while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                if (onb[bbase + mid] < x) lo = mid + 1; else hi = mid;
            }
            if (lo < nb && onb[bbase + lo] == x) {
                atomicAdd(&te[abase + i], 1u);
                atomicAdd(&te[bbase + lo], 1u);
                ++c;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0 && c) atomicAdd(&te[e], (uint32_t)c);
    }
}

// The per-edge triangle counts onto the full sorted lists: edge u -> v's
// count at v's entry in u's list and at u's entry in v's list (found by key);
// with the sum over edges of t (t - 1) in 128 bits.
__global__ void k_tf_scatter(const uint32_t *onb, const uint32_t *osrc, uint64_t ne, const uint32_t *te,
                             const uint64_t *foff, const uint64_t *fkey, const uint32_t *sdeg, uint32_t *tf,
                             unsigned long long *partial)
{
    unsigned long long lo = 0, hi = 0;
    GRID_STRIDE(e, ne)
    {
        const uint32_t u = osrc[e], v = onb[e], t = te[e];
        const uint64_t kv = ((uint64_t)sdeg[v] << 32) | v, ku = ((uint64_t)sdeg[u] << 32) | u;
        tf[foff[u] + lower_u64(fkey + foff[u], foff[u + 1] - foff[u], kv)] = t;
        tf[foff[v] + lower_u64(fkey + foff[v], foff[v + 1] - foff[v], ku)] = t;
        add128(lo, hi, (unsigned long long)t * (t ? t - 1 : 0), 0);
    }
    block_sum128_store(lo, hi, partial);
}

// sum over vertices v of S1^2 - S2, S1 = sum of (d_x - 1) and S2 = sum of
// (d_x - 1)^2 over the neighbours x of v (degrees read from the keys): the
// ordered pairs of distinct neighbours of v weighted by (d_x - 1)(d_y - 1).
// One wavefront per vertex; per-wave 128-bit partials as (lo, hi) pairs.
__global__ __launch_bounds__(256) void k_c5_pairs(const uint64_t *foff, const uint64_t *fkey, uint32_t n,
                                                  unsigned long long *partial)
{
    const uint32_t lane = threadIdx.x & 63;
    unsigned long long lo = 0, hi = 0;
    WAVE_STRIDE(v, (uint64_t)n)
    {
        const uint64_t vb = foff[v], dv = foff[v + 1] - vb;
        unsigned long long s1 = 0, s2lo = 0, s2hi = 0;
        for (uint64_t j = lane; j < dv; j += 64) {
            const unsigned long long a = (fkey[vb + j] >> 32) - 1;
            s1 += a;
            const unsigned __int128 q = (unsigned __int128)a * a;
            add128(s2lo, s2hi, (unsigned long long)q, (unsigned long long)(q >> 64));
        }
        for (int off = 32; off > 0; off >>= 1) {
            s1 += __shfl_xor(s1, off, 64);
            const unsigned long long ol = __shfl_xor(s2lo, off, 64), oh = __shfl_xor(s2hi, off, 64);
            add128(s2lo, s2hi, ol, oh);
        }
        if (lane == 0) {
            const unsigned __int128 w = (unsigned __int128)s1 * s1 - (((unsigned __int128)s2hi << 64) | s2lo);
            add128(lo, hi, (unsigned long long)w, (unsigned long long)(w >> 64));
        }
    }
    if (lane == 0) {
        const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
        partial[2 * wave] = lo;
        partial[2 * wave + 1] = hi;
    }
}

// sum of n (lo, hi) pairs written by block_sum128_store
int host_sum128(const unsigned long long *d, int n, hipStream_t st, unsigned __int128 *out)
{
    std::vector<unsigned long long> h(2 * (size_t)n);
    ZG_HIP(hipMemcpyAsync(h.data(), d, 16ull * n, hipMemcpyDeviceToHost, st));
    ZG_HIP(hipStreamSynchronize(st));
    unsigned __int128 s = 0;
    for (int i = 0; i < n; ++i) s += ((unsigned __int128)h[2 * i + 1] << 64) | h[2 * i];
    *out = s;
    return ZG_OK;
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

inline int ew_grid(uint64_t groups)
{
    const uint64_t waves = (groups + kEwTile - 1) / kEwTile, cap = (uint64_t)device_cus() * 64;
    return (int)std::max<uint64_t>(1, std::min((waves + 3) / 4, cap));
}

// Flags over the 2m list entries, one bit each (mask[groups] = 0), and their
// exclusive prefix per group (pre[groups] = the flagged total).
struct EwMask {
    XB<uint64_t> mask, pre;
    uint64_t total = 0;
};

int ew_alloc(zg_graph *g, EwMask &e)
{
    const uint64_t groups = (2 * (uint64_t)g->m + 63) >> 6;
    XALLOC(e.mask, groups + 1);
    XALLOC(e.pre, groups + 1);
    ZG_HIP(hipMemsetAsync(e.mask.p + groups, 0, 8, g->stream));
    return ZG_OK;
}

int ew_scan(zg_graph *g, EwMask &e)
{
    const uint64_t groups = (2 * (uint64_t)g->m + 63) >> 6;
    hipStream_t st = g->stream;
    const auto it = rocprim::make_transform_iterator(e.mask.p, Popc());
    size_t tb = 0;
    ZG_HIP(rocprim::exclusive_scan(nullptr, tb, it, e.pre.p, (uint64_t)0, (size_t)groups + 1,
                                   rocprim::plus<uint64_t>(), st));
    XB<char> tmp;
    XALLOC(tmp, tb);
    ZG_HIP(rocprim::exclusive_scan(tmp.p, tb, it, e.pre.p, (uint64_t)0, (size_t)groups + 1,
                                   rocprim::plus<uint64_t>(), st));
    ZG_HIP(hipMemcpyAsync(&e.total, e.pre.p + groups, 8, hipMemcpyDeviceToHost, st));
    ZG_HIP(hipStreamSynchronize(st));
    return ZG_OK;
}

// first-of-run flags (the full simple lists)
int ew_first(zg_graph *g, EwMask &e)
{
    const uint64_t m2 = 2 * (uint64_t)g->m, groups = (m2 + 63) >> 6;
    int rc = ew_alloc(g, e);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ew_first, dim3(ew_grid(groups)), dim3(256), 0, g->stream, g->off, g->sadj, g->n, m2,
                       e.mask.p);
    ZG_HIP(hipGetLastError());
    return ew_scan(g, e);
}

// out-list flags and, when in is given, in-list flags, from one pass
int ew_orient(zg_graph *g, const uint16_t *okey, EwMask &out, EwMask *in)
{
    const uint64_t m2 = 2 * (uint64_t)g->m, groups = (m2 + 63) >> 6;
    int rc = ew_alloc(g, out);
    if (!rc && in) rc = ew_alloc(g, *in);
    if (rc) return rc;
    if (in)
        hipLaunchKernelGGL(k_ew_orient<true>, dim3(ew_grid(groups)), dim3(256), 0, g->stream, g->off, g->sadj, g->n,
                           m2, okey, out.mask.p, in->mask.p);
    else
        hipLaunchKernelGGL(k_ew_orient<false>, dim3(ew_grid(groups)), dim3(256), 0, g->stream, g->off, g->sadj,
                           g->n, m2, okey, out.mask.p, nullptr);
    ZG_HIP(hipGetLastError());
    rc = ew_scan(g, out);
    if (!rc && in) rc = ew_scan(g, *in);
    return rc;
}

// The lists of the flagged entries, neighbours ascending by id: offsets (n + 1
// entries), neighbours and, when own is given, owners (pos_of_rank and
// edge_of_rank as for k_ew_compact).
int ew_lists(zg_graph *g, const EwMask &e, XB<uint64_t> &offs, XB<uint32_t> &nbr, XB<uint32_t> *own,
             uint32_t *pos_of_rank = nullptr, const uint32_t *edge_of_rank = nullptr, uint32_t *rank_out = nullptr)
{
    const uint64_t m2 = 2 * (uint64_t)g->m, groups = (m2 + 63) >> 6;
    hipStream_t st = g->stream;
    XALLOC(offs, (uint64_t)g->n + 1);
    XALLOC(nbr, e.total);
    if (own) XALLOC(*own, e.total);
    hipLaunchKernelGGL(k_ew_offsets, dim3(gridx((uint64_t)g->n + 1)), dim3(256), 0, st, g->off, g->n, e.mask.p,
                       e.pre.p, offs.p);
    hipLaunchKernelGGL(k_ew_compact, dim3(ew_grid(groups)), dim3(256), 0, st, g->off, g->sadj, g->n, m2, e.mask.p,
                       e.pre.p, nbr.p, own ? own->p : nullptr, pos_of_rank, edge_of_rank, rank_out);
    ZG_HIP(hipGetLastError());
    ZG_HIP(hipStreamSynchronize(st));   // before a mask is freed
    return ZG_OK;
}

}  // namespace
}  // namespace zg

using namespace zg;

extern "C" int zg_exact_counts(zg_graph *g, int what, uint64_t *out)
{
    if (!g || !out) return set_error(ZG_EINVAL, "null argument");
    if (what < 0 || what > 7)
        return set_error(ZG_EINVAL, "what must be 0 (triangles), 1 (4-cliques), 2 (wedges), 3 (3-edge paths), "
                                    "4 (5-stars), 5 (5-houses), 6 (5-vertex chains), 7 (4-cycles)");
    const uint32_t n = g->n;
    hipStream_t st = g->stream;
    if (g->m == 0) { *out = 0; return ZG_OK; }
    auto read_u64 = [&](const void *p, uint64_t *v) -> int {
        ZG_HIP(hipMemcpyAsync(v, p, 8, hipMemcpyDeviceToHost, st));
        ZG_HIP(hipStreamSynchronize(st));
        return ZG_OK;
    };
    // Simple degrees: the first-of-run entries of each list, which the full
    // simple lists (5-houses, 5-chains, 4-cycles) hold; without parallel edges
    // the list lengths themselves.
    XB<uint32_t> sdeg;
    XB<uint64_t> foff;          // full simple lists: offsets, neighbours, owners
    XB<uint32_t> fnb, fsrc;
    uint64_t nf = 2 * (uint64_t)g->m;
    XALLOC(sdeg, n);
    if (what >= 5) {
        EwMask fm;
        int rc = ew_first(g, fm);
        if (!rc) rc = ew_lists(g, fm, foff, fnb, &fsrc);
        if (rc) return rc;
        nf = fm.total;
    } else if (!g->simple) {
        EwMask fm;
        const int rc = ew_first(g, fm);
        if (rc) return rc;
        nf = fm.total;
        XALLOC(foff, (uint64_t)n + 1);
        hipLaunchKernelGGL(k_ew_offsets, dim3(gridx((uint64_t)n + 1)), dim3(256), 0, st, g->off, n, fm.mask.p,
                           fm.pre.p, foff.p);
        ZG_HIP(hipStreamSynchronize(st));   // before the mask is freed
    }
    hipLaunchKernelGGL(k_deg_from_offsets, dim3(gridx(n)), dim3(256), 0, st, foff.p ? foff.p : g->off, n, sdeg.p);
    ZG_HIP(hipGetLastError());
    if (what == 4) {
        const int grid = gridx(n);
        XB<unsigned long long> part;
        XALLOC(part, (uint64_t)grid * 8);   // (lo, hi) per wave, 4 waves per block
        hipLaunchKernelGGL(k_stars4, dim3(grid), dim3(256), 0, st, sdeg.p, n, part.p);
        ZG_HIP(hipGetLastError());
        std::vector<unsigned long long> hp((size_t)grid * 8);
        ZG_HIP(hipMemcpyAsync(hp.data(), part.p, 8ull * hp.size(), hipMemcpyDeviceToHost, st));
        ZG_HIP(hipStreamSynchronize(st));
        unsigned __int128 total = 0;
        for (size_t i = 0; i < hp.size(); i += 2) total += ((unsigned __int128)hp[i + 1] << 64) | hp[i];
        if (total >> 64) return set_error(ZG_EOVERFLOW, "5-star count exceeds 64 bits");
        *out = (uint64_t)total;
        return ZG_OK;
    }
    if (what == 2 || what == 3) {
        const int grid = gridx(n), pgrid = ew_grid((2 * (uint64_t)g->m + 63) >> 6);
        XB<unsigned long long> pw, pp;
        XALLOC(pw, grid); XALLOC(pp, pgrid);
        hipLaunchKernelGGL(k_wedges, dim3(grid), dim3(256), 0, st, sdeg.p, n, pw.p);
        hipLaunchKernelGGL(k_ew_paths, dim3(pgrid), dim3(256), 0, st, g->off, g->sadj, n, 2 * (uint64_t)g->m, sdeg.p,
                           pp.p);
        ZG_HIP(hipGetLastError());
        unsigned long long w = 0, p = 0;
        int rc = host_sum(pw.p, grid, st, &w);
        if (!rc) rc = host_sum(pp.p, pgrid, st, &p);
        if (rc) return rc;
        if (what == 2) { *out = w; return ZG_OK; }
        uint64_t T = 0;
        rc = zg_exact_counts(g, 0, &T);
        if (rc) return rc;
        *out = p - 3 * T;
        return ZG_OK;
    }
    // oriented out-lists (owners for the 4-clique rows and the weighted
    // triangles), and for triangles and 4-cliques the in-lists: for every head
    // v the tails u -> v (for 4-cliques the index of edge u -> v in the
    // out-lists instead of u)
    XB<uint64_t> ooff, iin;
    XB<uint32_t> onb, osrc, inb;
    uint64_t ne = 0, maxout = 0;
    if (what != 7) {
        XB<uint16_t> okey;
        XALLOC(okey, n);
        hipLaunchKernelGGL(k_order_key, dim3(gridx(n)), dim3(256), 0, st, sdeg.p, n, okey.p);
        EwMask omask, imask;
        int rc = ew_orient(g, okey.p, omask, what <= 1 ? &imask : nullptr);
        if (rc) return rc;
        ne = omask.total;
        if (what == 1 && ne >= (1ull << 32)) return set_error(ZG_EINVAL, "4-clique counts need m < 2^32");
        XB<uint32_t> pos_of_rank, orank;   // 4-cliques: out-list index and edge rank of each out-list entry
        if (what == 1) {
            XALLOC(pos_of_rank, g->m);
            XALLOC(orank, ne);
        }
        rc = ew_lists(g, omask, ooff, onb, what == 0 ? nullptr : &osrc, pos_of_rank.p, nullptr, orank.p);
        if (rc) return rc;
        omask.mask.reset();
        omask.pre.reset();
        if (what == 0) {
            rc = ew_lists(g, imask, iin, inb, nullptr);
            if (rc) return rc;
        }
        if (what == 1) {   // out-lists of kSortMin < d <= kSortMax entries into the orientation's order
            XB<uint32_t> slist;
            XB<unsigned int> scount;
            XALLOC(slist, n);
            XALLOC(scount, 1);
            ZG_HIP(hipMemsetAsync(scount.p, 0, sizeof(unsigned int), st));
            hipLaunchKernelGGL(k_collect_sorted, dim3(gridx(n)), dim3(256), 0, st, ooff.p, n, slist.p, scount.p);
            unsigned int ns = 0;
            ZG_HIP(hipMemcpyAsync(&ns, scount.p, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
            ZG_HIP(hipStreamSynchronize(st));
            if (ns)
                hipLaunchKernelGGL(k_sort_lists, dim3((int)std::min<uint64_t>(ns, (uint64_t)device_cus() * 8)),
                                   dim3(256), 0, st, ooff.p, slist.p, ns, okey.p, onb.p, orank.p, pos_of_rank.p);
            ZG_HIP(hipGetLastError());
            ZG_HIP(hipStreamSynchronize(st));   // before the temporaries are freed
            orank.reset();
            rc = ew_lists(g, imask, iin, inb, nullptr, nullptr, pos_of_rank.p);
            if (rc) return rc;
        }
        okey.reset();
        XB<unsigned long long> mx;
        XALLOC(mx, 1);
        ZG_HIP(hipMemsetAsync(mx.p, 0, 8, st));
        hipLaunchKernelGGL(k_max_gap, dim3(gridx(n)), dim3(256), 0, st, ooff.p, n, mx.p);
        ZG_HIP(hipGetLastError());
        rc = read_u64(mx.p, &maxout);
        if (rc) return rc;
    }
    // work items of the head-table kernels: up to kTriChunk (light) or
    // kTriChunkHeavy (heavy) in-list entries of one head with out-neighbours
    XB<uint64_t> items[2];
    uint64_t nitems[2] = {0, 0};
    if (what <= 1) {
        XB<uint64_t> icnt, ipre;
        XALLOC(icnt, (uint64_t)n + 1);
        XALLOC(ipre, (uint64_t)n + 1);
        for (int heavy = 0; heavy < 2; ++heavy) {
            if (heavy)
                hipLaunchKernelGGL(k_tri_items_count<true>, dim3(gridx((uint64_t)n + 1)), dim3(256), 0, st, ooff.p,
                                   iin.p, n, icnt.p);
            else
                hipLaunchKernelGGL(k_tri_items_count<false>, dim3(gridx((uint64_t)n + 1)), dim3(256), 0, st,
                                   ooff.p, iin.p, n, icnt.p);
            size_t tb = 0;
            ZG_HIP(rocprim::exclusive_scan(nullptr, tb, icnt.p, ipre.p, (uint64_t)0, (size_t)n + 1,
                                           rocprim::plus<uint64_t>(), st));
            XB<char> tmp;
            XALLOC(tmp, tb);
            ZG_HIP(rocprim::exclusive_scan(tmp.p, tb, icnt.p, ipre.p, (uint64_t)0, (size_t)n + 1,
                                           rocprim::plus<uint64_t>(), st));
            int rc = read_u64(ipre.p + n, &nitems[heavy]);
            if (rc) return rc;
            XALLOC(items[heavy], nitems[heavy]);
            if (nitems[heavy])
                hipLaunchKernelGGL(k_tri_items_fill, dim3(gridx(n)), dim3(256), 0, st, ipre.p, n, items[heavy].p);
            ZG_HIP(hipGetLastError());
            ZG_HIP(hipStreamSynchronize(st));   // before the temporaries are freed
        }
    }
    const int hgrid[2] = {
        (int)std::min<uint64_t>((uint64_t)device_cus() * 8, std::max<uint64_t>(1, (nitems[0] + 3) / 4)),
        (int)std::min<uint64_t>((uint64_t)device_cus() * 2, std::max<uint64_t>(1, nitems[1]))};
    auto head_args = [&](int heavy) {
        HeadArgs A{};
        A.ooff = ooff.p;
        A.onb = onb.p;
        A.iin = iin.p;
        A.inb = inb.p;
        A.items = items[heavy].p;
        A.nitems = nitems[heavy];
        A.osrc = osrc.p;
        return A;
    };
    if (what == 0) {
        unsigned long long total = 0;
        for (int heavy = 0; heavy < 2; ++heavy) {
            if (nitems[heavy] == 0) continue;
            XB<unsigned long long> tpart;
            XALLOC(tpart, hgrid[heavy]);
            if (heavy)
                hipLaunchKernelGGL(k_head_heavy<0>, dim3(hgrid[1]), dim3(kTriHeavyThreads), 0, st, head_args(1),
                                   tpart.p);
            else
                hipLaunchKernelGGL(k_head_light<0>, dim3(hgrid[0]), dim3(256), 0, st, head_args(0), tpart.p);
            ZG_HIP(hipGetLastError());
            unsigned long long s = 0;
            const int rc = host_sum(tpart.p, hgrid[heavy], st, &s);
            if (rc) return rc;
            total += s;
        }
        *out = total;
        return ZG_OK;
    }
    if (what >= 5) {
        // full simple adjacency as order keys, every list sorted by key
        if (nf >= (1ull << 32)) return set_error(ZG_EINVAL, "house, 4-cycle and 5-chain counts need 2m < 2^32");
        XB<uint64_t> fkey;
        XALLOC(fkey, nf);
        {
            XB<uint64_t> kin;
            XALLOC(kin, nf);
            hipLaunchKernelGGL(k_keyify, dim3(gridx(nf)), dim3(256), 0, st, fnb.p, sdeg.p, nf, kin.p);
            ZG_HIP(hipGetLastError());
            size_t ts = 0;
            ZG_HIP(rocprim::segmented_radix_sort_keys(nullptr, ts, kin.p, fkey.p, (unsigned int)nf, n, foff.p,
                                                      foff.p + 1, 0, 64, st));
            XB<char> tmps;
            XALLOC(tmps, ts);
            ZG_HIP(rocprim::segmented_radix_sort_keys(tmps.p, ts, kin.p, fkey.p, (unsigned int)nf, n, foff.p,
                                                      foff.p + 1, 0, 64, st));
            ZG_HIP(hipStreamSynchronize(st));   // before the temporaries are freed
        }
        fnb.reset();
        fsrc.reset();
        // 5-houses = sum_e t_e c4_e - 2 sum_e t_e (t_e - 1), t_e and c4_e the
        // triangles and 4-cycles through edge e (a house is an edge in both a
        // triangle and a 4-cycle whose apex is off the cycle; over its two
        // automorphisms).  sum_e t_e c4_e is the 4-cycle count below with each
        // cycle weighted by its four edges' triangle counts (tf).
        const bool weighted = what == 5;
        XB<uint32_t> tf;
        unsigned __int128 tt = 0;   // sum over edges of t_e (t_e - 1)
        if (weighted) {
            XB<uint32_t> te;
            XALLOC(te, ne ? ne : 1);
            XALLOC(tf, nf);
            hipLaunchKernelGGL(k_c4d_clear, dim3(gridx(ne)), dim3(256), 0, st, te.p, ne);
            const int tgrid = (int)std::min<uint64_t>((uint64_t)device_cus() * 16, std::max<uint64_t>(1, (ne + 3) / 4));
            hipLaunchKernelGGL(k_tri_edges, dim3(tgrid), dim3(256), 0, st, ooff.p, onb.p, osrc.p, ne, te.p);
            ZG_HIP(hipGetLastError());
            const int sgrid = gridx(ne);
            XB<unsigned long long> tpart;
            XALLOC(tpart, 2 * (uint64_t)sgrid);
            hipLaunchKernelGGL(k_tf_scatter, dim3(sgrid), dim3(256), 0, st, onb.p, osrc.p, ne, te.p, foff.p, fkey.p,
                               sdeg.p, tf.p, tpart.p);
            ZG_HIP(hipGetLastError());
            const int rc = host_sum128(tpart.p, sgrid, st, &tt);   // synchronizes before te is freed
            if (rc) return rc;
        }
        // 4-cycles: LDS hashes, then the vertices that overflowed them with global hashes
        const int cgrid = (int)std::min<uint64_t>((uint64_t)device_cus() * 8, std::max<uint64_t>(1, ((uint64_t)n + 3) / 4));
        XB<unsigned long long> cpart;
        XB<uint32_t> ovf;
        XB<unsigned int> novf_d;
        XALLOC(cpart, 2 * (uint64_t)cgrid); XALLOC(ovf, n); XALLOC(novf_d, 1);
        ZG_HIP(hipMemsetAsync(novf_d.p, 0, sizeof(unsigned int), st));
        if (weighted)
            hipLaunchKernelGGL(k_c4_lds<true>, dim3(cgrid), dim3(256), 0, st, foff.p, fkey.p, sdeg.p, n, cpart.p,
                               ovf.p, novf_d.p, tf.p);
        else
            hipLaunchKernelGGL(k_c4_lds<false>, dim3(cgrid), dim3(256), 0, st, foff.p, fkey.p, sdeg.p, n, cpart.p,
                               ovf.p, novf_d.p, nullptr);
        ZG_HIP(hipGetLastError());
        unsigned __int128 c4 = 0;   // 4-cycles (houses: their weights), summed in 128 bits
        int rc = host_sum128(cpart.p, cgrid, st, &c4);
        if (rc) return rc;
        unsigned int novf = 0;
        ZG_HIP(hipMemcpyAsync(&novf, novf_d.p, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        ZG_HIP(hipStreamSynchronize(st));
        if (novf) {
            XB<uint64_t> below, wedges;
            XALLOC(below, novf); XALLOC(wedges, novf);
            hipLaunchKernelGGL(k_c4_wedges, dim3(gridx((uint64_t)novf * 64)), dim3(256), 0, st, foff.p, fkey.p,
                               sdeg.p, ovf.p, novf, below.p, wedges.p);
            ZG_HIP(hipGetLastError());
            std::vector<uint64_t> hbelow(novf), hwedge(novf);
            std::vector<uint32_t> hlist(novf);
            ZG_HIP(hipMemcpyAsync(hbelow.data(), below.p, 8ull * novf, hipMemcpyDeviceToHost, st));
            ZG_HIP(hipMemcpyAsync(hwedge.data(), wedges.p, 8ull * novf, hipMemcpyDeviceToHost, st));
            ZG_HIP(hipMemcpyAsync(hlist.data(), ovf.p, 4ull * novf, hipMemcpyDeviceToHost, st));
            ZG_HIP(hipStreamSynchronize(st));
            size_t free_b = 0, total_b = 0;
            ZG_HIP(hipMemGetInfo(&free_b, &total_b));
            const uint64_t budget = std::min<uint64_t>(16ull << 30, std::max<uint64_t>(1ull << 30, free_b / 4));
            // sorted batches of consecutive vertices, up to `cap` wedges each (about 24 bytes per
            // wedge); a vertex with more wedges than that goes to the dense counters instead
            // (houses: keys, 16-byte {count, weight} values and the reduction's outputs, up to 88 bytes per wedge)
            uint64_t cap = std::min<uint64_t>(budget / (weighted ? 88 : 24), 0xF0000000ull);
            if (const char *ev = getenv("ZGRAPH_C4_SORT_WEDGES")) cap = std::min<uint64_t>(cap, strtoull(ev, nullptr, 10));
            // keys (batch index << wbits | w): 32-bit while ceil(log2 n) + the batch index's bits fit,
            // which caps a batch at 2^(32 - wbits) vertices; 64-bit keys past n = 2^30
            uint32_t wbits = 1;
            while ((1ull << wbits) < n) ++wbits;
            // ZGRAPH_C4_WIDE_KEYS=1 forces the 64-bit keys and ZGRAPH_C4_BATCH_VERTICES caps the
            // vertices per batch (tests reach those branches on small graphs with them)
            const char *ew = getenv("ZGRAPH_C4_WIDE_KEYS");
            const bool narrow = wbits <= 30 && !(ew && ew[0] == '1');
            uint32_t nbcap = narrow ? (1u << (32 - wbits)) : 0xffffffffu;
            if (const char *ev = getenv("ZGRAPH_C4_BATCH_VERTICES"))
                nbcap = std::max<uint32_t>(1, std::min<uint32_t>(nbcap, (uint32_t)strtoul(ev, nullptr, 10)));
            std::vector<std::pair<uint32_t, uint32_t>> batches;   // [first, end) in the list
            std::vector<uint32_t> dense;
            uint64_t tmax = 0;
            uint32_t nbmax = 0;
            for (uint32_t i = 0; i < novf;) {
                if (hwedge[i] > cap) { dense.push_back(hlist[i]); ++i; continue; }
                uint64_t tot = 0;
                uint32_t j = i;
                while (j < novf && j - i < nbcap && hwedge[j] <= cap && tot + hwedge[j] <= cap) tot += hwedge[j++];
                batches.push_back({i, j});
                tmax = std::max(tmax, tot);
                nbmax = std::max(nbmax, j - i);
                i = j;
            }
            // houses: the hubs' weighted wedges, (key, {count, weight}) pairs sorted and reduced by key;
            // 32-bit keys (batch index << wbits | w) while they fit, else 64-bit ones
            // values packed in 8 bytes while the maximum degree bounds them (see cw_make)
            uint32_t cbits = 1;
            while ((1ull << cbits) <= g->max_degree) ++cbits;   // a run's count < 2^cbits
            const uint32_t sb = 1 + 2 * cbits;                  // a run's weight < 2 maxdeg^2 < 2^sb
            // (ZGRAPH_HOUSE_WIDE_VALUES=1 forces the 16-byte values, for tests)
            const char *hv = getenv("ZGRAPH_HOUSE_WIDE_VALUES");
            const bool packed = cbits + sb <= 64 && !(hv && hv[0] == '1');
            auto weighted_batches = [&](auto key0, auto val0) -> int {
                using Key = decltype(key0);
                using Val = decltype(val0);
                XB<Key> ka, kb, uk;
                XB<uint64_t> d_ipre;
                XB<Val> va, vb, agg;
                XB<unsigned int> nruns;
                XB<unsigned long long> cursor;
                XALLOC(ka, tmax); XALLOC(kb, tmax); XALLOC(uk, tmax);
                XALLOC(va, tmax); XALLOC(vb, tmax); XALLOC(agg, tmax); XALLOC(nruns, 1);
                XALLOC(d_ipre, (uint64_t)nbmax + 1); XALLOC(cursor, nbmax);
                const uint32_t kw = sizeof(Key) == 4 ? wbits : 32u;   // bits of w in the key
                size_t ts = 0, tr = 0;
                ZG_HIP(rocprim::radix_sort_pairs(nullptr, ts, ka.p, kb.p, va.p, vb.p, (size_t)tmax, 0,
                                                 (unsigned)(8 * sizeof(Key)), st));
                using Op = typename std::conditional<std::is_same<Val, CW>::value, CWPlus,
                                                     rocprim::plus<unsigned long long>>::type;
                ZG_HIP(rocprim::reduce_by_key(nullptr, tr, kb.p, vb.p, (size_t)tmax, uk.p, agg.p, nruns.p, Op(),
                                              rocprim::equal_to<Key>(), st));
                XB<char> tmps;
                XALLOC(tmps, std::max(ts, tr));
                const int pgrid = gridx(tmax);
                XB<unsigned long long> spart;
                XALLOC(spart, 2 * (uint64_t)pgrid);
                for (const auto &b : batches) {
                    const uint32_t nb = b.second - b.first;
                    std::vector<uint64_t> ipre(nb + 1, 0), cur(nb, 0);
                    uint64_t tot = 0;
                    for (uint32_t i = 0; i < nb; ++i) {
                        ipre[i + 1] = ipre[i] + hbelow[b.first + i];
                        cur[i] = tot;
                        tot += hwedge[b.first + i];
                    }
                    if (tot == 0) continue;
                    uint32_t kbits = 1;
                    while ((1ull << kbits) < nb) ++kbits;
                    const uint32_t ebits = kw + kbits;   // <= 32 for 32-bit keys (nbcap)
                    ZG_HIP(hipMemcpyAsync(d_ipre.p, ipre.data(), 8ull * (nb + 1), hipMemcpyHostToDevice, st));
                    ZG_HIP(hipMemcpyAsync(cursor.p, cur.data(), 8ull * nb, hipMemcpyHostToDevice, st));
                    const int fgrid = (int)std::min<uint64_t>((uint64_t)device_cus() * 16,
                                                              std::max<uint64_t>(1, (ipre[nb] + 3) / 4));
                    hipLaunchKernelGGL((k_c4s_fill_w<Key, Val>), dim3(fgrid), dim3(256), 0, st, foff.p, fkey.p, sdeg.p,
                                       ovf.p + b.first, d_ipre.p, nb, kw, cursor.p, tf.p, sb, ka.p, va.p);
                    ZG_HIP(hipGetLastError());
                    size_t t1 = 0, t2 = 0;
                    ZG_HIP(rocprim::radix_sort_pairs(nullptr, t1, ka.p, kb.p, va.p, vb.p, (size_t)tot, 0, ebits, st));
                    ZG_HIP(rocprim::reduce_by_key(nullptr, t2, kb.p, vb.p, (size_t)tot, uk.p, agg.p, nruns.p,
                                                  Op(), rocprim::equal_to<Key>(), st));
                    if (std::max(t1, t2) > std::max(ts, tr))
                        return set_error(ZG_ENOMEM, "house count: sort scratch larger than sized");
                    t1 = t2 = std::max(ts, tr);
                    ZG_HIP(rocprim::radix_sort_pairs(tmps.p, t1, ka.p, kb.p, va.p, vb.p, (size_t)tot, 0, ebits, st));
                    ZG_HIP(rocprim::reduce_by_key(tmps.p, t2, kb.p, vb.p, (size_t)tot, uk.p, agg.p, nruns.p,
                                                  Op(), rocprim::equal_to<Key>(), st));
                    const int g2 = gridx(tot);
                    hipLaunchKernelGGL(k_c4s_pairs_w<Val>, dim3(g2), dim3(256), 0, st, agg.p, nruns.p, sb, spart.p);
                    ZG_HIP(hipGetLastError());
                    unsigned __int128 g4 = 0;
                    const int rc2 = host_sum128(spart.p, g2, st, &g4);   // synchronizes before ipre and cur change
                    if (rc2) return rc2;
                    c4 += g4;
                }
                return ZG_OK;
            };
            if (weighted && !batches.empty() && tmax > 0) {
                const CW cw0{0, 0};
                const unsigned long long pv0 = 0;
                if (narrow) rc = packed ? weighted_batches(uint32_t(0), pv0) : weighted_batches(uint32_t(0), cw0);
                else rc = packed ? weighted_batches(uint64_t(0), pv0) : weighted_batches(uint64_t(0), cw0);
                if (rc) return rc;
            } else if (!batches.empty() && tmax > 0) {
                XB<uint64_t> ka, kb, d_ipre;
                XB<unsigned int> counts, nruns;
                XB<unsigned long long> cursor;
                XALLOC(ka, tmax); XALLOC(kb, tmax); XALLOC(counts, tmax); XALLOC(nruns, 1);
                XALLOC(d_ipre, (uint64_t)nbmax + 1); XALLOC(cursor, nbmax);
                uint32_t *const ka32 = reinterpret_cast<uint32_t *>(ka.p);
                uint32_t *const kb32 = reinterpret_cast<uint32_t *>(kb.p);
                size_t ts = 0, tr = 0;
                if (narrow) {
                    ZG_HIP(rocprim::radix_sort_keys(nullptr, ts, ka32, kb32, (size_t)tmax, 0, 32, st));
                    ZG_HIP(rocprim::run_length_encode(nullptr, tr, kb32, (unsigned int)tmax, ka32, counts.p,
                                                      nruns.p, st));
                } else {
                    ZG_HIP(rocprim::radix_sort_keys(nullptr, ts, ka.p, kb.p, (size_t)tmax, 0, 64, st));
                    ZG_HIP(rocprim::run_length_encode(nullptr, tr, kb.p, (unsigned int)tmax, ka.p, counts.p,
                                                      nruns.p, st));
                }
                XB<char> tmps;
                XALLOC(tmps, std::max(ts, tr));
                const int pgrid = gridx(tmax);
                XB<unsigned long long> spart;
                XALLOC(spart, 2 * (uint64_t)pgrid);
                for (const auto &b : batches) {
                    const uint32_t nb = b.second - b.first;
                    std::vector<uint64_t> ipre(nb + 1, 0), cur(nb, 0);
                    uint64_t tot = 0;
                    for (uint32_t i = 0; i < nb; ++i) {
                        ipre[i + 1] = ipre[i] + hbelow[b.first + i];
                        cur[i] = tot;
                        tot += hwedge[b.first + i];
                    }
                    if (tot == 0) continue;
                    uint32_t kbits = 1;
                    while ((1ull << kbits) < nb) ++kbits;
                    ZG_HIP(hipMemcpyAsync(d_ipre.p, ipre.data(), 8ull * (nb + 1), hipMemcpyHostToDevice, st));
                    ZG_HIP(hipMemcpyAsync(cursor.p, cur.data(), 8ull * nb, hipMemcpyHostToDevice, st));
                    const int fgrid = (int)std::min<uint64_t>((uint64_t)device_cus() * 16,
                                                              std::max<uint64_t>(1, (ipre[nb] + 3) / 4));
                    const uint32_t ebits = narrow ? wbits + kbits : 32 + kbits;   // <= 32 when narrow
                    size_t t1 = 0, t2 = 0;
                    if (narrow) {
                        hipLaunchKernelGGL(k_c4s_fill<uint32_t>, dim3(fgrid), dim3(256), 0, st, foff.p, fkey.p,
                                           sdeg.p, ovf.p + b.first, d_ipre.p, nb, wbits, cursor.p, ka32);
                        ZG_HIP(hipGetLastError());
                        ZG_HIP(rocprim::radix_sort_keys(nullptr, t1, ka32, kb32, (size_t)tot, 0, ebits, st));
                        ZG_HIP(rocprim::run_length_encode(nullptr, t2, kb32, (unsigned int)tot, ka32, counts.p,
                                                          nruns.p, st));
                    } else {
                        hipLaunchKernelGGL(k_c4s_fill<uint64_t>, dim3(fgrid), dim3(256), 0, st, foff.p, fkey.p,
                                           sdeg.p, ovf.p + b.first, d_ipre.p, nb, 32u, cursor.p, ka.p);
                        ZG_HIP(hipGetLastError());
                        ZG_HIP(rocprim::radix_sort_keys(nullptr, t1, ka.p, kb.p, (size_t)tot, 0, ebits, st));
                        ZG_HIP(rocprim::run_length_encode(nullptr, t2, kb.p, (unsigned int)tot, ka.p, counts.p,
                                                          nruns.p, st));
                    }
                    if (std::max(t1, t2) > std::max(ts, tr))
                        return set_error(ZG_ENOMEM, "4-cycle count: sort scratch larger than sized");
                    t1 = t2 = std::max(ts, tr);
                    if (narrow) {
                        ZG_HIP(rocprim::radix_sort_keys(tmps.p, t1, ka32, kb32, (size_t)tot, 0, ebits, st));
                        ZG_HIP(rocprim::run_length_encode(tmps.p, t2, kb32, (unsigned int)tot, ka32, counts.p,
                                                          nruns.p, st));
                    } else {
                        ZG_HIP(rocprim::radix_sort_keys(tmps.p, t1, ka.p, kb.p, (size_t)tot, 0, ebits, st));
                        ZG_HIP(rocprim::run_length_encode(tmps.p, t2, kb.p, (unsigned int)tot, ka.p, counts.p,
                                                          nruns.p, st));
                    }
                    const int g2 = gridx(tot);
                    hipLaunchKernelGGL(k_c4s_pairs, dim3(g2), dim3(256), 0, st, counts.p, nruns.p, spart.p);
                    ZG_HIP(hipGetLastError());
                    unsigned __int128 g4 = 0;
                    rc = host_sum128(spart.p, g2, st, &g4);   // synchronizes before ipre and cur change
                    if (rc) return rc;
                    c4 += g4;
                }
            }
            if (!dense.empty()) {
                // one dense counter over the n vertices per vertex, `per` vertices at a time
                const uint32_t nd = (uint32_t)dense.size();
                const uint64_t per_vertex = (weighted ? 12ull : 4ull) * n;   // counts (+ weights) over the n vertices
                const uint32_t per = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(nd, budget / per_vertex));
                XB<uint32_t> cnt, dlist;
                XB<unsigned long long> wsum;
                XB<uint64_t> d_ipre;
                XALLOC(cnt, (uint64_t)per * n); XALLOC(dlist, nd); XALLOC(d_ipre, (uint64_t)per + 1);
                if (weighted) XALLOC(wsum, (uint64_t)per * n);
                ZG_HIP(hipMemcpyAsync(dlist.p, dense.data(), 4ull * nd, hipMemcpyHostToDevice, st));
                std::vector<uint64_t> dbelow(nd);
                for (uint32_t i = 0, k = 0; i < novf; ++i)
                    if (hwedge[i] > cap) dbelow[k++] = hbelow[i];
                const int sgrid = gridx((uint64_t)per * n);
                XB<unsigned long long> gpart;
                XALLOC(gpart, 2 * (uint64_t)sgrid);
                for (uint32_t i0 = 0; i0 < nd; i0 += per) {
                    const uint32_t nb = std::min<uint32_t>(per, nd - i0);
                    std::vector<uint64_t> ipre(nb + 1, 0);
                    for (uint32_t i = 0; i < nb; ++i) ipre[i + 1] = ipre[i] + dbelow[i0 + i];
                    const uint64_t total = (uint64_t)nb * n;
                    ZG_HIP(hipMemcpyAsync(d_ipre.p, ipre.data(), 8ull * (nb + 1), hipMemcpyHostToDevice, st));
                    hipLaunchKernelGGL(k_c4d_clear, dim3(gridx(total)), dim3(256), 0, st, cnt.p, total);
                    if (weighted) hipLaunchKernelGGL(k_clear_u64, dim3(gridx(total)), dim3(256), 0, st, wsum.p, total);
                    ZG_HIP(hipGetLastError());
                    const int igrid = (int)std::min<uint64_t>((uint64_t)device_cus() * 16,
                                                              std::max<uint64_t>(1, (ipre[nb] + 3) / 4));
                    if (weighted)
                        hipLaunchKernelGGL(k_c4d_insert<true>, dim3(igrid), dim3(256), 0, st, foff.p, fkey.p, sdeg.p,
                                           dlist.p + i0, d_ipre.p, nb, n, cnt.p, tf.p, wsum.p);
                    else
                        hipLaunchKernelGGL(k_c4d_insert<false>, dim3(igrid), dim3(256), 0, st, foff.p, fkey.p, sdeg.p,
                                           dlist.p + i0, d_ipre.p, nb, n, cnt.p, nullptr, nullptr);
                    ZG_HIP(hipGetLastError());
                    const int g2 = gridx(total);
                    if (weighted)
                        hipLaunchKernelGGL(k_c4d_sum<true>, dim3(g2), dim3(256), 0, st, cnt.p, total, wsum.p, gpart.p);
                    else
                        hipLaunchKernelGGL(k_c4d_sum<false>, dim3(g2), dim3(256), 0, st, cnt.p, total, nullptr,
                                           gpart.p);
                    ZG_HIP(hipGetLastError());
                    unsigned __int128 g4 = 0;
                    rc = host_sum128(gpart.p, g2, st, &g4);   // synchronizes before ipre changes
                    if (rc) return rc;
                    c4 += g4;
                }
            }
        }
        if (weighted) {
            if (c4 < 2 * tt) return set_error(ZG_EINVAL, "house count: inconsistent per-edge counts");
            const unsigned __int128 h = c4 - 2 * tt;
            if (h >> 64) return set_error(ZG_EOVERFLOW, "5-house count exceeds 64 bits");
            *out = (uint64_t)h;
            return ZG_OK;
        }
        // A simple graph with m edges has at most C(m, 2) 4-cycles (each holds two pairs of
        // disjoint edges, and each such pair closes at most two 4-cycles), so with 2m < 2^32
        // the count stays below 2^61; the check guards the summation itself.
        if (c4 >> 64) return set_error(ZG_EOVERFLOW, "4-cycle count exceeds 64 bits");
        if (what == 7) { *out = (uint64_t)c4; return ZG_OK; }
        // 5-vertex chains (FiveChain.cpp semantics): 2P = sum_v (S1^2 - S2)
        // - 4 sum_v (d_v - 1) t_v + 6 T - 8 C4
        const int tgrid = (int)std::min<uint64_t>((uint64_t)device_cus() * 16, std::max<uint64_t>(1, (ne + 3) / 4));
        XB<unsigned long long> tpart, wpart;
        XALLOC(tpart, tgrid); XALLOC(wpart, tgrid);
        hipLaunchKernelGGL(k_tri_wave<true>, dim3(tgrid), dim3(256), 0, st, ooff.p, onb.p, osrc.p, ne, sdeg.p,
                           tpart.p, wpart.p);
        ZG_HIP(hipGetLastError());
        unsigned long long T = 0, TW = 0;
        rc = host_sum(tpart.p, tgrid, st, &T);
        if (!rc) rc = host_sum(wpart.p, tgrid, st, &TW);
        if (rc) return rc;
        const int pgrid = gridx((uint64_t)n * 64);
        XB<unsigned long long> ppart;
        XALLOC(ppart, (uint64_t)pgrid * 8);   // (lo, hi) per wave, 4 waves per block
        hipLaunchKernelGGL(k_c5_pairs, dim3(pgrid), dim3(256), 0, st, foff.p, fkey.p, n, ppart.p);
        ZG_HIP(hipGetLastError());
        std::vector<unsigned long long> hp((size_t)pgrid * 8);
        ZG_HIP(hipMemcpyAsync(hp.data(), ppart.p, 8ull * hp.size(), hipMemcpyDeviceToHost, st));
        ZG_HIP(hipStreamSynchronize(st));
        unsigned __int128 W = 0;
        for (size_t i = 0; i < hp.size(); i += 2) W += ((unsigned __int128)hp[i + 1] << 64) | hp[i];
        const __int128 two_p = (__int128)W - 4 * (__int128)TW + 6 * (__int128)T - 8 * (__int128)c4;
        if (two_p < 0 || (two_p >> 65) != 0)
            return set_error(ZG_EOVERFLOW, "5-vertex chain count exceeds 64 bits");
        *out = (uint64_t)(two_p / 2);
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
    uint64_t total_words = 0;
    int rc = read_u64(boff.p + n, &total_words);
    if (rc) return rc;
    if (total_words == 0) { *out = 0; return ZG_OK; }
    size_t free_b = 0, total_b = 0;
    ZG_HIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t one_vertex = maxout * ((maxout + 63) >> 6);
    // row words per pass: a third of the free memory, at most 2^32 (32 GiB, the
    // row-word offsets are 32-bit)
    uint64_t budget = std::min<uint64_t>(1ull << 32, std::max<uint64_t>(1ull << 27, free_b / 24));
    if (const char *ev = getenv("ZGRAPH_K4_SCRATCH_WORDS")) budget = strtoull(ev, nullptr, 10);   // smaller passes
    budget = std::max(budget, one_vertex);
    XB<uint64_t> bits;
    XALLOC(bits, std::min(total_words, budget));
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
            uint64_t b1 = 0;
            rc = read_u64(boff.p + u1, &b1);
            if (rc) return rc;
            ZG_HIP(hipMemsetAsync(bits.p, 0, 8 * (b1 - b0), st));   // rows of edges into heads with no out-list stay 0
            for (int heavy = 0; heavy < 2; ++heavy) {
                if (nitems[heavy] == 0) continue;
                HeadArgs A = head_args(heavy);
                A.boff = boff.p;
                A.bbase = b0;
                A.u0 = u0;
                A.u1 = u1;
                A.rows = bits.p;
                if (heavy)
                    hipLaunchKernelGGL(k_head_heavy<1>, dim3(hgrid[1]), dim3(kTriHeavyThreads), 0, st, A, nullptr);
                else
                    hipLaunchKernelGGL(k_head_light<1>, dim3(hgrid[0]), dim3(256), 0, st, A, nullptr);
                ZG_HIP(hipGetLastError());
            }
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
