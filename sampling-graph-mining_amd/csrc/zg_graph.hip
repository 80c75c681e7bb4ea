// zg_graph.hip -- device CSR construction, the ELP graph subsampler, the
// predicate filter and the seeded synthetic generators.
//
// CSRConstructor::construct_bitmap_assisted_csr (src/Graph.cpp:214-377)
// shuffles the edge list, takes rank := post-shuffle index, counts degrees,
// prefix-sums them and fills every list in rank order.  Here the shuffle is a
// keyed random permutation (radix sort of a Philox key per input edge, ties by
// index), the fill is a stable radix sort of the 2m list entries by owner
// (entries enter in rank order, so lists come out rank-ascending), and a
// second stable sort by (owner, neighbour) builds the neighbour-ordered copy
// used by the closing checks.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "zg_internal.h"

namespace zg {

namespace {

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t k)
    {
        n = k;
        return hipMalloc(&p, sizeof(T) * (k ? k : 1));
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

#define ZG_ALLOC(buf, k)                                                          \
    do {                                                                          \
        hipError_t _e = (buf).alloc(k);                                           \
        if (_e != hipSuccess) return set_error(ZG_ENOMEM, "device allocation failed"); \
    } while (0)

inline unsigned bits_for(uint64_t x)
{
    unsigned b = 1;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

inline int grid_for(uint64_t n, int block = 256)
{
    uint64_t g = (n + block - 1) / block;
    const uint64_t cap = (uint64_t)device_cus() * 64;
    if (g > cap) g = cap;
    return (int)(g ? g : 1);
}

#define GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

__global__ void k_validate(const uint32_t *src, const uint32_t *dst, uint64_t m, uint32_t n, uint32_t *bad)
{
    GRID_STRIDE(i, m)
    {
        const uint32_t s = src[i], d = dst[i];
        if (s >= n || d >= n) atomicOr(bad, 2u);
        else if (s == d) atomicOr(bad, 1u);
    }
}

__global__ void k_shuffle_keys(uint64_t *keys, uint32_t *idx, uint64_t m, uint64_t seed)
{
    GRID_STRIDE(i, m)
    {
        keys[i] = key64(seed, TAG_SHUFFLE, i);
        idx[i] = (uint32_t)i;
    }
}

// gather by rank, count degrees, emit the 2m list entries in rank order
__global__ void k_gather(const uint32_t *src, const uint32_t *dst, const float *w, const uint32_t *perm, uint64_t m,
                         ERec *edges, float *wout, unsigned long long *deg, uint32_t *own, uint64_t *ent)
{
    GRID_STRIDE(r, m)
    {
        const uint32_t i = perm[r];
        const uint32_t s = src[i], d = dst[i];
        ERec e;
        e.src = s; e.dst = d;
        e.s_off = e.s_deg = e.d_off = e.d_deg = e.s_pos = e.d_pos = 0;
        edges[r] = e;
        if (w) wout[r] = w[i];
        atomicAdd(&deg[s], 1ull);
        atomicAdd(&deg[d], 1ull);
        own[2 * r] = s;
        ent[2 * r] = (r << 32) | d;
        own[2 * r + 1] = d;
        ent[2 * r + 1] = (r << 32) | s;
    }
}

// each rank's endpoint lists (start, degree, position of the rank) into its
// record; neighbour-order sort keys
__global__ void k_positions(const uint32_t *own, const uint64_t *adj, const uint64_t *off, uint64_t n2, ERec *edges,
                            uint64_t *skey, uint32_t *sval)
{
    GRID_STRIDE(i, n2)
    {
        const uint32_t o = own[i];
        const uint64_t e = adj[i];
        const uint32_t r = (uint32_t)(e >> 32), nb = (uint32_t)e;
        const uint64_t b = off[o];
        const uint32_t p = (uint32_t)(i - b), dg = (uint32_t)(off[o + 1] - b);
        if (edges[r].src == o) { edges[r].s_off = (uint32_t)b; edges[r].s_deg = dg; edges[r].s_pos = p; }
        else { edges[r].d_off = (uint32_t)b; edges[r].d_deg = dg; edges[r].d_pos = p; }
        skey[i] = ((uint64_t)o << 32) | nb;
        sval[i] = r;
    }
}

__global__ void k_sadj(const uint64_t *skey, const uint32_t *sval, uint64_t n2, uint64_t *sadj)
{
    GRID_STRIDE(i, n2) { sadj[i] = ((skey[i] & 0xffffffffull) << 32) | sval[i]; }
}

// any two equal consecutive (owner, neighbour) keys: a pair joined by parallel edges
__global__ void k_parallel_flag(const uint64_t *skey, uint64_t n2, uint32_t *flag)
{
    GRID_STRIDE(i, n2 - 1)
    {
        if (skey[i] == skey[i + 1]) *flag = 1u;
    }
}

// edge-set hash: one slot per rank.  A bucket's slots are claimed with an
// atomic add on its fill count; the claimant writes the slot and then ORs its
// tag into the header.  A pair whose bucket is full marks it as passed over
// and tries the next one.  Which slot a copy lands in depends on the order of
// the atomics; a lookup reads every slot whose tag matches, so its answer
// does not.
__global__ void k_hash_insert(const ERec *edges, uint64_t m, HBucket *tab, uint32_t nb)
{
    GRID_STRIDE(r, m)
    {
        const uint32_t a = edges[r].src, b = edges[r].dst;
        const uint64_t key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
        const uint64_t h = hmix(key);
        const uint32_t t = htag_of(h);
        uint32_t bk = hbucket_of(h, nb);
        for (;;) {
            const uint32_t s = atomicAdd(&tab[bk].fill, 1u);
            if (s < (uint32_t)kBucketSlots) {
                tab[bk].slot[s] = HSlot{(uint32_t)key, (uint32_t)(key >> 32), (uint32_t)r};
                atomicOr(&tab[bk].tag[s >> 1], t << (16 * (s & 1)));
                break;
            }
            atomicOr(&tab[bk].tag[0], kBucketPassed);
            bk = (bk + 1 == nb) ? 0 : bk + 1;
        }
    }
}

__global__ void k_deg_to_u64(const unsigned long long *deg, uint64_t *out, uint64_t n1)
{
    GRID_STRIDE(i, n1) { out[i] = deg[i]; }
}

}  // namespace

void free_graph(zg_graph *g)
{
    if (!g) return;
    if (g->edges) (void)hipFree(g->edges);
    if (g->off) (void)hipFree(g->off);
    if (g->adj) (void)hipFree(g->adj);
    if (g->sadj) (void)hipFree(g->sadj);
    if (g->htab) (void)hipFree(g->htab);
    if (g->w) (void)hipFree(g->w);
    if (g->partials) (void)hipFree(g->partials);
    if (g->result) (void)hipFree(g->result);
    if (g->outbuf) (void)hipFree(g->outbuf);
    for (uint32_t *q : g->ord)
        if (q) (void)hipFree(q);
    if (g->ord_tmp) (void)hipFree(g->ord_tmp);
    if (g->ev0) (void)hipEventDestroy(g->ev0);
    if (g->ev1) (void)hipEventDestroy(g->ev1);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

int build_graph(const uint32_t *d_src, const uint32_t *d_dst, const float *d_w, uint64_t m, uint32_t n,
                uint64_t seed, zg_graph **out)
{
    if (m >= (1ull << 31)) return set_error(ZG_EINVAL, "graphs need fewer than 2^31 edges (u32 list offsets)");
    if (n == 0 || n == 0xffffffffu) return set_error(ZG_EINVAL, "vertex count must be in [1, 2^32-1)");
    zg_graph *g = new zg_graph();
    struct Guard {
        zg_graph *g;
        ~Guard() { if (g) free_graph(g); }
    } guard{g};
    ZG_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    ZG_HIP(hipEventCreate(&g->ev0));
    ZG_HIP(hipEventCreate(&g->ev1));
    ZG_HIP(hipMalloc(&g->result, sizeof(Acc) * 2));
    hipStream_t st = g->stream;
    g->n = n; g->m = m; g->num_edges = m;
    const uint64_t n2 = 2 * m;

    {   // input validation: self-loops (src/Graph.cpp:454) and ids
        DevBuf<uint32_t> bad;
        ZG_ALLOC(bad, 1);
        ZG_HIP(hipMemsetAsync(bad.p, 0, 4, st));
        if (m) hipLaunchKernelGGL(k_validate, dim3(grid_for(m)), dim3(256), 0, st, d_src, d_dst, m, n, bad.p);
        uint32_t hb = 0;
        ZG_HIP(hipMemcpyAsync(&hb, bad.p, 4, hipMemcpyDeviceToHost, st));
        ZG_HIP(hipStreamSynchronize(st));
        if (hb & 2u) return set_error(ZG_EINVAL, "edge endpoint >= vertex count");
        if (hb & 1u) return set_error(ZG_EINVAL, "self-loops are not allowed in the input graph");
    }

    ZG_HIP(hipMalloc(&g->edges, sizeof(ERec) * (m ? m : 1)));
    ZG_HIP(hipMalloc(&g->off, sizeof(uint64_t) * ((uint64_t)n + 1)));
    ZG_HIP(hipMalloc(&g->adj, sizeof(uint64_t) * (n2 ? n2 : 1)));
    ZG_HIP(hipMalloc(&g->sadj, sizeof(uint64_t) * (n2 ? n2 : 1)));
    if (d_w) ZG_HIP(hipMalloc(&g->w, sizeof(float) * (m ? m : 1)));
    g->bytes = sizeof(ERec) * m + 8ull * (n + 1) + 16ull * n2 + (d_w ? 4ull * m : 0);

    if (m == 0) {   // no edges: empty lists, zero offsets
        ZG_HIP(hipMemsetAsync(g->off, 0, 8ull * (n + 1), st));
        ZG_HIP(hipStreamSynchronize(st));
        g->max_degree = 0;
        guard.g = nullptr;
        *out = g;
        return ZG_OK;
    }

    // 1. rank permutation
    DevBuf<uint32_t> perm;
    {
        DevBuf<uint64_t> k0, k1;
        DevBuf<uint32_t> i1;
        ZG_ALLOC(k0, m); ZG_ALLOC(k1, m); ZG_ALLOC(perm, m); ZG_ALLOC(i1, m);
        if (m) hipLaunchKernelGGL(k_shuffle_keys, dim3(grid_for(m)), dim3(256), 0, st, k0.p, perm.p, m, seed);
        rocprim::double_buffer<uint64_t> kb(k0.p, k1.p);
        rocprim::double_buffer<uint32_t> vb(perm.p, i1.p);
        size_t tb = 0;
        ZG_HIP(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, m, 0, 64, st));
        DevBuf<char> tmp;
        ZG_ALLOC(tmp, tb);
        ZG_HIP(rocprim::radix_sort_pairs(tmp.p, tb, kb, vb, m, 0, 64, st));
        if (vb.current() != perm.p)
            ZG_HIP(hipMemcpyAsync(perm.p, vb.current(), 4 * m, hipMemcpyDeviceToDevice, st));
        ZG_HIP(hipStreamSynchronize(st));
    }

    // 2. gather by rank, degrees, list entries
    DevBuf<unsigned long long> deg;
    DevBuf<uint32_t> own0, own1;
    DevBuf<uint64_t> ent1;
    ZG_ALLOC(deg, (uint64_t)n + 1);
    ZG_ALLOC(own0, n2); ZG_ALLOC(own1, n2); ZG_ALLOC(ent1, n2);
    ZG_HIP(hipMemsetAsync(deg.p, 0, 8ull * (n + 1), st));
    if (m)
        hipLaunchKernelGGL(k_gather, dim3(grid_for(m)), dim3(256), 0, st, d_src, d_dst, d_w, perm.p, m, g->edges,
                           g->w, deg.p, own0.p, g->adj);
    ZG_HIP(hipGetLastError());

    // 3. offsets = exclusive prefix of degrees; max degree
    {
        DevBuf<uint64_t> d64;
        ZG_ALLOC(d64, (uint64_t)n + 1);
        hipLaunchKernelGGL(k_deg_to_u64, dim3(grid_for(n + 1ull)), dim3(256), 0, st, deg.p, d64.p, n + 1ull);
        size_t tb = 0;
        ZG_HIP(rocprim::exclusive_scan(nullptr, tb, d64.p, g->off, (uint64_t)0, (size_t)n + 1,
                                       rocprim::plus<uint64_t>(), st));
        DevBuf<char> tmp;
        ZG_ALLOC(tmp, tb);
        ZG_HIP(rocprim::exclusive_scan(tmp.p, tb, d64.p, g->off, (uint64_t)0, (size_t)n + 1,
                                       rocprim::plus<uint64_t>(), st));
        DevBuf<uint64_t> mx;
        ZG_ALLOC(mx, 1);
        size_t tr = 0;
        ZG_HIP(rocprim::reduce(nullptr, tr, d64.p, mx.p, (uint64_t)0, (size_t)n, rocprim::maximum<uint64_t>(), st));
        DevBuf<char> tmp2;
        ZG_ALLOC(tmp2, tr);
        ZG_HIP(rocprim::reduce(tmp2.p, tr, d64.p, mx.p, (uint64_t)0, (size_t)n, rocprim::maximum<uint64_t>(), st));
        ZG_HIP(hipMemcpyAsync(&g->max_degree, mx.p, 8, hipMemcpyDeviceToHost, st));
        ZG_HIP(hipStreamSynchronize(st));
    }

    // 4. stable sort of the list entries by owner -> rank-ascending lists
    {
        rocprim::double_buffer<uint32_t> kb(own0.p, own1.p);
        rocprim::double_buffer<uint64_t> vb(g->adj, ent1.p);
        const unsigned eb = bits_for(n);
        size_t tb = 0;
        ZG_HIP(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, n2, 0, eb, st));
        DevBuf<char> tmp;
        ZG_ALLOC(tmp, tb);
        ZG_HIP(rocprim::radix_sort_pairs(tmp.p, tb, kb, vb, n2, 0, eb, st));
        if (vb.current() != g->adj)
            ZG_HIP(hipMemcpyAsync(g->adj, vb.current(), 8 * n2, hipMemcpyDeviceToDevice, st));
        if (kb.current() != own0.p)
            ZG_HIP(hipMemcpyAsync(own0.p, kb.current(), 4 * n2, hipMemcpyDeviceToDevice, st));
        ZG_HIP(hipStreamSynchronize(st));
    }

    // 5. positions of each rank in its endpoint lists + neighbour-ordered copy
    {
        DevBuf<uint64_t> sk0, sk1;
        DevBuf<uint32_t> sv0, sv1;
        ZG_ALLOC(sk0, n2); ZG_ALLOC(sk1, n2); ZG_ALLOC(sv0, n2); ZG_ALLOC(sv1, n2);
        if (n2)
            hipLaunchKernelGGL(k_positions, dim3(grid_for(n2)), dim3(256), 0, st, own0.p, g->adj, g->off, n2, g->edges,
                               sk0.p, sv0.p);
        ZG_HIP(hipGetLastError());
        rocprim::double_buffer<uint64_t> kb(sk0.p, sk1.p);
        rocprim::double_buffer<uint32_t> vb(sv0.p, sv1.p);
        const unsigned eb = 32 + bits_for(n);
        size_t tb = 0;
        ZG_HIP(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, n2, 0, eb, st));
        DevBuf<char> tmp;
        ZG_ALLOC(tmp, tb);
        ZG_HIP(rocprim::radix_sort_pairs(tmp.p, tb, kb, vb, n2, 0, eb, st));
        if (n2) hipLaunchKernelGGL(k_sadj, dim3(grid_for(n2)), dim3(256), 0, st, kb.current(), vb.current(), n2, g->sadj);
        ZG_HIP(hipGetLastError());
        DevBuf<uint32_t> par;
        ZG_ALLOC(par, 1);
        ZG_HIP(hipMemsetAsync(par.p, 0, 4, st));
        if (n2 > 1)
            hipLaunchKernelGGL(k_parallel_flag, dim3(grid_for(n2 - 1)), dim3(256), 0, st, kb.current(), n2, par.p);
        ZG_HIP(hipGetLastError());
        uint32_t hpar = 0;
        ZG_HIP(hipMemcpyAsync(&hpar, par.p, 4, hipMemcpyDeviceToHost, st));
        ZG_HIP(hipStreamSynchronize(st));
        g->simple = hpar == 0;
    }
    // 6. edge-set hash for the closing checks (4 edges per 8-slot bucket on average)
    {
        const uint64_t nb = (m + kEdgesPerBucket - 1) / kEdgesPerBucket;
        if (nb >= (1ull << 32)) return set_error(ZG_EINVAL, "too many edges for the edge hash");
        g->hnb = (uint32_t)nb;
        ZG_HIP(hipMalloc(&g->htab, sizeof(HBucket) * nb));
        ZG_HIP(hipMemsetAsync(g->htab, 0, sizeof(HBucket) * nb, st));
        hipLaunchKernelGGL(k_hash_insert, dim3(grid_for(m)), dim3(256), 0, st, g->edges, m, g->htab, g->hnb);
        ZG_HIP(hipGetLastError());
        ZG_HIP(hipStreamSynchronize(st));
        g->bytes += sizeof(HBucket) * nb;
    }
    guard.g = nullptr;
    *out = g;
    return ZG_OK;
}

// ---- subsampler / filter -------------------------------------------------
namespace {

__global__ void k_subsample_flags(uint64_t m, uint64_t seed, uint64_t thr, uint8_t *flags)
{
    GRID_STRIDE(r, m)
    {
        Rng rng;
        rng.init(seed, TAG_SUBSAMPLE, r);
        flags[r] = (1 + rng.uniform(1ull << 30)) <= thr;   // src/Graph.cpp:652-667
    }
}

__global__ void k_filter_flags(const float *w, uint64_t m, float thr, int keep_ge, uint8_t *flags)
{
    GRID_STRIDE(r, m) { flags[r] = keep_ge ? (w[r] >= thr) : (w[r] < thr); }
}

__global__ void k_gather_kept(const ERec *edges, const float *w, const uint32_t *sel, uint64_t k, uint32_t *s,
                              uint32_t *d, float *wo)
{
    GRID_STRIDE(i, k)
    {
        const ERec &e = edges[sel[i]];
        s[i] = e.src;
        d[i] = e.dst;
        if (w) wo[i] = w[sel[i]];
    }
}

int rebuild_kept(const zg_graph *in, const uint8_t *flags, uint64_t seed, zg_graph **out)
{
    hipStream_t st = in->stream;
    const uint64_t m = in->m;
    DevBuf<uint32_t> sel;
    DevBuf<uint64_t> cnt;
    ZG_ALLOC(sel, m); ZG_ALLOC(cnt, 1);
    size_t tb = 0;
    rocprim::counting_iterator<uint32_t> it(0);
    ZG_HIP(rocprim::select(nullptr, tb, it, flags, sel.p, cnt.p, (size_t)m, st));
    DevBuf<char> tmp;
    ZG_ALLOC(tmp, tb);
    ZG_HIP(rocprim::select(tmp.p, tb, it, flags, sel.p, cnt.p, (size_t)m, st));
    uint64_t k = 0;
    ZG_HIP(hipMemcpyAsync(&k, cnt.p, 8, hipMemcpyDeviceToHost, st));
    ZG_HIP(hipStreamSynchronize(st));
    DevBuf<uint32_t> s, d;
    DevBuf<float> w;
    ZG_ALLOC(s, k); ZG_ALLOC(d, k);
    if (in->w) ZG_ALLOC(w, k);
    if (k)
        hipLaunchKernelGGL(k_gather_kept, dim3(grid_for(k)), dim3(256), 0, st, in->edges, in->w, sel.p, k, s.p, d.p,
                           in->w ? w.p : nullptr);
    ZG_HIP(hipStreamSynchronize(st));
    return build_graph(s.p, d.p, in->w ? w.p : nullptr, k, in->n, seed, out);
}

}  // namespace

int subsample_graph(const zg_graph *in, double rate, uint64_t seed, zg_graph **out)
{
    if (!(rate >= 0.0 && rate <= 1.0)) return set_error(ZG_EINVAL, "sample rate must be in [0, 1]");
    const uint64_t thr = (uint64_t)(rate * (double)(1ull << 30));
    DevBuf<uint8_t> flags;
    ZG_ALLOC(flags, in->m);
    if (in->m)
        hipLaunchKernelGGL(k_subsample_flags, dim3(grid_for(in->m)), dim3(256), 0, in->stream, in->m, seed, thr,
                           flags.p);
    ZG_HIP(hipGetLastError());
    return rebuild_kept(in, flags.p, seed, out);
}

int filter_graph(const zg_graph *in, float thr, int keep_ge, uint64_t seed, zg_graph **out)
{
    if (!in->w) return set_error(ZG_EINVAL, "graph has no edge weights to filter on");
    DevBuf<uint8_t> flags;
    ZG_ALLOC(flags, in->m);
    if (in->m)
        hipLaunchKernelGGL(k_filter_flags, dim3(grid_for(in->m)), dim3(256), 0, in->stream, in->w, in->m, thr,
                           keep_ge, flags.p);
    ZG_HIP(hipGetLastError());
    return rebuild_kept(in, flags.p, seed, out);
}

}  // namespace zg
