// zg_gen.hip -- seeded synthetic graphs built directly in HBM (the benchmark
// inputs; the reference reads SNAP files through datasets/*/scripts instead).
// Candidate pair i draws from Philox (seed, TAG_GEN, i); the graph keeps the
// first m distinct undirected non-loop pairs in candidate order, stored as
// (min, max).  The result depends only on (parameters, seed).
#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "zg_internal.h"

namespace zg {
namespace {

enum : uint32_t { TAG_WEIGHT = 4 };

template <class T>
struct DB {
    T *p = nullptr;
    hipError_t alloc(size_t k) { return hipMalloc(&p, sizeof(T) * (k ? k : 1)); }
    ~DB() { if (p) (void)hipFree(p); }
    T *release() { T *q = p; p = nullptr; return q; }
};
#define GALLOC(b, k)                                                                   \
    do {                                                                               \
        if ((b).alloc(k) != hipSuccess) return set_error(ZG_ENOMEM, "device allocation failed"); \
    } while (0)
#define GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

inline int gridn(uint64_t n)
{
    uint64_t g = (n + 255) / 256, cap = (uint64_t)device_cus() * 64;
    return (int)std::max<uint64_t>(1, std::min(g, cap));
}

__device__ __forceinline__ uint64_t pair_key(uint32_t u, uint32_t v)
{
    if (u == v) return ~0ull;
    const uint32_t a = u < v ? u : v, b = u < v ? v : u;
    return ((uint64_t)a << 32) | b;
}

struct GenER { uint32_t n; };
struct GenCL { uint32_t n; const double *prob; const uint32_t *alias; };
struct GenRMAT { int scale; double a, ab, abc; };

__device__ __forceinline__ uint32_t draw_er(Rng &r, const GenER &p) { return (uint32_t)r.uniform(p.n); }
__device__ __forceinline__ uint32_t draw_cl(Rng &r, const GenCL &p)
{
    const uint32_t j = (uint32_t)r.uniform(p.n);
    const uint64_t hi = r.next(), lo = r.next();
    const double coin = (double)((hi << 21) ^ (lo >> 11)) * 0x1p-53;   // 53 random bits
    return coin < p.prob[j] ? j : p.alias[j];
}

__global__ void k_cand_er(GenER p, uint64_t seed, uint64_t M, uint64_t *key, uint32_t *idx)
{
    GRID_STRIDE(i, M)
    {
        Rng r;
        r.init(seed, TAG_GEN, i);
        const uint32_t u = draw_er(r, p), v = draw_er(r, p);
        key[i] = pair_key(u, v);
        idx[i] = (uint32_t)i;
    }
}
__global__ void k_cand_cl(GenCL p, uint64_t seed, uint64_t M, uint64_t *key, uint32_t *idx)
{
    GRID_STRIDE(i, M)
    {
        Rng r;
        r.init(seed, TAG_GEN, i);
        const uint32_t u = draw_cl(r, p), v = draw_cl(r, p);
        key[i] = pair_key(u, v);
        idx[i] = (uint32_t)i;
    }
}
__global__ void k_cand_rmat(GenRMAT p, uint64_t seed, uint64_t M, uint64_t *key, uint32_t *idx)
{
    GRID_STRIDE(i, M)
    {
        Rng r;
        r.init(seed, TAG_GEN, i);
        uint32_t u = 0, v = 0;
        for (int l = 0; l < p.scale; ++l) {
            const double t = (double)r.next() * 0x1p-32;
            const uint32_t bu = t >= p.ab, bv = (t >= p.a && t < p.ab) || t >= p.abc;
            u = (u << 1) | bu;
            v = (v << 1) | bv;
        }
        key[i] = pair_key(u, v);
        idx[i] = (uint32_t)i;
    }
}

__global__ void k_first_flags(const uint64_t *key, uint64_t M, uint8_t *flag)
{
    GRID_STRIDE(j, M) { flag[j] = key[j] != ~0ull && (j == 0 || key[j] != key[j - 1]); }
}

__global__ void k_split(const uint64_t *key, uint64_t m, uint32_t *s, uint32_t *d)
{
    GRID_STRIDE(i, m)
    {
        s[i] = (uint32_t)(key[i] >> 32);
        d[i] = (uint32_t)key[i];
    }
}

__global__ void k_weights(uint64_t m, uint64_t seed, float *w)
{
    GRID_STRIDE(i, m)
    {
        Rng r;
        r.init(seed, TAG_WEIGHT, i);
        w[i] = (float)(r.next() >> 8) * 0x1p-24f;   // U[0,1) on the float grid
    }
}

// keeps the first m distinct pairs by candidate index; returns ZG_ESTATE if
// the M candidates held fewer than m distinct pairs
template <class Launch>
int dedupe_first(Launch launch, uint64_t M, uint64_t m, hipStream_t st, zg_edges *out)
{
    DB<uint64_t> k0, k1;
    DB<uint32_t> i0, i1;
    GALLOC(k0, M); GALLOC(k1, M); GALLOC(i0, M); GALLOC(i1, M);
    launch(k0.p, i0.p);
    ZG_HIP(hipGetLastError());
    {
        rocprim::double_buffer<uint64_t> kb(k0.p, k1.p);
        rocprim::double_buffer<uint32_t> vb(i0.p, i1.p);
        size_t tb = 0;
        ZG_HIP(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, M, 0, 64, st));
        DB<char> tmp;
        GALLOC(tmp, tb);
        ZG_HIP(rocprim::radix_sort_pairs(tmp.p, tb, kb, vb, M, 0, 64, st));
        ZG_HIP(hipStreamSynchronize(st));
        if (kb.current() != k0.p) std::swap(k0.p, k1.p);
        if (vb.current() != i0.p) std::swap(i0.p, i1.p);
    }
    DB<uint8_t> flag;
    DB<uint64_t> cnt;
    GALLOC(flag, M); GALLOC(cnt, 1);
    hipLaunchKernelGGL(k_first_flags, dim3(gridn(M)), dim3(256), 0, st, k0.p, M, flag.p);
    {   // compact (idx, key) of first occurrences into (i1, k1)
        size_t tb = 0;
        ZG_HIP(rocprim::select(nullptr, tb, i0.p, flag.p, i1.p, cnt.p, (size_t)M, st));
        DB<char> tmp;
        GALLOC(tmp, tb);
        ZG_HIP(rocprim::select(tmp.p, tb, i0.p, flag.p, i1.p, cnt.p, (size_t)M, st));
        size_t tb2 = 0;
        ZG_HIP(rocprim::select(nullptr, tb2, k0.p, flag.p, k1.p, cnt.p, (size_t)M, st));
        DB<char> tmp2;
        GALLOC(tmp2, tb2);
        ZG_HIP(rocprim::select(tmp2.p, tb2, k0.p, flag.p, k1.p, cnt.p, (size_t)M, st));
    }
    uint64_t U = 0;
    ZG_HIP(hipMemcpyAsync(&U, cnt.p, 8, hipMemcpyDeviceToHost, st));
    ZG_HIP(hipStreamSynchronize(st));
    if (U < m) return ZG_ESTATE;
    {   // order the distinct pairs by candidate index
        rocprim::double_buffer<uint32_t> kb(i1.p, i0.p);
        rocprim::double_buffer<uint64_t> vb(k1.p, k0.p);
        size_t tb = 0;
        ZG_HIP(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, U, 0, 32, st));
        DB<char> tmp;
        GALLOC(tmp, tb);
        ZG_HIP(rocprim::radix_sort_pairs(tmp.p, tb, kb, vb, U, 0, 32, st));
        DB<uint32_t> s, d;
        GALLOC(s, m); GALLOC(d, m);
        if (m) hipLaunchKernelGGL(k_split, dim3(gridn(m)), dim3(256), 0, st, vb.current(), m, s.p, d.p);
        ZG_HIP(hipStreamSynchronize(st));
        out->d_src = s.release();
        out->d_dst = d.release();
        out->d_w = nullptr;
        out->m = m;
    }
    return ZG_OK;
}

template <class MakeLaunch>
int generate(MakeLaunch make, uint64_t m, double first_factor, zg_edges *out)
{
    int rc = ensure_init();
    if (rc) return rc;
    hipStream_t st;
    ZG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    double f = first_factor;
    rc = ZG_ESTATE;
    for (int attempt = 0; attempt < 12 && rc == ZG_ESTATE; ++attempt) {
        uint64_t M = (uint64_t)std::ceil((double)m * f) + 1024;
        if (M >= (1ull << 32)) M = (1ull << 32) - 1;
        rc = dedupe_first(make(st, M), M, m, st, out);
        if (rc == ZG_ESTATE && M == (1ull << 32) - 1) break;
        f *= 1.5;
    }
    (void)hipStreamDestroy(st);
    if (rc == ZG_ESTATE) return set_error(ZG_EINVAL, "generator could not reach the requested distinct edge count");
    return rc;
}

}  // namespace
}  // namespace zg

using namespace zg;

extern "C" int zg_gen_erdos_renyi(uint32_t n, uint64_t m, uint64_t seed, zg_edges *out)
{
    if (!out || n < 2) return set_error(ZG_EINVAL, "bad arguments");
    if ((double)m > 0.5 * (double)n * (double)(n - 1)) return set_error(ZG_EINVAL, "m exceeds n(n-1)/2");
    out->n = n;
    GenER p{n};
    return generate(
        [&](hipStream_t st, uint64_t M) {
            return [=](uint64_t *key, uint32_t *idx) {
                hipLaunchKernelGGL(k_cand_er, dim3(gridn(M)), dim3(256), 0, st, p, seed, M, key, idx);
            };
        },
        m, 1.05, out);
}

extern "C" int zg_gen_chung_lu(uint32_t n, uint64_t m, double beta, double max_deg, uint64_t seed, zg_edges *out)
{
    if (!out || n < 2 || !(beta > 1.0) || !(max_deg >= 1.0)) return set_error(ZG_EINVAL, "bad arguments");
    int rc = ensure_init();
    if (rc) return rc;
    out->n = n;
    const double alpha = 1.0 / (beta - 1.0);
    // choose i0 so that 2m * w_0 / sum(w) = max_deg (bisection on log i0)
    auto top_share = [&](double i0) {
        double s = 0;
        for (uint32_t i = 0; i < n; ++i) s += std::pow(i + i0, -alpha);
        return std::pow(i0, -alpha) / s;
    };
    const double target = max_deg / (2.0 * (double)m);
    double lo = 1e-3, hi = (double)n;
    for (int it = 0; it < 60; ++it) {
        const double mid = std::sqrt(lo * hi);
        if (top_share(mid) > target) lo = mid; else hi = mid;
    }
    const double i0 = std::sqrt(lo * hi);
    // Vose alias table over w_i = (i + i0)^-alpha
    std::vector<double> w(n), prob(n);
    std::vector<uint32_t> alias(n);
    double s = 0;
    for (uint32_t i = 0; i < n; ++i) { w[i] = std::pow(i + i0, -alpha); s += w[i]; }
    std::vector<uint32_t> small, large;
    small.reserve(n); large.reserve(n);
    for (uint32_t i = 0; i < n; ++i) {
        w[i] = w[i] * n / s;
        (w[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        const uint32_t l = small.back(); small.pop_back();
        const uint32_t g = large.back();
        prob[l] = w[l]; alias[l] = g;
        w[g] = (w[g] + w[l]) - 1.0;
        if (w[g] < 1.0) { large.pop_back(); small.push_back(g); }
    }
    for (uint32_t g : large) { prob[g] = 1.0; alias[g] = g; }
    for (uint32_t l : small) { prob[l] = 1.0; alias[l] = l; }
    DB<double> dprob;
    DB<uint32_t> dalias;
    GALLOC(dprob, n); GALLOC(dalias, n);
    ZG_HIP(hipMemcpy(dprob.p, prob.data(), 8ull * n, hipMemcpyHostToDevice));
    ZG_HIP(hipMemcpy(dalias.p, alias.data(), 4ull * n, hipMemcpyHostToDevice));
    GenCL p{n, dprob.p, dalias.p};
    return generate(
        [&](hipStream_t st, uint64_t M) {
            return [=](uint64_t *key, uint32_t *idx) {
                hipLaunchKernelGGL(k_cand_cl, dim3(gridn(M)), dim3(256), 0, st, p, seed, M, key, idx);
            };
        },
        m, 1.3, out);
}

extern "C" int zg_gen_rmat(int scale, uint64_t m, double a, double b, double c, uint64_t seed, zg_edges *out)
{
    if (!out || scale < 1 || scale > 31 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return set_error(ZG_EINVAL, "bad arguments");
    out->n = 1u << scale;
    GenRMAT p{scale, a, a + b, a + b + c};
    return generate(
        [&](hipStream_t st, uint64_t M) {
            return [=](uint64_t *key, uint32_t *idx) {
                hipLaunchKernelGGL(k_cand_rmat, dim3(gridn(M)), dim3(256), 0, st, p, seed, M, key, idx);
            };
        },
        m, 1.25, out);
}

extern "C" int zg_edges_add_uniform_weights(zg_edges *e, uint64_t seed)
{
    if (!e) return set_error(ZG_EINVAL, "bad arguments");
    if (e->d_w) ZG_HIP(hipFree(e->d_w));
    e->d_w = nullptr;
    ZG_HIP(hipMalloc(&e->d_w, 4 * (e->m ? e->m : 1)));
    if (e->m) hipLaunchKernelGGL(k_weights, dim3(gridn(e->m)), dim3(256), 0, 0, e->m, seed, e->d_w);
    ZG_HIP(hipGetLastError());
    ZG_HIP(hipDeviceSynchronize());
    return ZG_OK;
}

extern "C" int zg_edges_download(const zg_edges *e, uint32_t *src, uint32_t *dst, float *w)
{
    if (!e) return set_error(ZG_EINVAL, "bad arguments");
    if (src && e->m) ZG_HIP(hipMemcpy(src, e->d_src, 4 * e->m, hipMemcpyDeviceToHost));
    if (dst && e->m) ZG_HIP(hipMemcpy(dst, e->d_dst, 4 * e->m, hipMemcpyDeviceToHost));
    if (w && e->m) {
        if (!e->d_w) return set_error(ZG_EINVAL, "edges have no weights");
        ZG_HIP(hipMemcpy(w, e->d_w, 4 * e->m, hipMemcpyDeviceToHost));
    }
    return ZG_OK;
}

extern "C" int zg_edges_free(zg_edges *e)
{
    if (!e) return ZG_OK;
    if (e->d_src) (void)hipFree(e->d_src);
    if (e->d_dst) (void)hipFree(e->d_dst);
    if (e->d_w) (void)hipFree(e->d_w);
    e->d_src = e->d_dst = nullptr;
    e->d_w = nullptr;
    e->m = 0;
    return ZG_OK;
}
