// zg_api.hip -- the C ABI (include/zgraph_amd.h): error plumbing, device
// initialisation, graph handles, plans, runs and the regression
// error-latency profile.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "zg_internal.h"

namespace zg {

static thread_local std::string g_err;
static std::mutex g_init_mu;
static int g_inited = 0, g_cus = 0, g_lds = 0, g_warp = 0;

// Device scratch of zg_run_noalloc: kNoAllocPartials block partial records
// followed by the two-channel result record, in a file-scope device array
// (each GPU holds its own copy until the process ends; about 1 MiB).  Host
// code must not pass `scratch` itself to a kernel: its device address is
// looked up once per device with hipGetSymbolAddress (in ensure_init, with
// the device current) and kept in g_scratch_dev[device].
constexpr size_t kNoAllocPartials = 16384;
constexpr size_t kNoAllocRecords = kNoAllocPartials + 2;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) Acc scratch[kNoAllocRecords];
static Acc *g_scratch_dev[kMaxDevices];

int set_error(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hip_error(hipError_t e, const char *what)
{
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? ZG_ENOMEM : ZG_EHIP;
}

int ensure_init()
{
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_inited) return ZG_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return set_error(ZG_ESTATE, "no HIP device available");
    ZG_HIP(hipSetDevice(0));
    hipDeviceProp_t prop;
    ZG_HIP(hipGetDeviceProperties(&prop, 0));
    g_cus = prop.multiProcessorCount;
    g_lds = (int)prop.maxSharedMemoryPerMultiProcessor;
    g_warp = prop.warpSize;
    if (g_warp != 64) return set_error(ZG_ESTATE, "expected 64-wide wavefronts (CDNA)");
    // the library's handles all live on device 0, the current device here
    void *sp = nullptr;
    ZG_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
    if (!sp) return set_error(ZG_ESTATE, "no device address for the no-allocation scratch");
    g_scratch_dev[0] = static_cast<Acc *>(sp);
    g_inited = 1;
    return ZG_OK;
}

int device_cus() { return g_cus > 0 ? g_cus : 256; }

}  // namespace zg

using namespace zg;

// ---------------------------------------------------------------------------
extern "C" int zg_last_error(char *buf, size_t len)
{
    if (!buf || !len) return ZG_EINVAL;
    const size_t k = std::min(len - 1, g_err.size());
    memcpy(buf, g_err.data(), k);
    buf[k] = 0;
    return ZG_OK;
}

extern "C" int zg_init(void) { return ensure_init(); }

extern "C" int zg_device_info(int *num_cus, int *lds_per_cu, int *warp_size)
{
    int rc = ensure_init();
    if (rc) return rc;
    if (num_cus) *num_cus = g_cus;
    if (lds_per_cu) *lds_per_cu = g_lds;
    if (warp_size) *warp_size = g_warp;
    return ZG_OK;
}

extern "C" int zg_device_memory(uint64_t *free_bytes, uint64_t *total_bytes)
{
    int rc = ensure_init();
    if (rc) return rc;
    size_t f = 0, t = 0;
    ZG_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return ZG_OK;
}

// ---- graphs ----------------------------------------------------------------
extern "C" int zg_graph_create(const uint32_t *src, const uint32_t *dst, const float *weights, uint64_t m,
                               uint32_t n, uint64_t shuffle_seed, zg_graph **out)
{
    if (!out || (m && (!src || !dst))) return set_error(ZG_EINVAL, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    uint32_t *ds = nullptr, *dd = nullptr;
    float *dw = nullptr;
    auto cleanup = [&]() {
        if (ds) (void)hipFree(ds);
        if (dd) (void)hipFree(dd);
        if (dw) (void)hipFree(dw);
    };
    if (hipMalloc(&ds, 4 * (m ? m : 1)) != hipSuccess || hipMalloc(&dd, 4 * (m ? m : 1)) != hipSuccess ||
        (weights && hipMalloc(&dw, 4 * (m ? m : 1)) != hipSuccess)) {
        cleanup();
        return set_error(ZG_ENOMEM, "device allocation failed");
    }
    hipError_t e = hipSuccess;
    if (m) {
        e = hipMemcpy(ds, src, 4 * m, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dd, dst, 4 * m, hipMemcpyHostToDevice);
        if (e == hipSuccess && weights) e = hipMemcpy(dw, weights, 4 * m, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        cleanup();
        return hip_error(e, "hipMemcpy (edge upload)");
    }
    rc = build_graph(ds, dd, dw, m, n, shuffle_seed, out);
    cleanup();
    return rc;
}

extern "C" int zg_graph_create_device(const uint32_t *d_src, const uint32_t *d_dst, const float *d_weights,
                                      uint64_t m, uint32_t n, uint64_t shuffle_seed, zg_graph **out)
{
    if (!out || (m && (!d_src || !d_dst))) return set_error(ZG_EINVAL, "null argument");
    int rc = ensure_init();
    if (rc) return rc;
    return build_graph(d_src, d_dst, d_weights, m, n, shuffle_seed, out);
}

extern "C" int zg_graph_destroy(zg_graph *g)
{
    if (g) {
        (void)hipStreamSynchronize(g->stream);
        free_graph(g);
    }
    return ZG_OK;
}

extern "C" int zg_graph_info_get(const zg_graph *g, zg_graph_info *out)
{
    if (!g || !out) return set_error(ZG_EINVAL, "null argument");
    out->n = g->n;
    out->m = g->m;
    out->num_edges = g->num_edges;
    out->max_degree = g->max_degree;
    out->device_bytes = g->bytes;
    out->has_weights = g->w != nullptr;
    out->has_parallel_edges = g->simple ? 0 : 1;
    return ZG_OK;
}

extern "C" int zg_graph_export(const zg_graph *g, uint32_t *src, uint32_t *dst, float *w, uint64_t *off,
                               uint32_t *adj_rank, uint32_t *adj_nbr)
{
    if (!g) return set_error(ZG_EINVAL, "null graph");
    const uint64_t m = g->m, n2 = 2 * m;
    if (src || dst) {
        std::vector<zg::ERec> e(m);
        if (m) ZG_HIP(hipMemcpy(e.data(), g->edges, sizeof(zg::ERec) * m, hipMemcpyDeviceToHost));
        for (uint64_t r = 0; r < m; ++r) {
            if (src) src[r] = e[r].src;
            if (dst) dst[r] = e[r].dst;
        }
    }
    if (w) {
        if (!g->w) return set_error(ZG_EINVAL, "graph has no weights");
        if (m) ZG_HIP(hipMemcpy(w, g->w, 4 * m, hipMemcpyDeviceToHost));
    }
    if (off) ZG_HIP(hipMemcpy(off, g->off, 8ull * (g->n + 1), hipMemcpyDeviceToHost));
    if (adj_rank || adj_nbr) {
        std::vector<uint64_t> a(n2);
        if (n2) ZG_HIP(hipMemcpy(a.data(), g->adj, 8 * n2, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < n2; ++i) {
            if (adj_rank) adj_rank[i] = (uint32_t)(a[i] >> 32);
            if (adj_nbr) adj_nbr[i] = (uint32_t)a[i];
        }
    }
    return ZG_OK;
}

extern "C" int zg_graph_subsample(const zg_graph *in, double rate, uint64_t seed, zg_graph **out)
{
    if (!in || !out) return set_error(ZG_EINVAL, "null argument");
    return subsample_graph(in, rate, seed, out);
}

extern "C" int zg_graph_filter(const zg_graph *in, float threshold, int keep_ge, uint64_t shuffle_seed,
                               zg_graph **out)
{
    if (!in || !out) return set_error(ZG_EINVAL, "null argument");
    return filter_graph(in, threshold, keep_ge, shuffle_seed, out);
}

// ---- plans -----------------------------------------------------------------
namespace {

struct PatEdges {
    std::vector<std::pair<uint32_t, uint32_t>> e;
};

bool pat_has(const PatEdges &p, uint32_t a, uint32_t b)
{
    for (auto &x : p.e)
        if ((x.first == a && x.second == b) || (x.first == b && x.second == a)) return true;
    return false;
}

// all k! relabelings in generate_mapping's DFS order (src/SamplerGenerator.cpp:262-284)
void perms(int k, int idx, std::vector<uint32_t> &map, std::vector<int> &used, const PatEdges &base,
           std::vector<PatEdges> &out)
{
    if (idx == k) {
        PatEdges p;
        for (auto &x : base.e) p.e.push_back({map[x.first], map[x.second]});
        out.push_back(p);
        return;
    }
    for (int v = 0; v < k; ++v)
        if (!used[v]) {
            used[v] = 1;
            map[idx] = (uint32_t)v;
            perms(k, idx + 1, map, used, base, out);
            used[v] = 0;
        }
}

inline int pair_bit(uint32_t a, uint32_t b)
{
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return (int)(hi * (hi - 1) / 2 + lo);
}

// unique relabelings of a pattern (SamplerGenerator ctor, src/SamplerGenerator.cpp:311-363):
// the edge masks of the relabelings that survive Pattern::is_automorphic (:225-242)
// deduplication, in generate_mapping order.  Host only.
int pattern_classes(const zg_plan_desc *d, std::vector<uint32_t> &uniq_masks, int *ntotal)
{
    const int k = d->pattern_vertices, E = d->pattern_edges;
    if (k < 2 || k > 6) return set_error(ZG_EINVAL, "pattern vertices must be in [2, 6]");
    if (E < 1 || E > ZG_MAX_PATTERN_EDGES) return set_error(ZG_EINVAL, "pattern edges must be in [1, 15]");
    if (E > k * (k - 1) / 2) return set_error(ZG_EINVAL, "more pattern edges than vertex pairs");
    PatEdges base;
    for (int i = 0; i < E; ++i) {
        const uint32_t a = d->pattern_src[i], b = d->pattern_dst[i];
        if ((int)a >= k || (int)b >= k || a == b) return set_error(ZG_EINVAL, "bad pattern edge");
        if (pat_has(base, a, b)) return set_error(ZG_EINVAL, "repeated pattern edge");
        base.e.push_back({a, b});
    }
    std::vector<PatEdges> all;
    std::vector<uint32_t> map(k);
    std::vector<int> used(k, 0);
    perms(k, 0, map, used, base, all);
    std::vector<char> dup(all.size(), 0);
    uniq_masks.clear();
    for (size_t i = 0; i < all.size(); ++i) {
        for (size_t j = 0; j < i; ++j) {
            if (dup[j]) continue;
            bool same = true;
            for (auto &x : all[i].e) same = same && pat_has(all[j], x.first, x.second);
            if (same) { dup[i] = 1; break; }
        }
        if (!dup[i]) {
            uint32_t mk = 0;
            for (auto &x : all[i].e) mk |= 1u << pair_bit(x.first, x.second);
            uniq_masks.push_back(mk);
        }
    }
    *ntotal = (int)all.size();
    return ZG_OK;
}

}  // namespace

extern "C" int zg_pattern_unique(const zg_plan_desc *d, int *unique, int *total)
{
    if (!d) return set_error(ZG_EINVAL, "null argument");
    std::vector<uint32_t> masks;
    int nt = 0;
    int rc = pattern_classes(d, masks, &nt);
    if (rc) return rc;
    if (unique) *unique = (int)masks.size();
    if (total) *total = nt;
    return ZG_OK;
}

extern "C" int zg_plan_create(const zg_plan_desc *d, zg_plan **out)
{
    if (!d || !out) return set_error(ZG_EINVAL, "null argument");
    std::unique_ptr<zg_plan> p(new zg_plan());
    p->desc = *d;
    switch (d->kind) {
    case ZG_PLAN_TRIANGLE: case ZG_PLAN_THREE_CHAIN: p->K = 1; p->cond_calls = 1; break;
    case ZG_PLAN_THREE_MOTIF: p->K = 1; p->cond_calls = 1; p->nch = 2; break;
    case ZG_PLAN_CHAIN:
        if (d->order < 3 || d->order > 9) return set_error(ZG_EINVAL, "chain order must be in [3, 9]");
        p->K = d->order - 2; p->cond_calls = d->order - 2;
        break;
    case ZG_PLAN_FOUR_CHAIN: p->K = 2; p->cond_calls = 2; break;
    case ZG_PLAN_K4_CLOSE: p->K = 0; p->cond_calls = 2; break;
    // each channel keeps its own plan's exponent, so its sums equal that plan's word for word
    case ZG_PLAN_TRIANGLE_K4: p->K = 1; p->K1 = 0; p->cond_calls = 2; p->nch = 2; break;
    case ZG_PLAN_K4_SPLIT: case ZG_PLAN_TRIANGLE_K4_SPLIT: case ZG_PLAN_HOUSE_SPLIT: {
        const int J = d->completions;
        if (J < 1 || J > 4096 || (J & (J - 1)))
            return set_error(ZG_EINVAL, "completions must be a power of two in [1, 4096]");
        const int lg = __builtin_ctz((unsigned)J);   // every split success has denominator J
        p->cond_calls = d->kind == ZG_PLAN_HOUSE_SPLIT ? 3 : 2;   // house: the roof, x and y
        if (d->kind != ZG_PLAN_TRIANGLE_K4_SPLIT) {
            p->K = lg;
        } else {
            p->K = 1;   // the triangle channel keeps the triangle plan's exponent
            p->K1 = lg;
            p->nch = 2;
        }
        break;
    }
    case ZG_PLAN_PATTERN: {
        const int k = d->pattern_vertices, E = d->pattern_edges;
        std::vector<uint32_t> uniq_masks;
        int rc = pattern_classes(d, uniq_masks, &p->ntotal);
        if (rc) return rc;
        p->nuniq = (int)uniq_masks.size();
        const int nbits = k * (k - 1) / 2;
        std::vector<uint32_t> table(((1u << nbits) + 31) / 32, 0);
        for (uint32_t P : uniq_masks)
            for (uint32_t s = P;; s = (s - 1) & P) {   // every sub-mask of P
                table[s >> 5] |= 1u << (s & 31);
                if (!s) break;
            }
        rc = ensure_init();
        if (rc) return rc;
        ZG_HIP(hipMalloc(&p->d_accept, 4 * table.size()));
        ZG_HIP(hipMemcpy(p->d_accept, table.data(), 4 * table.size(), hipMemcpyHostToDevice));
        p->K = E - 1;
        p->cond_calls = E - 1;   // src/SamplerGenerator.cpp:471-474
        break;
    }
    default: return set_error(ZG_EINVAL, "unknown plan kind");
    }
    *out = p.release();
    return ZG_OK;
}

extern "C" int zg_plan_destroy(zg_plan *p)
{
    if (p) {
        if (p->d_accept) (void)hipFree(p->d_accept);
        delete p;
    }
    return ZG_OK;
}

extern "C" int zg_plan_channels(const zg_plan *p) { return p ? p->nch : 0; }
extern "C" double zg_plan_cond_calls(const zg_plan *p) { return p ? p->cond_calls : 0.0; }
extern "C" int zg_plan_unique_patterns(const zg_plan *p, int *unique, int *total)
{
    if (!p || p->desc.kind != ZG_PLAN_PATTERN) return set_error(ZG_EINVAL, "not a pattern plan");
    if (unique) *unique = p->nuniq;
    if (total) *total = p->ntotal;
    return ZG_OK;
}

// ---- runs ------------------------------------------------------------------
namespace {

constexpr uint64_t kMaxPerLaunch = 1ull << 31;   // digit sums stay below 2^64

struct HostAcc {
    unsigned __int128 dig[5];
    uint64_t succ, ovf;
    double sq;
};

// ext (may be null): device scratch of kNoAllocRecords records that replaces
// the handle's partials and result buffers, so that nothing is allocated
int run_chunks(zg_graph *g, const zg_plan *p, uint64_t seed, uint64_t first, uint64_t count, zg_outcome *d_out,
               const zg_launch *launch, zg_sums *sums, float *kernel_ms, Acc *ext = nullptr)
{
    if (count > ~0ull - first) return set_error(ZG_EINVAL, "estimator ids: first_id + count must not exceed 2^64 - 1");
    const int nch = p->nch;
    HostAcc h[2];
    memset(h, 0, sizeof(h));
    const int block = launch ? launch->block_size : 0;
    const int bpc = launch ? launch->blocks_per_cu : 0;
    const int order = d_out ? 1 : resolve_order(g, p, count, launch ? launch->order : 0);
    const uint64_t per_launch = order >= 2 ? kMaxOrderedLaunch : kMaxPerLaunch;
    ZG_HIP(hipEventRecord(g->ev0, g->stream));
    for (uint64_t done = 0; done < count || (count == 0 && done == 0);) {
        const uint64_t c = std::min(per_launch, count - done);
        int rc = launch_run(g, p, seed, first + done, c, d_out ? d_out + done * nch : nullptr, block, bpc, nullptr,
                            order, ext, ext ? kNoAllocPartials : 0, ext ? ext + kNoAllocPartials : nullptr);
        if (rc) return rc;
        Acc r[2];
        ZG_HIP(hipMemcpyAsync(r, ext ? ext + kNoAllocPartials : g->result, sizeof(Acc) * nch, hipMemcpyDeviceToHost,
                              g->stream));
        ZG_HIP(hipStreamSynchronize(g->stream));
        for (int ch = 0; ch < nch; ++ch) {
            for (int i = 0; i < 5; ++i) h[ch].dig[i] += r[ch].dig[i];
            h[ch].succ += r[ch].succ;
            h[ch].ovf += r[ch].ovf;
            h[ch].sq += r[ch].sq;
        }
        done += c;
        if (count == 0) break;
    }
    ZG_HIP(hipEventRecord(g->ev1, g->stream));
    ZG_HIP(hipEventSynchronize(g->ev1));
    if (kernel_ms) ZG_HIP(hipEventElapsedTime(kernel_ms, g->ev0, g->ev1));
    for (int ch = 0; ch < nch; ++ch) {
        // fold the 128-bit digit sums into 64-bit digits for finish_sums
        Acc a;
        unsigned __int128 carry = 0;
        uint64_t limbs32[7];
        for (int i = 0; i < 7; ++i) {
            unsigned __int128 t = carry + (i < 5 ? h[ch].dig[i] : 0);
            limbs32[i] = (uint64_t)(uint32_t)t;
            carry = t >> 32;
        }
        // re-express as digits < 2^32 (limbs32[5], [6] fold into dig[4])
        for (int i = 0; i < 4; ++i) a.dig[i] = limbs32[i];
        a.dig[4] = limbs32[4] | (limbs32[5] << 32);
        if (limbs32[6] || carry) return set_error(ZG_EOVERFLOW, "estimate sum exceeds 192 bits");
        a.succ = h[ch].succ;
        a.ovf = h[ch].ovf;
        a.sq = h[ch].sq;
        const int k = p->channel_k(ch);
        finish_sums(&a, 1, k, k, count, &sums[ch]);
    }
    return ZG_OK;
}

}  // namespace

extern "C" int zg_run(zg_graph *g, const zg_plan *p, uint64_t seed, uint64_t first_id, uint64_t count,
                      const zg_launch *launch, zg_sums *sums, float *kernel_ms)
{
    if (!g || !p || !sums) return set_error(ZG_EINVAL, "null argument");
    int rc = run_chunks(g, p, seed, first_id, count, nullptr, launch, sums, kernel_ms);
    if (rc) return rc;
    for (int c = 0; c < p->nch; ++c)
        if (sums[c].status) return set_error(ZG_EOVERFLOW, "an estimator value exceeded 128 bits");
    return ZG_OK;
}

extern "C" int zg_run_noalloc(zg_graph *g, const zg_plan *p, uint64_t seed, uint64_t first_id, uint64_t count,
                              const zg_launch *launch, zg_sums *sums, float *kernel_ms)
{
    if (!g || !p || !sums) return set_error(ZG_EINVAL, "null argument");
    if (launch && launch->order >= 2)
        return set_error(ZG_EINVAL, "zg_run_noalloc runs in estimator-id order only (order 0 or 1)");
    int dev = 0;
    ZG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices || !g_scratch_dev[dev])
        return set_error(ZG_ESTATE, "zg_run_noalloc: no scratch on the current device (the library runs on device 0)");
    int rc = run_chunks(g, p, seed, first_id, count, nullptr, launch, sums, kernel_ms, g_scratch_dev[dev]);
    if (rc) return rc;
    for (int c = 0; c < p->nch; ++c)
        if (sums[c].status) return set_error(ZG_EOVERFLOW, "an estimator value exceeded 128 bits");
    return ZG_OK;
}

extern "C" int zg_run_outcomes(zg_graph *g, const zg_plan *p, uint64_t seed, uint64_t first_id, uint64_t count,
                               zg_outcome *host_out, zg_sums *sums)
{
    if (!g || !p || !host_out) return set_error(ZG_EINVAL, "null argument");
    const uint64_t need = count * p->nch;
    if (need > g->outbuf_cap) {
        if (g->outbuf) ZG_HIP(hipFree(g->outbuf));
        g->outbuf = nullptr;
        g->outbuf_cap = 0;
        ZG_HIP(hipMalloc(&g->outbuf, sizeof(zg_outcome) * (need ? need : 1)));
        g->outbuf_cap = need;
    }
    zg_sums tmp[2];
    int rc = run_chunks(g, p, seed, first_id, count, g->outbuf, nullptr, sums ? sums : tmp, nullptr);
    if (rc) return rc;
    if (need) ZG_HIP(hipMemcpy(host_out, g->outbuf, sizeof(zg_outcome) * need, hipMemcpyDeviceToHost));
    return ZG_OK;
}

extern "C" int zg_count_reads(zg_graph *g, const zg_plan *p, uint64_t seed, uint64_t first_id, uint64_t count,
                              uint64_t *reads)
{
    if (!g || !p || !reads) return set_error(ZG_EINVAL, "null argument");
    if (count > kMaxPerLaunch) return set_error(ZG_EINVAL, "count too large for one counting launch");
    if (count > ~0ull - first_id) return set_error(ZG_EINVAL, "estimator ids: first_id + count must not exceed 2^64 - 1");
    unsigned long long *d = nullptr;
    ZG_HIP(hipMalloc(&d, 8));
    int rc = ZG_OK;
    hipError_t e = hipMemsetAsync(d, 0, 8, g->stream);
    if (e == hipSuccess) rc = launch_run(g, p, seed, first_id, count, nullptr, 0, 0, d);
    else rc = hip_error(e, "hipMemsetAsync");
    unsigned long long h = 0;
    if (!rc) {
        e = hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, g->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
        if (e != hipSuccess) rc = hip_error(e, "read count copy");
    }
    (void)hipFree(d);
    *reads = h;
    return rc;
}

// ---- regression error-latency profile --------------------------------------
extern "C" long double zg_regression_k(const long double *x, const long double *y, unsigned len)
{   // UnbiasedLinearRegression::solve (src/Regression.cpp:17-28)
    long double t0 = 0, t1 = 0;
    for (unsigned i = 0; i < len; ++i) t0 += x[i] * y[i];
    for (unsigned i = 0; i < len; ++i) t1 += x[i] * x[i];
    return t0 / t1;
}

extern "C" double zg_elp_bound(double eps, double delta, double a, double b, double est, uint64_t max_degree,
                               uint64_t num_edges)
{   // ErrorProfile::calculate_bound (src/ErrorLatencyProfile.cpp:51-59)
    const double f0 = 1. / pow(eps, 2.);
    const double f1 = pow((double)max_degree, b) * pow((double)num_edges, a) / est;
    const double f2 = log(2. / delta);
    return f0 * f1 * f2;
}

extern "C" double zg_elp_from_pilots(const long double *est, uint64_t npts, long double F, double a, double b,
                                     uint64_t samp_maxdeg, uint64_t samp_m, uint64_t maxdeg, uint64_t m, double eps,
                                     double conf, long double *K_out)
{   // ErrorProfile::get_num_estimators (src/ErrorLatencyProfile.cpp:127-200)
    const unsigned long long per = 16;
    std::vector<long double> X(npts ? npts : 1), Y(npts ? npts : 1);
    long double ema = 100;   // "a very large number" (:128)
    const double error_times = 1, alpha = 0.04;
    const double Fd = (double)F;
    for (uint64_t i = 0; i < npts; ++i) {
        long double err = (est[i] - Fd) / Fd * 5;
        if (err < 0) err = -err;
        ema = ema * (1 - alpha) + err * alpha;
        const double xt = (double)((long double)zg_elp_bound((double)(ema * error_times), 1 - conf, a, b, Fd,
                                                             samp_maxdeg, samp_m) *
                                   powl(ema * error_times, 2));
        X[i] = sqrt(xt / (double)((i + 1) * per));
        Y[i] = ema;
    }
    const double K = (double)zg_regression_k(X.data(), Y.data(), (unsigned)npts);
    const long double rough = (long double)Fd / powl(0.1L, (long double)(a + b));
    double N = zg_elp_bound(eps, 1. - conf, a, b, (double)rough, maxdeg, m);
    N *= (K * K);
    if (K_out) *K_out = K;
    return N;
}

extern "C" int zg_elp_regression(zg_graph *g, const zg_plan *p, double eps, double conf, uint64_t seed,
                                 int pilot_scale, zg_elp_result *out)
{
    if (!g || !p || !out) return set_error(ZG_EINVAL, "null argument");
    if (p->nch != 1) return set_error(ZG_EINVAL, "the ELP sizes single-channel plans");
    if (pilot_scale < 0 || pilot_scale > 10) return set_error(ZG_EINVAL, "pilot_scale must be in [0, 10]");
    const auto t0 = std::chrono::steady_clock::now();
    // get_num_conditional_sample_edge_calls (:470-488) runs 1024 estimates and
    // returns NaN when none succeeds, which makes the reference's budget NaN.
    // Here b is a constant of the plan and the pilot only guards that case; it
    // widens (x32, up to 2^20 estimates) before reporting it, since 1024
    // estimates can all fail on large sparse graphs.
    zg_sums s1;
    int rc = ZG_OK;
    for (uint64_t pil = 1024;; pil <<= 5) {
        rc = zg_run(g, p, seed ^ 0x9E3779B97F4A7C15ull, 0, pil, nullptr, &s1, nullptr);
        if (rc) return rc;
        if (s1.successes) break;
        if (pil >= (1ull << 20))
            return set_error(ZG_ESTATE, "no successful estimator among 2^20 pilots (the reference's budget is NaN)");
    }
    const double a = 1.0, b = p->cond_calls;
    zg_graph *sg = nullptr;
    rc = subsample_graph(g, 0.1, seed, &sg);   // sample_rate 0.1 (:83)
    if (rc) return rc;
    struct G { zg_graph *g; ~G() { if (g) free_graph(g); } } sgg{sg};
    const uint64_t npts = (1ull << 14) >> pilot_scale, per = 16, post = (1ull << 23) >> pilot_scale;
    std::vector<zg_outcome> o(npts * per);
    rc = zg_run_outcomes(sg, p, seed, 0, npts * per, o.data(), nullptr);
    if (rc) return rc;
    // running means after each batch of 16 (Estimation::update, src/ZGraphInstance.cpp:24-27)
    std::vector<long double> est(npts);
    unsigned __int128 acc = 0;
    const long double scale = ldexpl(1.0L, -p->K);
    for (uint64_t i = 0; i < npts; ++i) {
        for (uint64_t j = 0; j < per; ++j) {
            const zg_outcome &r = o[i * per + j];
            if (r.status != 1) continue;
            const unsigned __int128 v = ((unsigned __int128)r.num_hi << 64) | r.num_lo;
            acc += v << (p->K - r.log2den);
        }
        est[i] = (long double)acc * scale / (long double)((i + 1) * per);
    }
    zg_sums s2;
    rc = zg_run(sg, p, seed, npts * per, post, nullptr, &s2, nullptr);
    if (rc) return rc;
    const long double tot_pilot = (long double)acc * scale;
    const long double tot_post = ((long double)s2.sum[2] * 0x1p128L + (long double)s2.sum[1] * 0x1p64L +
                                  (long double)s2.sum[0]) * scale;
    const long double F = (tot_pilot + tot_post) / (long double)(npts * per + post);
    if (!(F > 0)) return set_error(ZG_ESTATE, "zero estimate on the subsampled graph; the bound is undefined");
    long double K = 0;
    const double N = zg_elp_from_pilots(est.data(), npts, F, a, b, sg->max_degree, sg->num_edges, g->max_degree,
                                        g->num_edges, eps, conf, &K);
    out->num_estimators = N;
    out->K = (double)K;
    out->sampled_estimate = (double)F;
    out->rough_estimate = (double)((long double)(double)F / powl(0.1L, (long double)(a + b)));
    out->sampled_m = sg->num_edges;
    out->sampled_max_degree = sg->max_degree;
    out->elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return ZG_OK;
}
