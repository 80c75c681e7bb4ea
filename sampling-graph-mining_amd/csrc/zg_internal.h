// zg_internal.h -- host-side structures shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "zg_device.h"
#include "../../include/zgraph_amd.h"

namespace zg {

// per-channel reduction record (block partials and final result)
struct Acc {
    uint64_t dig[5];   // sum of 32-bit digits of value * 2^(K - log2den)
    uint64_t succ;
    double sq;
    uint64_t ovf;
};
static_assert(sizeof(Acc) == 64, "Acc layout");
static_assert(sizeof(ERec) == 32, "edge record is 32 bytes");
static_assert(sizeof(zg_outcome) == 24, "zg_outcome is a 24-byte record");

struct PlanView {
    int kind, order, pk, pe, K, nch;
    const uint32_t *accept;   // ZG_PLAN_PATTERN: bit table over relabeled edge masks
    int K1;                   // channel 1's denominator exponent (two-channel plans)
};

// one query of a batched launch (device descriptor)
struct BQuery {
    GView g;
    PlanView P;
    uint64_t seed, first, count;
    uint64_t chunk0;   // index of the query's first chunk in the launch
    uint64_t chunk;    // estimators per chunk (a multiple of 64)
    uint64_t rec0;     // first chunk record (P.nch records per chunk)
    uint32_t res0;     // first result record (P.nch records)
    int kk;            // kernel kind: the plan kind, or 100 + k for k-vertex patterns
};

}  // namespace zg

struct zg_graph {
    uint32_t n = 0;
    uint64_t m = 0, num_edges = 0, max_degree = 0;
    zg::ERec *edges = nullptr;
    uint64_t *off = nullptr, *adj = nullptr, *sadj = nullptr;
    float *w = nullptr;
    zg::HBucket *htab = nullptr;
    uint32_t hnb = 0;
    bool simple = true;           // no parallel edges (set by the build)
    uint64_t bytes = 0;
    hipStream_t stream = nullptr;
    zg::Acc *partials = nullptr;  // scratch for block partials
    size_t partials_cap = 0;
    zg::Acc *result = nullptr;    // device result (2 channels)
    zg_outcome *outbuf = nullptr;
    size_t outbuf_cap = 0;
    uint32_t *ord[4] = {nullptr, nullptr, nullptr, nullptr};   // first-edge ordering: keys, offsets (x2)
    size_t ord_cap = 0;
    void *ord_tmp = nullptr;
    size_t ord_tmp_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    zg::GView view() const { return zg::GView{edges, off, adj, sadj, htab, m, n, hnb, simple ? 1u : 0u, 0u, nullptr}; }
};

struct zg_plan {
    zg_plan_desc desc;
    int nch = 1, K = 0;
    int K1 = -1;              // channel 1's exponent when it differs from channel 0's
    double cond_calls = 0;
    int nuniq = 0, ntotal = 0;
    uint32_t *d_accept = nullptr;
    zg::PlanView view() const
    {
        return zg::PlanView{desc.kind, desc.order, desc.pattern_vertices, desc.pattern_edges, K, nch, d_accept,
                            K1 < 0 ? K : K1};
    }
    int channel_k(int ch) const { return ch == 1 && K1 >= 0 ? K1 : K; }
};

namespace zg {

// error plumbing (thread-local message)
int set_error(int code, const std::string &msg);
int hip_error(hipError_t e, const char *what);
#define ZG_HIP(call)                                        \
    do {                                                    \
        hipError_t _e = (call);                             \
        if (_e != hipSuccess) return zg::hip_error(_e, #call); \
    } while (0)

int ensure_init();
int device_cus();

// graph construction from device edge arrays (zg_graph.hip)
int build_graph(const uint32_t *d_src, const uint32_t *d_dst, const float *d_w, uint64_t m, uint32_t n,
                uint64_t seed, zg_graph **out);
int subsample_graph(const zg_graph *in, double rate, uint64_t seed, zg_graph **out);
int filter_graph(const zg_graph *in, float thr, int keep_ge, uint64_t seed, zg_graph **out);
void free_graph(zg_graph *g);

// estimator launch (zg_kernels.hip): enqueues on g->stream, results in g->result
// order: 0 automatic, 1 estimator-id order, 2..16 first-edge order with 2^order buckets
// ext_partials / ext_result (both or neither): caller-provided block partials
// (ext_cap records) and result record instead of the handle's own buffers; the
// launch then allocates nothing and fails if its grid needs more than ext_cap
int launch_run(zg_graph *g, const zg_plan *p, uint64_t seed, uint64_t first, uint64_t count,
               zg_outcome *d_out, int block, int blocks_per_cu, unsigned long long *reads = nullptr,
               int order = 1, Acc *ext_partials = nullptr, size_t ext_cap = 0, Acc *ext_result = nullptr);
// launches of at most this many estimators when they run in first-edge order
constexpr uint64_t kMaxOrderedLaunch = 1ull << 28;
int resolve_order(const zg_graph *g, const zg_plan *p, uint64_t count, int order);
void finish_sums(const Acc *host_acc, int nch, int K0, int K1, uint64_t count, zg_sums *out);
// a batch's launches (queries in chunk order): counter reset, the work-queue
// kernel over every query of every plan kind, the per-query reduction
int launch_batch(const BQuery *d_q, int nq, uint64_t nchunks, unsigned long long *counter, Acc *chunk_recs,
                 Acc *result, int grid, int block, hipStream_t st);

}  // namespace zg
