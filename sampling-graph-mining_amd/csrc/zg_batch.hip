// zg_batch.hip -- many estimator queries (graph, plan, seed, id range) in one
// replayable HIP graph: per plan kind, a counter reset, one persistent
// work-queue launch over all of that kind's queries (zg_kernels.hip,
// k_batch) and a per-query reduction; then one copy of every query's record
// to pinned host memory.  The reference answers one query per
// ZGraphInstance::calculate call (src/ZGraphInstance.cpp:256-297); serving
// many small queries that way is bound by per-call launch and synchronisation
// cost, which the captured graph pays once per batch.
//
// The autotuner picks the launch shape (threads per workgroup, persistent
// workgroups per CU, estimators per claimed chunk) by timing graph replays;
// every shape gives the same exact sums (only the floating-point sum of
// squares can differ in its last bits, as the chunking changes its order).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "zg_internal.h"

using namespace zg;

namespace {
constexpr uint64_t kMaxQueryCount = 1ull << 31;   // digit sums of one query stay below 2^63
}

struct zg_batch {
    struct Group {
        int kind = 0, nch = 1;
        std::vector<int> qidx;          // positions in the caller's query array
        std::vector<BQuery> hq;         // host copies of the descriptors
        BQuery *dq = nullptr;
        uint64_t nchunks = 0;
        unsigned long long *counter = nullptr;
        Acc *partials = nullptr;
        size_t partials_cap = 0;        // in Acc records
        size_t result_off = 0;          // first Acc record of this group in the result buffer
    };
    std::vector<Group> groups;
    std::vector<zg_query> queries;
    std::vector<size_t> rec_off;        // first output record of each query
    size_t nrec = 0;
    zg_batch_config cfg{};
    hipStream_t stream = nullptr;
    Acc *d_result = nullptr, *h_result = nullptr;   // h_result is pinned
    size_t nres = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;

    ~zg_batch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &g : groups) {
            if (g.dq) (void)hipFree(g.dq);
            if (g.counter) (void)hipFree(g.counter);
            if (g.partials) (void)hipFree(g.partials);
        }
        if (d_result) (void)hipFree(d_result);
        if (h_result) (void)hipHostFree(h_result);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int check_config(const zg_batch_config &c)
{
    if (c.block_size % 64 || c.block_size < 64 || c.block_size > 256)
        return set_error(ZG_EINVAL, "block_size must be 64, 128, 192 or 256");
    if (c.blocks_per_cu < 1 || c.blocks_per_cu > 64) return set_error(ZG_EINVAL, "blocks_per_cu must be in [1, 64]");
    if (c.chunk_estimators < 256 || c.chunk_estimators > (1u << 24))
        return set_error(ZG_EINVAL, "chunk_estimators must be in [256, 2^24]");
    return ZG_OK;
}

zg_batch_config defaults(const zg_batch_config *c)
{
    zg_batch_config d{256, 8, 16384};
    if (c) {
        if (c->block_size) d.block_size = c->block_size;
        if (c->blocks_per_cu) d.blocks_per_cu = c->blocks_per_cu;
        if (c->chunk_estimators) d.chunk_estimators = c->chunk_estimators;
    }
    return d;
}

// chunking, descriptor upload and graph capture for the batch's current config
int configure(zg_batch *b)
{
    if (b->exec) { ZG_HIP(hipGraphExecDestroy(b->exec)); b->exec = nullptr; }
    if (b->graph) { ZG_HIP(hipGraphDestroy(b->graph)); b->graph = nullptr; }
    const uint64_t ch = b->cfg.chunk_estimators;
    for (auto &g : b->groups) {
        uint64_t c = 0;
        for (auto &q : g.hq) {
            q.chunk0 = c;
            c += (q.count + ch - 1) / ch;
        }
        g.nchunks = c;
        const size_t need = std::max<size_t>(1, (size_t)c * g.nch);
        if (need > g.partials_cap) {
            if (g.partials) ZG_HIP(hipFree(g.partials));
            g.partials = nullptr;
            if (hipMalloc(&g.partials, sizeof(Acc) * need) != hipSuccess)
                return set_error(ZG_ENOMEM, "batch chunk records: device allocation failed");
            g.partials_cap = need;
        }
        ZG_HIP(hipMemcpyAsync(g.dq, g.hq.data(), sizeof(BQuery) * g.hq.size(), hipMemcpyHostToDevice, b->stream));
    }
    ZG_HIP(hipStreamSynchronize(b->stream));
    const int grid = device_cus() * b->cfg.blocks_per_cu;
    ZG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = ZG_OK;
    for (auto &g : b->groups) {
        rc = launch_batch(g.kind, g.dq, (int)g.hq.size(), g.nchunks, (uint32_t)ch, g.counter, g.partials,
                          b->d_result + g.result_off, grid, b->cfg.block_size, b->stream);
        if (rc) break;
    }
    if (!rc && hipMemcpyAsync(b->h_result, b->d_result, sizeof(Acc) * b->nres, hipMemcpyDeviceToHost, b->stream) !=
                   hipSuccess)
        rc = set_error(ZG_EHIP, "batch: result copy could not be captured");
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(b->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return hip_error(ec, "hipStreamEndCapture");
    b->graph = graph;
    ZG_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return ZG_OK;
}

int replay(zg_batch *b, float *ms)
{
    ZG_HIP(hipEventRecord(b->e0, b->stream));
    ZG_HIP(hipGraphLaunch(b->exec, b->stream));
    ZG_HIP(hipEventRecord(b->e1, b->stream));
    ZG_HIP(hipStreamSynchronize(b->stream));
    if (ms) ZG_HIP(hipEventElapsedTime(ms, b->e0, b->e1));
    return ZG_OK;
}

}  // namespace

extern "C" int zg_batch_create(const zg_query *q, int nq, const zg_batch_config *config, zg_batch **out)
{
    if (!q || nq <= 0 || !out) return set_error(ZG_EINVAL, "need at least one query");
    int rc = ensure_init();
    if (rc) return rc;
    const zg_batch_config cfg = defaults(config);
    rc = check_config(cfg);
    if (rc) return rc;
    std::unique_ptr<zg_batch> b(new zg_batch());
    b->cfg = cfg;
    b->queries.assign(q, q + nq);
    std::map<int, size_t> by_kind;
    for (int i = 0; i < nq; ++i) {
        const zg_query &x = q[i];
        if (!x.graph || !x.plan) return set_error(ZG_EINVAL, "query without graph or plan");
        if (x.count > kMaxQueryCount) return set_error(ZG_EINVAL, "a batched query runs at most 2^31 estimators");
        if (x.count > ~0ull - x.first_id) return set_error(ZG_EINVAL, "estimator ids: first_id + count must not exceed 2^64 - 1");
        if (x.graph->m == 0 && x.count) return set_error(ZG_EINVAL, "query on an empty graph");
        const int kind = x.plan->desc.kind == ZG_PLAN_PATTERN ? 100 + x.plan->desc.pattern_vertices
                                                             : x.plan->desc.kind;   // one kernel per pattern size
        auto it = by_kind.find(kind);
        if (it == by_kind.end()) {
            it = by_kind.emplace(kind, b->groups.size()).first;
            b->groups.emplace_back();
            b->groups.back().kind = kind;
            b->groups.back().nch = x.plan->nch;
        }
        auto &g = b->groups[it->second];
        BQuery d;
        std::memset(&d, 0, sizeof d);
        d.g = x.graph->view();
        d.P = x.plan->view();
        d.seed = x.seed;
        d.first = x.first_id;
        d.count = x.count;
        g.qidx.push_back(i);
        g.hq.push_back(d);
        b->rec_off.push_back(b->nrec);
        b->nrec += x.plan->nch;
    }
    size_t off = 0;
    for (auto &g : b->groups) {
        g.result_off = off;
        off += g.hq.size() * g.nch;
        if (hipMalloc(&g.dq, sizeof(BQuery) * g.hq.size()) != hipSuccess ||
            hipMalloc(&g.counter, sizeof(unsigned long long)) != hipSuccess)
            return set_error(ZG_ENOMEM, "batch descriptors: device allocation failed");
    }
    b->nres = off;
    if (hipMalloc(&b->d_result, sizeof(Acc) * off) != hipSuccess)
        return set_error(ZG_ENOMEM, "batch results: device allocation failed");
    ZG_HIP(hipHostMalloc(&b->h_result, sizeof(Acc) * off, hipHostMallocDefault));
    ZG_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    ZG_HIP(hipEventCreate(&b->e0));
    ZG_HIP(hipEventCreate(&b->e1));
    rc = configure(b.get());
    if (rc) return rc;
    *out = b.release();
    return ZG_OK;
}

extern "C" int zg_batch_records(const zg_batch *b) { return b ? (int)b->nrec : 0; }

extern "C" int zg_batch_run(zg_batch *b, zg_sums *out, float *ms)
{
    if (!b || !out) return set_error(ZG_EINVAL, "null argument");
    int rc = replay(b, ms);
    if (rc) return rc;
    for (auto &g : b->groups)
        for (size_t j = 0; j < g.hq.size(); ++j) {
            const int qi = g.qidx[j];
            const zg_plan *pl = b->queries[qi].plan;
            finish_sums(b->h_result + g.result_off + j * g.nch, g.nch, pl->channel_k(0), pl->channel_k(1), g.hq[j].count,
                        out + b->rec_off[qi]);
        }
    for (size_t r = 0; r < b->nrec; ++r)
        if (out[r].status) return set_error(ZG_EOVERFLOW, "an estimator value exceeded 128 bits");
    return ZG_OK;
}

extern "C" int zg_batch_set_ids(zg_batch *b, const uint64_t *seeds, const uint64_t *first_ids)
{
    if (!b) return set_error(ZG_EINVAL, "null argument");
    if (first_ids)
        for (size_t qi = 0; qi < b->queries.size(); ++qi)
            if (b->queries[qi].count > ~0ull - first_ids[qi])
                return set_error(ZG_EINVAL, "estimator ids: first_id + count must not exceed 2^64 - 1");
    for (auto &g : b->groups) {
        for (size_t j = 0; j < g.hq.size(); ++j) {
            const int qi = g.qidx[j];
            if (seeds) g.hq[j].seed = b->queries[qi].seed = seeds[qi];
            if (first_ids) g.hq[j].first = b->queries[qi].first_id = first_ids[qi];
        }
        ZG_HIP(hipMemcpyAsync(g.dq, g.hq.data(), sizeof(BQuery) * g.hq.size(), hipMemcpyHostToDevice, b->stream));
    }
    ZG_HIP(hipStreamSynchronize(b->stream));
    return ZG_OK;
}

extern "C" int zg_batch_autotune(zg_batch *b, int trials, zg_batch_config *chosen, float *best_ms)
{
    if (!b) return set_error(ZG_EINVAL, "null argument");
    if (trials < 1) trials = 3;
    const int blocks[] = {64, 128, 256};
    const int bpcs[] = {2, 4, 8, 16};
    const uint32_t chunks[] = {4096, 16384, 65536};
    auto measure = [&](const zg_batch_config &c, float *t) -> int {
        b->cfg = c;
        int rc = configure(b);
        if (rc) return rc;
        std::vector<float> v;
        for (int i = 0; i < trials; ++i) {
            float ms = 0;
            rc = replay(b, &ms);
            if (rc) return rc;
            v.push_back(ms);
        }
        std::sort(v.begin(), v.end());
        *t = v[v.size() / 2];
        return ZG_OK;
    };
    zg_batch_config best = b->cfg;
    float best_t = 0;
    int rc = measure(best, &best_t);
    if (rc) return rc;
    // coordinate descent: chunk size, then workgroups per CU, then block size, twice
    for (int pass = 0; pass < 2; ++pass) {
        for (int dim = 0; dim < 3; ++dim) {
            const int n = dim == 0 ? 3 : dim == 1 ? 4 : 3;
            for (int k = 0; k < n; ++k) {
                zg_batch_config c = best;
                if (dim == 0) c.chunk_estimators = chunks[k];
                else if (dim == 1) c.blocks_per_cu = bpcs[k];
                else c.block_size = blocks[k];
                if (!std::memcmp(&c, &best, sizeof c)) continue;
                float t = 0;
                rc = measure(c, &t);
                if (rc) return rc;
                if (t < best_t) { best_t = t; best = c; }
            }
        }
    }
    b->cfg = best;
    rc = configure(b);
    if (rc) return rc;
    if (chosen) *chosen = best;
    if (best_ms) *best_ms = best_t;
    return ZG_OK;
}

extern "C" int zg_batch_config_get(const zg_batch *b, zg_batch_config *out)
{
    if (!b || !out) return set_error(ZG_EINVAL, "null argument");
    *out = b->cfg;
    return ZG_OK;
}

extern "C" int zg_batch_destroy(zg_batch *b)
{
    delete b;
    return ZG_OK;
}

extern "C" int zg_run_batch(const zg_query *q, int nq, const zg_batch_config *config, zg_sums *out, float *ms)
{
    zg_batch *b = nullptr;
    int rc = zg_batch_create(q, nq, config, &b);
    if (rc) return rc;
    rc = zg_batch_run(b, out, ms);
    zg_batch_destroy(b);
    return rc;
}
