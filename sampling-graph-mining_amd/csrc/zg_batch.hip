// zg_batch.hip -- many estimator queries (graph, plan, seed, id range) in one
// replayable HIP graph: a counter reset, ONE persistent work-queue launch over
// the chunks of every query of every plan kind (zg_kernels.hip, k_batch), a
// per-query reduction, and one copy of every query's record to pinned host
// memory.  The reference answers one query per ZGraphInstance::calculate call
// (src/ZGraphInstance.cpp:256-297); serving many small queries that way is
// bound by per-call launch and synchronisation cost, which the captured graph
// pays once per batch, and a launch per plan kind would idle the GPU on each
// kind's tail, which the single queue does not.
//
// Queries are laid out in the queue by decreasing estimated cost per
// estimator (random lines each estimator reads, larger graphs first on ties),
// so the long chunks are claimed first and the queue drains on short ones.
// The autotuner picks the launch shape (threads per workgroup, persistent
// workgroups per CU, estimators per wavefront per claim) by timing graph
// replays; every shape gives the same exact sums (only the floating-point sum
// of squares can differ in its last bits, as the chunking changes its order).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "zg_internal.h"

using namespace zg;

namespace {
constexpr uint64_t kMaxQueryCount = 1ull << 31;      // digit sums of one query stay below 2^63
constexpr uint64_t kMaxChunksPerQuery = 16384;       // a larger query takes longer chunks
constexpr uint32_t kDefaultChunk = 1024;             // 16 estimators per lane per claim
}

struct zg_batch {
    std::vector<zg_query> queries;      // the caller's order
    std::vector<int> order;             // queue position -> caller's index
    std::vector<BQuery> hq;             // descriptors in queue order
    std::vector<size_t> rec_off;        // first output record of each query (caller's order)
    size_t nrec = 0;
    uint64_t nchunks = 0;
    zg_batch_config cfg{};
    hipStream_t stream = nullptr;
    BQuery *dq = nullptr;
    unsigned long long *counter = nullptr;
    Acc *recs = nullptr;                // chunk records
    size_t recs_cap = 0;
    Acc *d_result = nullptr, *h_result = nullptr;   // h_result is pinned
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;

    ~zg_batch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (dq) (void)hipFree(dq);
        if (counter) (void)hipFree(counter);
        if (recs) (void)hipFree(recs);
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
    if (c.chunk_estimators < 64 || c.chunk_estimators > (1u << 24) || c.chunk_estimators % 64)
        return set_error(ZG_EINVAL, "chunk_estimators must be a multiple of 64 in [64, 2^24]");
    return ZG_OK;
}

zg_batch_config defaults(const zg_batch_config *c)
{
    zg_batch_config d{256, 8, kDefaultChunk};
    if (c) {
        if (c->block_size) d.block_size = c->block_size;
        if (c->blocks_per_cu) d.blocks_per_cu = c->blocks_per_cu;
        if (c->chunk_estimators) d.chunk_estimators = c->chunk_estimators;
    }
    return d;
}

// random 128-byte lines one estimator of the kernel kind reads (measured on
// power-law graphs; only the queue order depends on them)
double lines_per_estimator(const zg_plan *p)
{
    switch (p->desc.kind) {
    case ZG_PLAN_THREE_CHAIN: return 2.0;
    case ZG_PLAN_TRIANGLE: case ZG_PLAN_THREE_MOTIF: case ZG_PLAN_K4_CLOSE: case ZG_PLAN_TRIANGLE_K4: return 3.0;
    case ZG_PLAN_K4_SPLIT: case ZG_PLAN_TRIANGLE_K4_SPLIT: return 3.0 + 0.01 * p->desc.completions;   // ~0.1 % of first stages close, ~10 lines per completion
    case ZG_PLAN_HOUSE_SPLIT: return 3.0 + 0.003 * p->desc.completions;   // ~0.1 % of roofs close, 3 lines per completion
    case ZG_PLAN_FOUR_CHAIN: return 4.5;
    case ZG_PLAN_CHAIN: return 1.0 + 2.3 * (p->desc.order - 2);
    case ZG_PLAN_PATTERN: return 1.0 + 2.1 * (p->desc.pattern_edges - 1);
    default: return 3.0;
    }
}

// chunking, descriptor upload and graph capture for the batch's current config
int configure(zg_batch *b)
{
    if (b->exec) { ZG_HIP(hipGraphExecDestroy(b->exec)); b->exec = nullptr; }
    if (b->graph) { ZG_HIP(hipGraphDestroy(b->graph)); b->graph = nullptr; }
    uint64_t c = 0, r = 0;
    for (auto &q : b->hq) {
        const uint64_t need = (q.count + kMaxChunksPerQuery - 1) / kMaxChunksPerQuery;
        q.chunk = std::max<uint64_t>(b->cfg.chunk_estimators, (need + 63) / 64 * 64);
        const uint64_t nc = (q.count + q.chunk - 1) / q.chunk;
        q.chunk0 = c;
        q.rec0 = r;
        c += nc;
        r += nc * (uint64_t)q.P.nch;
    }
    b->nchunks = c;
    const size_t need = std::max<size_t>(1, (size_t)r);
    if (need > b->recs_cap) {
        if (b->recs) ZG_HIP(hipFree(b->recs));
        b->recs = nullptr;
        b->recs_cap = 0;
        if (hipMalloc(&b->recs, sizeof(Acc) * need) != hipSuccess)
            return set_error(ZG_ENOMEM, "batch chunk records: device allocation failed");
        b->recs_cap = need;
    }
    ZG_HIP(hipMemcpyAsync(b->dq, b->hq.data(), sizeof(BQuery) * b->hq.size(), hipMemcpyHostToDevice, b->stream));
    ZG_HIP(hipStreamSynchronize(b->stream));
    const int grid = device_cus() * b->cfg.blocks_per_cu;
    ZG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_batch(b->dq, (int)b->hq.size(), b->nchunks, b->counter, b->recs, b->d_result, grid,
                          b->cfg.block_size, b->stream);
    if (!rc && hipMemcpyAsync(b->h_result, b->d_result, sizeof(Acc) * b->nrec, hipMemcpyDeviceToHost, b->stream) !=
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
    std::vector<double> cost(nq);
    for (int i = 0; i < nq; ++i) {
        const zg_query &x = q[i];
        if (!x.graph || !x.plan) return set_error(ZG_EINVAL, "query without graph or plan");
        if (x.count > kMaxQueryCount) return set_error(ZG_EINVAL, "a batched query runs at most 2^31 estimators");
        if (x.count > ~0ull - x.first_id) return set_error(ZG_EINVAL, "estimator ids: first_id + count must not exceed 2^64 - 1");
        if (x.graph->m == 0 && x.count) return set_error(ZG_EINVAL, "query on an empty graph");
        cost[i] = lines_per_estimator(x.plan);
        b->rec_off.push_back(b->nrec);
        b->nrec += x.plan->nch;
    }
    b->order.resize(nq);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) {
        if (cost[a] != cost[c]) return cost[a] > cost[c];
        return q[a].graph->m > q[c].graph->m;
    });
    uint32_t res = 0;
    for (int i : b->order) {
        const zg_query &x = q[i];
        BQuery d;
        std::memset(&d, 0, sizeof d);
        d.g = x.graph->view();
        d.P = x.plan->view();
        d.seed = x.seed;
        d.first = x.first_id;
        d.count = x.count;
        d.res0 = res;
        d.kk = x.plan->desc.kind == ZG_PLAN_PATTERN ? 100 + x.plan->desc.pattern_vertices : x.plan->desc.kind;
        res += (uint32_t)x.plan->nch;
        b->hq.push_back(d);
    }
    if (hipMalloc(&b->dq, sizeof(BQuery) * nq) != hipSuccess ||
        hipMalloc(&b->counter, sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&b->d_result, sizeof(Acc) * b->nrec) != hipSuccess)
        return set_error(ZG_ENOMEM, "batch descriptors: device allocation failed");
    ZG_HIP(hipHostMalloc(&b->h_result, sizeof(Acc) * b->nrec, hipHostMallocDefault));
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
    for (size_t j = 0; j < b->hq.size(); ++j) {
        const int qi = b->order[j];
        const zg_plan *pl = b->queries[qi].plan;
        finish_sums(b->h_result + b->hq[j].res0, pl->nch, pl->channel_k(0), pl->channel_k(1), b->hq[j].count,
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
    for (size_t j = 0; j < b->hq.size(); ++j) {
        const int qi = b->order[j];
        if (seeds) b->hq[j].seed = b->queries[qi].seed = seeds[qi];
        if (first_ids) b->hq[j].first = b->queries[qi].first_id = first_ids[qi];
    }
    ZG_HIP(hipMemcpyAsync(b->dq, b->hq.data(), sizeof(BQuery) * b->hq.size(), hipMemcpyHostToDevice, b->stream));
    ZG_HIP(hipStreamSynchronize(b->stream));
    return ZG_OK;
}

extern "C" int zg_batch_autotune(zg_batch *b, int trials, zg_batch_config *chosen, float *best_ms)
{
    if (!b) return set_error(ZG_EINVAL, "null argument");
    if (trials < 1) trials = 3;
    const int blocks[] = {64, 128, 256};
    const int bpcs[] = {2, 4, 8, 16};
    const uint32_t chunks[] = {256, 1024, 4096, 16384};   // estimators per wavefront per claim
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
    // coordinate descent: estimators per claim, then workgroups per CU, then block size, twice
    for (int pass = 0; pass < 2; ++pass) {
        for (int dim = 0; dim < 3; ++dim) {
            const int n = dim == 0 ? 4 : dim == 1 ? 4 : 3;
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
