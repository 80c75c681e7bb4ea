// ZGraph.hpp -- C++ host API of the MI355X estimator library, with the class
// names, argument meaning and error behaviour of ZGraph's host API
// (include/ZGraphInstance.hpp, include/SamplerGenerator.hpp,
// include/ErrorLatencyProfile.hpp, include/Graph.hpp, include/GraphFilters.hpp,
// include/Regression.hpp), so the reference's application mains compile
// against it with their includes swapped.  Header-only; every call goes
// through the C ABI of libzgraph_amd.so (include/zgraph_amd.h).
//
// Differences from the reference, by design:
//   * estimate() bodies are device plans.  The built-in instances (triangle,
//     3-chain, 3-motif, chain mining, 4-chain, 4-clique, SamplerGenerator)
//     select a plan; a user subclass that only overrides estimate() cannot
//     run on the device and calculate() throws zgraph::Error for it.
//   * One worker (W = 1): partitioners, NUMA destroyers and RandomMapping are
//     accepted for source compatibility and do nothing (vertex relabeling
//     does not change any count).  get_scaling_factor is 1.
//   * Estimator ids are global: calculate(N, k, init) runs ids
//     [init.get_num_finished_estimators(), N), so a top-up of an Estimation
//     equals one run of N estimators.  Batches that must be independent (the
//     TimeProfile's) take disjoint id ranges.
//   * Failures throw zgraph::Error (std::runtime_error) carrying the ZG_E*
//     code and message, where the reference asserts or exits.
#ifndef ZGRAPH_AMD_HOST_HPP
#define ZGRAPH_AMD_HOST_HPP

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "zgraph_amd.h"

namespace zgraph {

typedef uint32_t VertexId;   // include/Type.hpp:6-10
typedef uint64_t EdgeId;
struct Empty {};

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

inline void check(int rc)
{
    if (rc != ZG_OK) {
        char buf[1024];
        zg_last_error(buf, sizeof buf);
        throw Error(rc, std::string("zgraph: ") + buf);
    }
}

// ---- meta data and seeds (src/Utilities.cpp:24-47) -------------------------
// |V| is the first integer on line 2 of <graph>.csv.
inline void load_graph_meta_data(const std::string &path, VertexId &num_vertices)
{
    std::ifstream f(path);
    std::string line;
    if (!f || !std::getline(f, line) || !std::getline(f, line))
        throw Error(ZG_EINVAL, "load_graph_meta_data: cannot read line 2 of " + path);
    size_t i = 0;
    while (i < line.size() && (line[i] < '0' || line[i] > '9')) ++i;
    if (i == line.size()) throw Error(ZG_EINVAL, "load_graph_meta_data: no vertex count in " + path);
    num_vertices = (VertexId)std::strtoul(line.c_str() + i, nullptr, 10);
}

// The reference seeds from time(NULL) on rank 0; here the seed is fixed (1)
// unless ZGRAPH_SEED is set, so runs are reproducible.
inline unsigned int get_globally_consistent_seed()
{
    const char *s = std::getenv("ZGRAPH_SEED");
    return s ? (unsigned int)std::strtoul(s, nullptr, 10) : 1u;
}

// ---- Graph<EdgeData> (include/Graph.hpp:148-166) ---------------------------
// A device-resident CSR.  EdgeData is Empty (8-byte edge units), or a 4-byte
// weight (VertexId / unsigned / float: 12-byte units); weights are stored as
// float on the device (exact for integers below 2^24).
template <typename EdgeData>
class Graph {
  public:
    VertexId num_vertices = 0;
    EdgeId num_edges = 0;         // graph.num_edges (ELP bound input)
    EdgeId num_local_edges = 0;   // stored edges
    EdgeId max_degree = 0;

    Graph() = default;
    Graph(const Graph &) = delete;
    Graph &operator=(const Graph &) = delete;
    Graph(Graph &&o) noexcept { *this = std::move(o); }
    Graph &operator=(Graph &&o) noexcept
    {
        if (this != &o) {
            reset();
            h_ = o.h_;
            o.h_ = nullptr;
            num_vertices = o.num_vertices;
            num_edges = o.num_edges;
            num_local_edges = o.num_local_edges;
            max_degree = o.max_degree;
        }
        return *this;
    }
    ~Graph() { reset(); }

    // upload host edges (src[i], dst[i]) [+ weights] and build the CSR in HBM
    void create(const std::vector<uint32_t> &src, const std::vector<uint32_t> &dst, const float *weights,
                VertexId n, uint64_t shuffle_seed)
    {
        if (src.size() != dst.size()) throw Error(ZG_EINVAL, "Graph::create: src/dst length mismatch");
        zg_graph *g = nullptr;
        check(zg_graph_create(src.data(), dst.data(), weights, src.size(), n, shuffle_seed, &g));
        adopt(g);
    }
    void adopt(zg_graph *g)
    {
        reset();
        h_ = g;
        zg_graph_info info;
        check(zg_graph_info_get(h_, &info));
        num_vertices = info.n;
        num_edges = info.num_edges;
        num_local_edges = info.m;
        max_degree = info.max_degree;
    }
    void reset()
    {
        if (h_) zg_graph_destroy(h_);
        h_ = nullptr;
        num_vertices = 0;
        num_edges = num_local_edges = max_degree = 0;
    }
    zg_graph *handle() const
    {
        if (!h_) throw Error(ZG_ESTATE, "graph not loaded");
        return h_;
    }
    bool loaded() const { return h_ != nullptr; }

    // edges in rank order (src, dst, weight-or-0), for filters and inspection
    void download_edges(std::vector<uint32_t> &src, std::vector<uint32_t> &dst, std::vector<float> *w) const
    {
        src.resize(num_local_edges);
        dst.resize(num_local_edges);
        if (w) w->resize(num_local_edges);
        check(zg_graph_export(handle(), src.data(), dst.data(), w ? w->data() : nullptr, nullptr, nullptr, nullptr));
    }
    uint64_t exact_count(int what) const
    {
        uint64_t r = 0;
        check(zg_exact_counts(handle(), what, &r));
        return r;
    }

  private:
    zg_graph *h_ = nullptr;
};

// Loader chain with the reference's names (src/Graph.cpp:125-156, 215-615).
class UniformGraphPartitioner {
  public:
    explicit UniformGraphPartitioner(VertexId num_vertices) : n_(num_vertices) {}
    VertexId num_vertices() const { return n_; }

  private:
    VertexId n_;
};
class RandomMapping {   // relabeling is a no-op at one worker
  public:
    RandomMapping(unsigned int seed, VertexId num_vertices) : seed_(seed), n_(num_vertices) {}
    unsigned int seed() const { return seed_; }

  private:
    unsigned int seed_;
    VertexId n_;
};
template <typename EdgeData> class CSRConstructor {};
template <typename EdgeData> class GraphDestroyer {
  public:
    virtual ~GraphDestroyer() = default;
    virtual void destroy_graph(Graph<EdgeData> &g) { g.reset(); }
};
template <typename EdgeData> class NumaAwareGraphDestroyer : public GraphDestroyer<EdgeData> {};

template <typename EdgeData>
constexpr unsigned edge_unit_bytes() { return std::is_same<EdgeData, Empty>::value ? 8u : 8u + sizeof(EdgeData); }

// DistGraphLoader<E>::load_graph (src/Graph.cpp:386-615) at W = 1: reads the
// packed <graph>.biedgelist {u32 src, u32 dst[, 4-byte data]}, rejects
// self-loops and ids >= n, ranks the edges by a seeded random permutation
// and builds the rank-ordered CSR on the device.
template <typename EdgeData>
class DistGraphLoader {
    static_assert(std::is_same<EdgeData, Empty>::value || sizeof(EdgeData) == 4, "4-byte edge data only");

  public:
    DistGraphLoader(UniformGraphPartitioner *, RandomMapping *mapping, CSRConstructor<EdgeData> *,
                    GraphDestroyer<EdgeData> *destroyer)
        : seed_(mapping ? mapping->seed() : 1u), destroyer_(destroyer)
    {
    }
    void load_graph(const std::string &path, VertexId num_vertices, Graph<EdgeData> &graph)
    {
        const unsigned unit = edge_unit_bytes<EdgeData>();
        std::ifstream f(path, std::ios::binary | std::ios::ate);
        if (!f) throw Error(ZG_EINVAL, "load_graph: cannot open " + path);
        const uint64_t size = (uint64_t)f.tellg();
        if (size % unit) throw Error(ZG_EINVAL, "load_graph: file size is not a multiple of the edge unit");
        const uint64_t m = size / unit;   // src/Graph.cpp:400-402
        f.seekg(0);
        std::vector<uint32_t> raw(size / 4);
        f.read(reinterpret_cast<char *>(raw.data()), (std::streamsize)size);
        const unsigned words = unit / 4;
        std::vector<uint32_t> src(m), dst(m);
        std::vector<float> w;
        if (words == 3) w.resize(m);
        for (uint64_t i = 0; i < m; ++i) {
            src[i] = raw[i * words];
            dst[i] = raw[i * words + 1];
            if (words == 3) w[i] = word_to_float(raw[i * words + 2]);
        }
        graph.create(src, dst, words == 3 ? w.data() : nullptr, num_vertices, seed_);
    }
    void destroy_graph(Graph<EdgeData> &graph)
    {
        if (destroyer_) destroyer_->destroy_graph(graph);
        else graph.reset();
    }

  private:
    static float word_to_float(uint32_t x)
    {
        if (std::is_same<EdgeData, float>::value) {
            float f;
            std::memcpy(&f, &x, 4);
            return f;
        }
        return (float)x;
    }
    unsigned int seed_;
    GraphDestroyer<EdgeData> *destroyer_;
};

// GraphSampler<E>::sample_graph (src/Graph.cpp:628-780): keep each edge with
// probability `rate`, rebuild the CSR with a fresh rank permutation.
template <typename EdgeData>
class GraphSampler {
  public:
    GraphSampler(CSRConstructor<EdgeData> *, GraphDestroyer<EdgeData> *, double rate, uint64_t seed = 1)
        : rate_(rate), seed_(seed)
    {
    }
    void sample_graph(const Graph<EdgeData> &input, Graph<EdgeData> &output)
    {
        zg_graph *g = nullptr;
        check(zg_graph_subsample(input.handle(), rate_, seed_, &g));
        output.adopt(g);
    }

  private:
    double rate_;
    uint64_t seed_;
};

// The threshold predicates of the reference's applications (edge_data >= t,
// and its negation) run on the device; any other std::function predicate is
// evaluated on the host over the edge records in rank order, and the kept
// edges are uploaded with the same shuffle seed (the same graph the device
// filter would build for the same keep set).
template <typename EdgeData>
struct ThresholdPredicate {
    float threshold;
    bool keep_ge;
    bool operator()(EdgeData d) const { return keep_ge ? (float)d >= threshold : (float)d < threshold; }
};
template <typename EdgeData>
ThresholdPredicate<EdgeData> at_least(float t) { return ThresholdPredicate<EdgeData>{t, true}; }
template <typename EdgeData>
ThresholdPredicate<EdgeData> below(float t) { return ThresholdPredicate<EdgeData>{t, false}; }

// GraphFilter<E>::generate_graph (src/GraphFilters.cpp:16-177)
template <typename EdgeData>
class GraphFilter {
    static_assert(!std::is_same<EdgeData, Empty>::value, "GraphFilter needs edge data");

  public:
    GraphFilter(CSRConstructor<EdgeData> *, GraphDestroyer<EdgeData> *, ThresholdPredicate<EdgeData> p,
                uint64_t seed = 1)
        : thr_(p), is_thr_(true), seed_(seed)
    {
    }
    GraphFilter(CSRConstructor<EdgeData> *, GraphDestroyer<EdgeData> *, std::function<bool(EdgeData)> p,
                uint64_t seed = 1)
        : fn_(std::move(p)), is_thr_(false), seed_(seed)
    {
    }
    void generate_graph(const Graph<EdgeData> &input, Graph<EdgeData> &output)
    {
        if (is_thr_) {
            zg_graph *g = nullptr;
            check(zg_graph_filter(input.handle(), thr_.threshold, thr_.keep_ge ? 1 : 0, seed_, &g));
            output.adopt(g);
            return;
        }
        std::vector<uint32_t> s, d, ks, kd;
        std::vector<float> w, kw;
        input.download_edges(s, d, &w);
        for (size_t i = 0; i < s.size(); ++i)
            if (fn_(from_float(w[i]))) {
                ks.push_back(s[i]);
                kd.push_back(d[i]);
                kw.push_back(w[i]);
            }
        output.create(ks, kd, kw.data(), input.num_vertices, seed_);
    }
    void destroy_graph(Graph<EdgeData> &g) { g.reset(); }

  private:
    static EdgeData from_float(float x) { return (EdgeData)x; }
    ThresholdPredicate<EdgeData> thr_{0.f, true};
    std::function<bool(EdgeData)> fn_;
    bool is_thr_;
    uint64_t seed_;
};

// ---- Estimation (src/ZGraphInstance.cpp:15-35) ------------------------------
// The running estimator-weighted mean of the reference, plus the exact
// integer sums, success count and sum of squares of every absorbed run (for
// the exact mean and the CLT interval).
class Estimation {
  public:
    Estimation() { reset(); }
    void reset()
    {
        num_finished_estimators = 0;
        current_estimation = 0;
        exact_hi = exact_lo = 0;
        log2_den = -1;
        sumsq = 0;
        successes = 0;
    }
    void update(unsigned long long new_num_estimators, double new_estimation)
    {
        const double n0 = (double)num_finished_estimators, n1 = (double)new_num_estimators;
        current_estimation = current_estimation * n0 / (n0 + n1) + new_estimation * n1 / (n0 + n1);
        num_finished_estimators += new_num_estimators;
    }
    unsigned long long get_num_finished_estimators() const { return num_finished_estimators; }
    double get_currnet_estimation() const { return current_estimation; }   // spelling of include/ZGraphInstance.hpp:27
    double get_current_estimation() const { return current_estimation; }

    void absorb(const zg_sums &s)
    {
        if (log2_den < 0) log2_den = s.log2_den;
        if (s.log2_den != log2_den) throw Error(ZG_ESTATE, "Estimation: runs of different plans mixed");
        // 192-bit add of s.sum into (exact_hi:exact_lo) as long double pieces
        const unsigned __int128 lo = ((unsigned __int128)s.sum[1] << 64) | s.sum[0];
        const unsigned __int128 old = exact_lo;
        exact_lo += lo;
        exact_hi += s.sum[2] + (exact_lo < old ? 1 : 0);
        sumsq += s.sumsq;
        successes += s.successes;
    }
    // exact mean of every absorbed estimator value
    double exact_mean() const
    {
        if (!num_finished_estimators || log2_den < 0) return 0.0;
        const long double tot = std::ldexp((long double)exact_hi, 128) + (long double)exact_lo;
        return (double)(std::ldexp(tot, -log2_den) / (long double)num_finished_estimators);
    }
    double get_sumsq() const { return sumsq; }
    unsigned long long get_successes() const { return successes; }
    // two-sided CLT half-width of the mean at `confidence`
    double half_width(double confidence) const;

  private:
    unsigned long long num_finished_estimators;
    double current_estimation;
    unsigned __int128 exact_lo;
    uint64_t exact_hi;
    int log2_den;
    double sumsq;
    unsigned long long successes;
};

// inverse standard-normal CDF by bisection on erfc
inline double normal_quantile_two_sided(double confidence)
{
    const double p = 1.0 - confidence;   // two-sided tail mass
    double lo = 0, hi = 40;
    for (int i = 0; i < 200; ++i) {
        const double mid = 0.5 * (lo + hi);
        if (std::erfc(mid / std::sqrt(2.0)) > p) lo = mid;
        else hi = mid;
    }
    return 0.5 * (lo + hi);
}

inline double Estimation::half_width(double confidence) const
{
    const double t = (double)num_finished_estimators;
    if (t < 2) return std::numeric_limits<double>::infinity();
    const double mu = exact_mean();
    const double var = std::max(sumsq / t - mu * mu, 0.0) * t / (t - 1);
    return normal_quantile_two_sided(confidence) * std::sqrt(var / t);
}

// ---- Pattern (include/SamplerGenerator.hpp:9-47) -----------------------------
template <typename EdgeData>
struct Edge {
    EdgeId edge_id;
    VertexId src;
    VertexId dst;
};

// ---- per-estimator state types (include/ZGraphInstance.hpp:30-105) ---------
// Host value types with the reference's semantics.  The device plans keep
// this state in registers; these serve host code that describes or checks
// sampled sub-graphs (and the reference's tests/TestSubgraph.cpp).
constexpr int MAX_LIST_SIZE = 16;

template <typename Data>
struct MyList {
    int list_size = 0;
    Data list[MAX_LIST_SIZE];
    void push_back(const Data &d)
    {
        if (list_size >= MAX_LIST_SIZE) throw Error(ZG_EINVAL, "MyList: more than 16 entries");
        list[list_size++] = d;
    }
};

template <typename EdgeData>
struct SubGraph {
    MyList<Edge<EdgeData>> edges;
    MyList<VertexId> vertices;
    EdgeId min_valid_edge_id = 0;       // one past the largest edge id added
    VertexId min_valid_vertex_id = 0;   // one past the largest vertex added

    bool is_in_vertices_list(VertexId v) const
    {
        for (int i = 0; i < vertices.list_size; ++i)
            if (vertices.list[i] == v) return true;
        return false;
    }
    // appends the edge, then its src and dst (in that order) if new (src/ZGraphInstance.cpp:56-73)
    void add_edge(const Edge<EdgeData> &e)
    {
        edges.push_back(e);
        if (e.edge_id >= min_valid_edge_id) min_valid_edge_id = e.edge_id + 1;
        add_vertex(e.src);
        add_vertex(e.dst);
    }
    void add_vertex(VertexId v)
    {
        if (is_in_vertices_list(v)) return;
        vertices.push_back(v);
        if (v >= min_valid_vertex_id) min_valid_vertex_id = v + 1;
    }
    const MyList<Edge<EdgeData>> &get_edges() const { return edges; }
    const MyList<VertexId> &get_vertices() const { return vertices; }
    EdgeId get_min_valid_edge_id() const { return min_valid_edge_id; }
    VertexId get_min_valid_vertex_id() const { return min_valid_vertex_id; }
    // the edges of this sub-graph whose endpoints (either orientation) are not an edge of `o`
    // (src/ZGraphInstance.cpp:96-113)
    SubGraph operator-(const SubGraph &o) const
    {
        SubGraph r;
        for (int i = 0; i < edges.list_size; ++i) {
            const Edge<EdgeData> &a = edges.list[i];
            bool drop = false;
            for (int j = 0; j < o.edges.list_size && !drop; ++j) {
                const Edge<EdgeData> &b = o.edges.list[j];
                drop = (a.src == b.src && a.dst == b.dst) || (a.src == b.dst && a.dst == b.src);
            }
            if (!drop) r.add_edge(a);
        }
        return r;
    }
};

class Pattern {
  public:
    Pattern(VertexId num_vertices_, EdgeId num_edges_, const Edge<Empty> *edges_) : nv_(num_vertices_)
    {
        if (num_edges_ > ZG_MAX_PATTERN_EDGES) throw Error(ZG_EINVAL, "Pattern: more than 15 edges");
        for (EdgeId i = 0; i < num_edges_; ++i) edges_v_.push_back(edges_[i]);
    }
    VertexId get_num_vertices() const { return nv_; }
    EdgeId get_num_edges() const { return edges_v_.size(); }
    const Edge<Empty> *get_edges() const { return edges_v_.data(); }
    zg_plan_desc desc() const
    {
        zg_plan_desc d;
        std::memset(&d, 0, sizeof d);
        d.kind = ZG_PLAN_PATTERN;
        d.pattern_vertices = (int)nv_;
        d.pattern_edges = (int)edges_v_.size();
        for (size_t i = 0; i < edges_v_.size(); ++i) {
            d.pattern_src[i] = edges_v_[i].src;
            d.pattern_dst[i] = edges_v_[i].dst;
        }
        return d;
    }
    // SamplerGenerator's relabeling dedupe (src/SamplerGenerator.cpp:311-363):
    // (unique relabelings, k!)
    std::pair<int, int> unique_relabelings() const
    {
        const zg_plan_desc d = desc();
        int u = 0, t = 0;
        check(zg_pattern_unique(&d, &u, &t));
        return {u, t};
    }
    // Pattern::is_automorphic (src/SamplerGenerator.cpp:225-242): same edge set
    bool is_automorphic(const Pattern &o) const
    {
        if (o.nv_ != nv_ || o.edges_v_.size() != edges_v_.size()) return false;
        for (const auto &e : o.edges_v_) {
            bool found = false;
            for (const auto &f : edges_v_)
                found = found || (e.src == f.src && e.dst == f.dst) || (e.src == f.dst && e.dst == f.src);
            if (!found) return false;
        }
        return true;
    }

  private:
    VertexId nv_;
    std::vector<Edge<Empty>> edges_v_;
};

// ---- Plans and instances (include/ZGraphInstance.hpp:107-140) ---------------
class Plan {
  public:
    explicit Plan(const zg_plan_desc &d) : kind_(d.kind) { check(zg_plan_create(&d, &h_)); }
    static zg_plan_desc of(int kind, int order = 0)
    {
        zg_plan_desc d;
        std::memset(&d, 0, sizeof d);
        d.kind = kind;
        d.order = order;
        return d;
    }
    // ZG_PLAN_K4_SPLIT (or, fused with the triangle channel, ZG_PLAN_TRIANGLE_K4_SPLIT)
    // with `completions` per closed triangle; also ZG_PLAN_HOUSE_SPLIT (per closed roof)
    static zg_plan_desc split_k4(int completions, int kind = ZG_PLAN_K4_SPLIT)
    {
        zg_plan_desc d = of(kind);
        d.completions = completions;
        return d;
    }
    Plan(const Plan &) = delete;
    Plan &operator=(const Plan &) = delete;
    ~Plan()
    {
        if (h_) zg_plan_destroy(h_);
    }
    zg_plan *handle() const { return h_; }
    int kind() const { return kind_; }
    int channels() const { return zg_plan_channels(h_); }
    double cond_calls() const { return zg_plan_cond_calls(h_); }

  private:
    int kind_ = 0;
    zg_plan *h_ = nullptr;
};

template <typename EdgeData>
class ZGraphInstance {
  public:
    ZGraphInstance() = default;
    explicit ZGraphInstance(Graph<EdgeData> *graph_) : graph(graph_) {}
    virtual ~ZGraphInstance() = default;

    virtual Graph<EdgeData> *get_graph() { return graph; }
    virtual void set_graph(Graph<EdgeData> *graph_) { graph = graph_; }

    // ZGraphInstance::calculate (src/ZGraphInstance.cpp:256-297): runs the
    // estimators of ids [init.finished, N) on the device and folds them into
    // init_estimation; returns the running estimate.
    virtual double calculate(unsigned long long num_estimators, VertexId num_vertices_in_pattern,
                             Estimation &init_estimation)
    {
        (void)num_vertices_in_pattern;
        const unsigned long long done = init_estimation.get_num_finished_estimators();
        if (done >= num_estimators) return init_estimation.get_currnet_estimation();
        const unsigned long long n = num_estimators - done;
        const zg_sums s = run_ids(done, n);
        init_estimation.absorb(s);
        init_estimation.update(n, s.mean);   // get_scaling_factor == 1 at one worker
        return init_estimation.get_currnet_estimation();
    }

    // the sums of the estimators of ids [first_id, first_id + count)
    zg_sums run_ids(unsigned long long first_id, unsigned long long count, float *kernel_ms = nullptr)
    {
        require_plan();
        std::vector<zg_sums> s((size_t)plan_->channels());
        zg_launch launch{0, 0, 0};
        check(zg_run(get_run_graph()->handle(), plan_->handle(), seed, first_id, count, &launch, s.data(),
                     kernel_ms));
        return s[(size_t)channel];
    }

    // get_num_sample_edge_calls (src/ZGraphInstance.cpp:449-468) returns 1
    virtual double get_num_sample_edge_calls() { return 1.0; }
    // get_num_conditional_sample_edge_calls (:470-488): the plan's constant
    virtual double get_num_conditional_sample_edge_calls()
    {
        require_plan();
        return plan_->cond_calls();
    }

    // Per-estimator host callbacks cannot run on the device; only the plans can.
    virtual double estimate()
    {
        throw Error(ZG_EINVAL, "estimate(): per-estimator host callbacks are not supported; use a plan "
                               "(built-in instance or SamplerGenerator)");
    }

    const Plan *plan() const { return plan_.get(); }
    // the graph the plan runs on (the filtered one for predicate instances)
    virtual Graph<EdgeData> *get_run_graph() { return graph; }
    uint64_t seed = 1;   // Philox key of the estimator streams
    int channel = 0;

  protected:
    void set_plan(const zg_plan_desc &d) { plan_.reset(new Plan(d)); }
    void require_plan() const
    {
        if (!plan_) throw Error(ZG_EINVAL, "this instance has no device plan (estimate() callbacks are host-only)");
        if (!graph) throw Error(ZG_ESTATE, "set_graph must be called first");
    }
    Graph<EdgeData> *graph = nullptr;

  private:
    std::unique_ptr<Plan> plan_;
};

// applications/Triangle.cpp:42-74 (class name spelled as the reference)
template <typename EdgeData>
class TriangleCoutingInstance : public ZGraphInstance<EdgeData> {
  public:
    explicit TriangleCoutingInstance(Graph<EdgeData> *g) : ZGraphInstance<EdgeData>(g)
    {
        this->set_plan(Plan::of(ZG_PLAN_TRIANGLE));
    }
};
template <typename EdgeData> using TriangleCountingInstance = TriangleCoutingInstance<EdgeData>;

// applications/ThreeChain.cpp:16-38
template <typename EdgeData>
class ThreeChainCountingInstance : public ZGraphInstance<EdgeData> {
  public:
    explicit ThreeChainCountingInstance(Graph<EdgeData> *g) : ZGraphInstance<EdgeData>(g)
    {
        this->set_plan(Plan::of(ZG_PLAN_THREE_CHAIN));
    }
};

// applications/ThreeMotif.cpp:48-121: 3-chains (channel 0) and triangles
// (channel 1) from the same wedges (the reference's cached sub-patterns).
template <typename EdgeData>
class ThreeMotifInstance : public ZGraphInstance<EdgeData> {
  public:
    explicit ThreeMotifInstance(Graph<EdgeData> *g, int channel_ = 0) : ZGraphInstance<EdgeData>(g)
    {
        this->set_plan(Plan::of(ZG_PLAN_THREE_MOTIF));
        this->channel = channel_;
    }
    // both counts from one pass over ids [0, N)
    std::pair<double, double> calculate_both(unsigned long long num_estimators)
    {
        this->require_plan();
        zg_sums s[2];
        zg_launch launch{0, 0, 0};
        check(zg_run(this->get_run_graph()->handle(), this->plan()->handle(), this->seed, 0, num_estimators, &launch,
                     s, nullptr));
        return {s[0].mean, s[1].mean};
    }
};

// Triangles (channel 0, TriangleCoutingInstance's estimator) and 4-cliques
// (channel 1, the three-shape plan) from one first edge, conditional sample and
// closing check per id: ThreeMotif's cached sub-pattern sharing
// (src/ZGraphInstance.cpp:596-833) applied to triangle -> 4-clique.  Each
// channel equals its own instance's calculate on the same ids.
template <typename EdgeData>
class TriangleFourCliqueInstance : public ZGraphInstance<EdgeData> {
  public:
    explicit TriangleFourCliqueInstance(Graph<EdgeData> *g, int channel_ = 0) : ZGraphInstance<EdgeData>(g)
    {
        this->set_plan(Plan::of(ZG_PLAN_TRIANGLE_K4));
        this->channel = channel_;
    }
    // channel 1 by the split 4-clique plan with `completions` per closed triangle
    // (ZG_PLAN_TRIANGLE_K4_SPLIT)
    TriangleFourCliqueInstance(Graph<EdgeData> *g, int channel_, int completions) : ZGraphInstance<EdgeData>(g)
    {
        this->set_plan(Plan::split_k4(completions, ZG_PLAN_TRIANGLE_K4_SPLIT));
        this->channel = channel_;
    }
    // both counts from one pass over ids [0, N)
    std::pair<double, double> calculate_both(unsigned long long num_estimators)
    {
        this->require_plan();
        zg_sums s[2];
        zg_launch launch{0, 0, 0};
        check(zg_run(this->get_run_graph()->handle(), this->plan()->handle(), this->seed, 0, num_estimators, &launch,
                     s, nullptr));
        return {s[0].mean, s[1].mean};
    }
};

// applications/ChainMining.cpp:18-85 (order = vertices of the path)
template <typename EdgeData>
class ChainMiningInstance : public ZGraphInstance<EdgeData> {
  public:
    ChainMiningInstance(Graph<EdgeData> *g, int order) : ZGraphInstance<EdgeData>(g)
    {
        this->set_plan(Plan::of(ZG_PLAN_CHAIN, order));
    }
};

// applications/FourChain.cpp:18-125
template <typename EdgeData>
class FourChainCountingInstance : public ZGraphInstance<EdgeData> {
  public:
    explicit FourChainCountingInstance(Graph<EdgeData> *g) : ZGraphInstance<EdgeData>(g)
    {
        this->set_plan(Plan::of(ZG_PLAN_FOUR_CHAIN));
    }
};

// 4-cliques by three spanning shapes and closing checks (unbiased; the
// reference's applications/FourClique.cpp:17-77 finds 3/5 of each clique)
template <typename EdgeData>
class FourCliqueCountingInstance : public ZGraphInstance<EdgeData> {
  public:
    explicit FourCliqueCountingInstance(Graph<EdgeData> *g) : ZGraphInstance<EdgeData>(g)
    {
        this->set_plan(Plan::of(ZG_PLAN_K4_CLOSE));
    }
};

// 4-cliques, split: FourCliqueCountingInstance's first stage, then
// `completions` independent draws of its three shapes per closed triangle,
// averaged (same mean, lower variance per estimator; no reference counterpart)
template <typename EdgeData>
class FourCliqueSplitCountingInstance : public ZGraphInstance<EdgeData> {
  public:
    FourCliqueSplitCountingInstance(Graph<EdgeData> *g, int completions = 64) : ZGraphInstance<EdgeData>(g)
    {
        this->set_plan(Plan::split_k4(completions));
    }
};

// 5-houses (the pattern of applications/FiveHouse.cpp:18-27), split: a base
// edge and its roof apex, closed by one pair lookup, then `completions` draws
// of the square's two other vertices per closed roof, averaged -- the house
// count SamplerGenerator estimates, with far less variance per estimator (no
// reference counterpart)
template <typename EdgeData>
class FiveHouseSplitCountingInstance : public ZGraphInstance<EdgeData> {
  public:
    FiveHouseSplitCountingInstance(Graph<EdgeData> *g, int completions = 256) : ZGraphInstance<EdgeData>(g)
    {
        this->set_plan(Plan::split_k4(completions, ZG_PLAN_HOUSE_SPLIT));
    }
};

// SamplerGenerator<E>(graph, pattern) (src/SamplerGenerator.cpp:311-476)
template <typename EdgeData>
class SamplerGenerator : public ZGraphInstance<EdgeData> {
  public:
    SamplerGenerator(Graph<EdgeData> *g, Pattern *pattern) : ZGraphInstance<EdgeData>(g), pattern_(pattern)
    {
        this->set_plan(pattern->desc());
    }
    double get_num_sample_edge_calls() override { return 1.0; }                     // :466-469
    double get_num_conditional_sample_edge_calls() override                           // :471-474
    {
        return (double)pattern_->get_num_edges() - 1;
    }

  private:
    Pattern *pattern_;
};

// AllPredicateZGraphInstance (src/ZGraphInstance.cpp:490-525): the plan runs
// on the sub-graph whose edge data satisfy the predicate; get_graph()
// returns the original graph.  `plan_kind` is the estimator (triangles in
// applications/AllPredicateWeightedTriangle.cpp:48-71).
template <typename EdgeData>
class AllPredicateZGraphInstance : public ZGraphInstance<EdgeData> {
  public:
    template <typename Pred>
    AllPredicateZGraphInstance(Graph<EdgeData> *g, Pred predicate, int plan_kind = ZG_PLAN_TRIANGLE,
                               uint64_t filter_seed = 1)
        : ZGraphInstance<EdgeData>(g), filter_(nullptr, nullptr, predicate, filter_seed)
    {
        filter_.generate_graph(*g, filtered_);
        this->set_plan(Plan::of(plan_kind));
    }
    Graph<EdgeData> *get_graph() override { return this->graph; }
    void set_graph(Graph<EdgeData> *g) override
    {
        this->graph = g;
        filter_.generate_graph(*g, filtered_);
    }
    Graph<EdgeData> &filtered_graph() { return filtered_; }
    Graph<EdgeData> *get_run_graph() override { return &filtered_; }

  private:
    GraphFilter<EdgeData> filter_;
    Graph<EdgeData> filtered_;
};
template <typename EdgeData>
class AllPredicateWeightedTriangleCountingInstance : public AllPredicateZGraphInstance<EdgeData> {
  public:
    template <typename Pred>
    AllPredicateWeightedTriangleCountingInstance(Graph<EdgeData> *g, Pred p)
        : AllPredicateZGraphInstance<EdgeData>(g, p, ZG_PLAN_TRIANGLE)
    {
    }
};

// ---- UnbiasedLinearRegression (src/Regression.cpp:17-28) --------------------
class UnbiasedLinearRegression {   // y = k * x
  public:
    void solve(long double *x, long double *y, unsigned int len)
    {
        k_ = zg_regression_k(x, y, len);
        solved_ = true;
    }
    long double get_k() const
    {
        if (!solved_) throw Error(ZG_ESTATE, "UnbiasedLinearRegression: solve() first");
        return k_;
    }

  private:
    bool solved_ = false;
    long double k_ = 0;
};

// ---- TimeProfile / ErrorProfile (src/ErrorLatencyProfile.cpp:12-207) --------
template <typename EdgeData>
class TimeProfile {
  public:
    explicit TimeProfile(unsigned long long basic_estimators = 1ull << 16) : basic_(basic_estimators) {}
    // Doubling batches while used + 2 * last <= limit (ms), then one batch
    // sized to the remaining time (:24-47).  Batches take disjoint id ranges.
    double run(ZGraphInstance<EdgeData> &inst, VertexId num_vertices_in_pattern, unsigned int time_limit)
    {
        (void)num_vertices_in_pattern;
        Estimation acc;
        batches_.clear();
        unsigned long long used = 0, cur = 0, next = basic_, first = 0;
        while (used + cur * 2 <= time_limit) {
            const auto t0 = std::chrono::steady_clock::now();
            const zg_sums s = inst.run_ids(first, next);
            cur = (unsigned long long)std::chrono::duration_cast<std::chrono::milliseconds>(
                      std::chrono::steady_clock::now() - t0).count();
            used += cur;
            acc.absorb(s);
            acc.update(next, s.mean);
            batches_.push_back({first, next});
            first += next;
            next <<= 1;
        }
        if (cur > 0 && time_limit > used) {
            const unsigned long long rest = next * (time_limit - used) / (cur * 2);
            if (rest) {
                const zg_sums s = inst.run_ids(first, rest);
                acc.absorb(s);
                acc.update(rest, s.mean);
                batches_.push_back({first, rest});
            }
        }
        last_ = acc;
        return acc.get_currnet_estimation();
    }
    // the last run: its Estimation (running mean, exact sums) and its batches (first id, count)
    const Estimation &last_estimation() const { return last_; }
    const std::vector<std::pair<unsigned long long, unsigned long long>> &batches() const { return batches_; }

  private:
    Estimation last_;
    std::vector<std::pair<unsigned long long, unsigned long long>> batches_;
    unsigned long long basic_;
};

template <typename EdgeData>
class ErrorProfile {
  public:
    explicit ErrorProfile(int pilot_scale = 0) : pilot_scale_(pilot_scale) {}
    // ErrorProfile::run (:61-69): size with get_num_estimators, then calculate
    double run(ZGraphInstance<EdgeData> &inst, VertexId num_vertices_in_pattern, double required_error_rate,
               double required_confidence, Estimation &init_estimation)
    {
        const unsigned long long n =
            get_num_estimators(inst, num_vertices_in_pattern, required_error_rate, required_confidence);
        return inst.calculate(n, num_vertices_in_pattern, init_estimation);
    }
    // ErrorProfile::get_num_estimators (:71-207) on the device
    unsigned long long get_num_estimators(ZGraphInstance<EdgeData> &inst, VertexId num_vertices_in_pattern,
                                          double required_error_rate, double required_confidence)
    {
        (void)num_vertices_in_pattern;
        if (!inst.plan()) throw Error(ZG_EINVAL, "ErrorProfile: instance has no device plan");
        check(zg_elp_regression(inst.get_run_graph()->handle(), inst.plan()->handle(), required_error_rate,
                                required_confidence, inst.seed, pilot_scale_, &last));
        if (!(last.num_estimators >= 0) || !std::isfinite(last.num_estimators)) return 0;
        if (last.num_estimators >= 1.8e19) return ~0ull;
        return (unsigned long long)last.num_estimators;
    }
    // calculate_bound (:51-59)
    static double calculate_bound(double epsilon, double delta, double a, double b, double estimation,
                                  EdgeId max_degree, EdgeId num_edges)
    {
        return zg_elp_bound(epsilon, delta, a, b, estimation, max_degree, num_edges);
    }
    zg_elp_result last{};

  private:
    int pilot_scale_;
};

// CLT judge: the stopping rule of sampling/IEStop.cpp:47-56
inline bool judge(const Estimation &e, double eps, double confidence, unsigned long long min_sample = 1ull << 12)
{
    const double mu = e.exact_mean();
    if (e.get_num_finished_estimators() < min_sample || !(mu > 0)) return false;
    return e.half_width(confidence) <= eps * mu;
}

// Runs `n` estimators, then tops the same id sequence up until the judge
// accepts (at most max_rounds more runs); returns the exact mean.
template <typename EdgeData>
double calculate_to_error(ZGraphInstance<EdgeData> &inst, unsigned long long n, double eps, double confidence,
                          Estimation &est, int max_rounds = 8)
{
    inst.calculate(n, 0, est);
    for (int r = 0; r < max_rounds && !judge(est, eps, confidence); ++r) {
        const double mu = est.exact_mean(), hw = est.half_width(confidence);
        double grow = (mu > 0 && std::isfinite(hw) && hw > 0) ? (hw / (eps * mu)) * (hw / (eps * mu)) * 1.05 : 4.0;
        grow = std::min(std::max(grow, 1.25), 64.0);
        n = (unsigned long long)std::ceil((double)n * grow);
        inst.calculate(n, 0, est);
    }
    return est.exact_mean();
}

// ---- learned error-latency profile (no reference counterpart) ---------------
// Sizes a query from a pilot (the query's own first ids), graph features and
// the trained network, then runs it under the judge; the regression profile
// is the fallback.  Same procedure and features as zgraph/elp.py, so the two
// size a query identically.  Weights: the flat file written by
// zgraph.elp.export_binary (elp_weights.bin next to libzgraph_amd.so).
template <typename EdgeData>
class LearnedErrorProfile {
  public:
    explicit LearnedErrorProfile(const std::string &weights_bin, unsigned long long pilot = 1ull << 16,
                                 unsigned long long max_pilot = 1ull << 28)
        : pilot_(pilot), max_pilot_(max_pilot)
    {
        std::ifstream f(weights_bin, std::ios::binary);
        char magic[8];
        uint32_t hdr[2];
        if (!f || !f.read(magic, 8) || (std::memcmp(magic, "ZGELP001", 8) && std::memcmp(magic, "ZGELP002", 8)) ||
            !f.read((char *)hdr, 8))
            throw Error(ZG_EINVAL, "LearnedErrorProfile: not a weights file: " + weights_bin);
        const bool v2 = !std::memcmp(magic, "ZGELP002", 8);
        if ((int)hdr[0] != zg_mlp_num_params() || (int)hdr[1] != zg_mlp_input_dim())
            throw Error(ZG_EINVAL, "LearnedErrorProfile: weights do not match the network");
        mu_.resize(hdr[1]);
        sd_.resize(hdr[1]);
        params_.resize(hdr[0]);
        f.read((char *)mu_.data(), 4 * hdr[1]);
        f.read((char *)sd_.data(), 4 * hdr[1]);
        f.read((char *)use_, 4);
        f.read((char *)&y_scale_, 4);
        if (v2) f.read((char *)val_rmse_, 16);
        f.read((char *)params_.data(), 4 * (std::streamsize)hdr[0]);
        if (!f) throw Error(ZG_EINVAL, "LearnedErrorProfile: truncated weights file");
        check(zg_mlp_create(1, &net_));
        check(zg_mlp_set_params(net_, params_.data()));
    }
    ~LearnedErrorProfile()
    {
        if (net_) zg_mlp_destroy(net_);
    }
    LearnedErrorProfile(const LearnedErrorProfile &) = delete;
    LearnedErrorProfile &operator=(const LearnedErrorProfile &) = delete;

    // pattern slot of a plan kind, -1 when the network was not trained on it
    static int slot(int kind)
    {
        switch (kind) {
        case ZG_PLAN_TRIANGLE: return 0;
        case ZG_PLAN_K4_CLOSE: return 1;
        case ZG_PLAN_THREE_CHAIN: return 2;
        case ZG_PLAN_FOUR_CHAIN: return 3;
        default: return -1;
        }
    }

    // Runs the pilot into `est` and returns the learned budget (from the
    // pilot's own CV for a plan the network has no slot for), or 0 when the
    // pilot cannot size this query (too few successes).
    unsigned long long size(ZGraphInstance<EdgeData> &inst, double eps, double conf, Estimation &est)
    {
        if (!inst.plan()) throw Error(ZG_EINVAL, "LearnedErrorProfile: instance has no device plan");
        const int sl = slot(plan_kind(inst));
        unsigned long long n = pilot_;
        for (;;) {
            inst.calculate(n, 0, est);
            if (est.get_successes() >= 32 || n >= max_pilot_) break;
            n = std::min(n * 8, max_pilot_);
        }
        return budget(inst.get_run_graph(), sl, n, est, eps, conf);
    }

    // Both channels of a two-channel plan (3-motif; fused triangle + 4-clique)
    // to +-eps at `conf` over one id range: one pilot, each channel sized with
    // its own pattern slot (a 2^24 start where the network cannot size it), and
    // the judge on both (zgraph/elp.py estimate_both_to_error).  Returns the ids run.
    unsigned long long run_both(ZGraphInstance<EdgeData> &inst, double eps, double conf, Estimation &e0,
                                Estimation &e1, unsigned long long max_estimators = 1ull << 40, int max_rounds = 8)
    {
        if (!inst.plan()) throw Error(ZG_EINVAL, "LearnedErrorProfile: instance has no device plan");
        int k0, k1;
        switch (plan_kind(inst)) {
        case ZG_PLAN_THREE_MOTIF: k0 = ZG_PLAN_THREE_CHAIN; k1 = ZG_PLAN_TRIANGLE; break;
        case ZG_PLAN_TRIANGLE_K4: k0 = ZG_PLAN_TRIANGLE; k1 = ZG_PLAN_K4_CLOSE; break;
        case ZG_PLAN_TRIANGLE_K4_SPLIT: k0 = ZG_PLAN_TRIANGLE; k1 = ZG_PLAN_K4_SPLIT; break;
        default: throw Error(ZG_EINVAL, "LearnedErrorProfile::run_both: not a two-channel plan");
        }
        Graph<EdgeData> *g = inst.get_run_graph();
        unsigned long long done = 0;
        auto extend = [&](unsigned long long n) {
            if (n <= done) return;
            zg_sums s[2];
            zg_launch launch{0, 0, 0};
            check(zg_run(g->handle(), inst.plan()->handle(), inst.seed, done, n - done, &launch, s, nullptr));
            e0.absorb(s[0]);
            e0.update(n - done, s[0].mean);
            e1.absorb(s[1]);
            e1.update(n - done, s[1].mean);
            done = n;
        };
        unsigned long long n = pilot_;
        for (;;) {
            extend(n);
            if (std::min(e0.get_successes(), e1.get_successes()) >= 32 || n >= max_pilot_) break;
            n = std::min(n * 8, max_pilot_);
        }
        unsigned long long b[2] = {budget(g, slot(k0), n, e0, eps, conf), budget(g, slot(k1), n, e1, eps, conf)};
        for (auto &x : b)
            if (!x) x = std::max(1ull << 24, n);
        unsigned long long N = std::min(std::max(b[0], b[1]), max_estimators);
        extend(N);
        for (int r = 0; r < max_rounds && N < max_estimators && !(judge(e0, eps, conf) && judge(e1, eps, conf));
             ++r) {
            double grow = 1.25;
            for (const Estimation *e : {&e0, &e1}) {
                const double mu = e->exact_mean(), hw = e->half_width(conf);
                grow = std::max(grow, (mu > 0 && std::isfinite(hw) && hw > 0) ? (hw / (eps * mu)) * (hw / (eps * mu)) * 1.05
                                                                             : 4.0);
            }
            N = std::min((unsigned long long)std::ceil((double)N * std::min(grow, 64.0)), max_estimators);
            extend(N);
        }
        last_sized = std::max(b[0], b[1]);
        return done;
    }

    unsigned long long get_num_estimators(ZGraphInstance<EdgeData> &inst, VertexId k, double eps, double conf)
    {
        Estimation est;
        const unsigned long long n = size(inst, eps, conf, est);
        return n ? n : fallback_.get_num_estimators(inst, k, eps, conf);
    }

    // size, run and judge (top-ups) into init_estimation; returns the exact mean
    double run(ZGraphInstance<EdgeData> &inst, VertexId k, double eps, double conf, Estimation &init_estimation)
    {
        unsigned long long n = size(inst, eps, conf, init_estimation);
        if (!n) {
            try {
                n = fallback_.get_num_estimators(inst, k, eps, conf);
            } catch (const Error &) {
                n = 1ull << 24;   // sizing undefined: let the judge find the budget
            }
        }
        last_sized = n;
        return calculate_to_error(inst, n, eps, conf, init_estimation);
    }
    unsigned long long last_sized = 0;

  private:
    static int plan_kind(ZGraphInstance<EdgeData> &inst) { return inst.plan()->kind(); }
    // the learned budget of a pattern slot from the pilot's statistics (slot
    // -1: the pilot's CV alone, no network), or 0 when undefined
    unsigned long long budget(Graph<EdgeData> *g, int sl, unsigned long long n, const Estimation &est, double eps,
                              double conf)
    {
        const double t = (double)est.get_num_finished_estimators(), mean = est.exact_mean();
        if (est.get_successes() < 32 || !(mean > 0)) return 0;
        const double var = std::max(est.get_sumsq() / t - mean * mean, 0.0);
        const double cv = std::sqrt(var) / mean, succ = (double)est.get_successes() / t;
        if (!(cv > 0) || !std::isfinite(cv)) return 0;
        const double z = normal_quantile_two_sided(conf);
        if (sl < 0) {
            const double q = z * cv / eps;
            return std::max((unsigned long long)std::ceil(q * q), n);
        }
        float x[16];
        features(g, sl, (double)n, cv, succ, x);
        float y = 0;
        for (int i = 0; i < 16; ++i) x[i] = (x[i] - mu_[i]) / sd_[i];
        check(zg_mlp_forward(net_, x, 1, &y));
        // aim one validation RMSE of the correction below the predicted CV: a short run costs
        // one top-up of the id sequence, a long one its excess
        const double r = (use_[sl] ? (double)(y * y_scale_) : 0.0) - aim_sigmas_ * (double)val_rmse_[sl];
        const double q = z * cv * std::pow(10.0, r) / eps;
        return std::max((unsigned long long)std::ceil(q * q), n);
    }
    void features(Graph<EdgeData> *g, int sl, double pilot_n, double cv, double succ, float *x)
    {
        double s2 = 0, s3 = 0;
        check(zg_graph_degree_moments(g->handle(), &s2, &s3));
        const double n = (double)g->num_vertices, m = (double)g->num_local_edges, dmax = (double)g->max_degree;
        auto lg = [](double v) { return std::log10(std::max(v, 1e-30)); };
        for (int i = 0; i < 16; ++i) x[i] = 0.f;
        x[0] = (float)lg(n);
        x[1] = (float)lg(m);
        x[2] = (float)lg(2 * m / std::max(n, 1.0));
        x[3] = (float)lg(s2 / std::max(2 * m, 1.0));
        x[4] = (float)lg(s3 / std::max(s2, 1.0));
        x[5] = (float)lg(dmax);
        x[6 + sl] = 1.f;
        x[10] = (float)lg(cv);
        x[11] = (float)lg(succ);
        x[12] = (float)lg(pilot_n);
        x[13] = (float)lg(cv * cv * succ);
        x[14] = (float)lg(s2 * s2 / std::max(m * s3, 1.0));
    }
    unsigned long long pilot_, max_pilot_;
    std::vector<float> mu_, sd_, params_;
    unsigned char use_[4] = {1, 1, 1, 1};
    float y_scale_ = 1.f;   // the network predicts r / y_scale
    float val_rmse_[4] = {0.f, 0.f, 0.f, 0.f};   // the correction's RMSE on validation graphs, per slot
    double aim_sigmas_ = 1.0;
    zg_mlp *net_ = nullptr;
    ErrorProfile<EdgeData> fallback_;
};

// ---- completions of the split plans (no reference counterpart) -------------
// Completions for a split plan (ZG_PLAN_K4_SPLIT, the 4-clique channel of
// ZG_PLAN_TRIANGLE_K4_SPLIT, ZG_PLAN_HOUSE_SPLIT), from two pilots over the
// same ids (from first_id on, away from a query's own) with j_lo and j_hi
// completions.  The j_lo pilot starts at `pilot` ids and grows (at least 4x,
// more when far short) until it holds min_successes successes and its kernel
// runs min_ms milliseconds (a launch-independent time), or reaches max_pilot
// ids.  If a +-eps answer at `conf` with j_lo completions, projected from that
// pilot, would take less time than the j_hi pilot can (j_hi / j_lo times the
// j_lo pilot's), the choice is j_lo (eps <= 0 turns this off); otherwise the
// j_hi pilot runs the same ids.  Per estimator Var(J) = Vb + Vw / J
// and the kernel time t(J) = c0 + c1 J, so the time to a given error,
// proportional to Var(J) t(J), is least at J* = sqrt(Vw c0 / (Vb c1)); the
// choice is the nearest power of two in [1, 4096] (1 when Vw <= 0, 4096 when
// Vb <= 0 or c1 <= 0).  Same rule as zgraph.elp.choose_split_completions.
struct SplitChoice {
    int completions = 0;
    double j_star = 0, v_between = 0, v_within = 0, c0_ms_per_id = 0, c1_ms_per_id_per_completion = 0,
           seconds = 0;
    unsigned long long pilot_ids = 0, pilot_successes[2] = {0, 0};
    double pilot_var[2] = {0, 0};
    bool hi_pilot_skipped = false;
    double projected_ms = 0;
};

template <typename EdgeData>
SplitChoice choose_split_completions(Graph<EdgeData> *g, uint64_t seed = 1, int kind = ZG_PLAN_K4_SPLIT, int j_lo = 64,
                                     int j_hi = 1024, unsigned long long min_successes = 32, double min_ms = 1.0,
                                     unsigned long long pilot = 1ull << 18, unsigned long long max_pilot = 1ull << 30,
                                     uint64_t first_id = 1ull << 60, double eps = 0.01, double conf = 0.95)
{
    if (kind != ZG_PLAN_K4_SPLIT && kind != ZG_PLAN_TRIANGLE_K4_SPLIT && kind != ZG_PLAN_HOUSE_SPLIT)
        throw Error(ZG_EINVAL, "choose_split_completions: not a split plan kind");
    const auto t0 = std::chrono::steady_clock::now();
    const int ch = kind == ZG_PLAN_TRIANGLE_K4_SPLIT ? 1 : 0;
    Plan lo(Plan::split_k4(j_lo, kind)), hi(Plan::split_k4(j_hi, kind));
    zg_launch launch{0, 0, 0};
    zg_sums s[2];
    for (const Plan *p : {&lo, &hi})   // code objects loaded before timing
        check(zg_run(g->handle(), p->handle(), seed, first_id, 1ull << 12, &launch, s, nullptr));
    SplitChoice c;
    unsigned long long n = pilot;
    float ms = 0;
    for (;;) {
        check(zg_run(g->handle(), lo.handle(), seed, first_id, n, &launch, s, &ms));
        if ((s[ch].successes >= min_successes && ms >= min_ms) || n >= max_pilot) break;
        const double short_by = std::max((double)min_successes / (double)std::max<uint64_t>(s[ch].successes, 1),
                                         min_ms / std::max((double)ms, 1e-3));
        n = std::min((unsigned long long)((double)n * std::min(std::max(1.5 * short_by, 4.0), 256.0)), max_pilot);
    }
    double ms_per_id[2];
    auto take = [&](int i, const zg_sums &x, float t) {
        const double mean = x.mean;
        c.pilot_successes[i] = x.successes;
        c.pilot_var[i] = std::max(x.sumsq / (double)n - mean * mean, 0.0);
        ms_per_id[i] = t / (double)n;
    };
    take(0, s[ch], ms);
    c.pilot_ids = n;
    if (eps > 0 && s[ch].mean > 0) {
        const double z = normal_quantile_two_sided(conf), mu = s[ch].mean;
        c.projected_ms = z * z * c.pilot_var[0] / (eps * mu * eps * mu) * ms_per_id[0];
        if (c.projected_ms < (double)ms * j_hi / j_lo) {
            c.hi_pilot_skipped = true;
            c.completions = j_lo;
            c.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            return c;
        }
    }
    check(zg_run(g->handle(), hi.handle(), seed, first_id, n, &launch, s, &ms));
    take(1, s[ch], ms);
    c.v_within = (c.pilot_var[0] - c.pilot_var[1]) / (1.0 / j_lo - 1.0 / j_hi);
    c.v_between = c.pilot_var[1] - c.v_within / j_hi;
    c.c1_ms_per_id_per_completion = (ms_per_id[1] - ms_per_id[0]) / (j_hi - j_lo);
    c.c0_ms_per_id = ms_per_id[0] - c.c1_ms_per_id_per_completion * j_lo;
    if (c.v_within <= 0) c.j_star = 1;
    else if (c.v_between <= 0 || c.c1_ms_per_id_per_completion <= 0) c.j_star = 4096;
    else
        c.j_star = std::sqrt(c.v_within * std::max(c.c0_ms_per_id, 1e-12) /
                             (c.v_between * c.c1_ms_per_id_per_completion));
    const double l = std::nearbyint(std::log2(std::max(c.j_star, 1.0)));
    c.completions = (int)std::min(4096.0, std::max(1.0, std::exp2(l)));
    c.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return c;
}

}  // namespace zgraph

#ifndef ZGRAPH_NO_GLOBAL_NAMES
// the reference's names at global scope, for its application mains
using zgraph::AllPredicateWeightedTriangleCountingInstance;
using zgraph::AllPredicateZGraphInstance;
using zgraph::ChainMiningInstance;
using zgraph::CSRConstructor;
using zgraph::DistGraphLoader;
using zgraph::Edge;
using zgraph::EdgeId;
using zgraph::Empty;
using zgraph::ErrorProfile;
using zgraph::Estimation;
using zgraph::FourChainCountingInstance;
using zgraph::FourCliqueCountingInstance;
using zgraph::FourCliqueSplitCountingInstance;
using zgraph::FiveHouseSplitCountingInstance;
using zgraph::get_globally_consistent_seed;
using zgraph::Graph;
using zgraph::MyList;
using zgraph::SubGraph;
using zgraph::LearnedErrorProfile;
using zgraph::GraphDestroyer;
using zgraph::GraphFilter;
using zgraph::GraphSampler;
using zgraph::load_graph_meta_data;
using zgraph::NumaAwareGraphDestroyer;
using zgraph::Pattern;
using zgraph::RandomMapping;
using zgraph::SamplerGenerator;
using zgraph::ThreeChainCountingInstance;
using zgraph::ThreeMotifInstance;
using zgraph::TriangleFourCliqueInstance;
using zgraph::TimeProfile;
using zgraph::TriangleCoutingInstance;
using zgraph::UniformGraphPartitioner;
using zgraph::UnbiasedLinearRegression;
using zgraph::VertexId;
using zgraph::ZGraphInstance;
#endif

#endif
