// test_host_api.cpp -- host-side checks of the C++ API (no GPU calls), in
// the style of the reference's tests/Test*.cpp programs: each prints its
// result and the program exits non-zero on the first mismatch.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "../ZGraph.hpp"

static int failures = 0;
#define EXPECT(cond)                                                        \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                     \
        }                                                                   \
    } while (0)

// tests/TestRegression.cpp: k^2 of the through-origin fit is 609.8565
static void test_regression()
{
    const double xs[7] = {0.09, 0.26, 0.15, 0.42, 0.57, 0.10, 2.62};
    const double ys[7] = {1, 2, 4, 8, 16, 32, 64};
    long double x[7], y[7];
    for (int i = 0; i < 7; ++i) {
        x[i] = xs[i] / 100;
        y[i] = ys[i] / 100;
    }
    UnbiasedLinearRegression r;
    r.solve(x, y, 7);
    char buf[32];
    std::snprintf(buf, sizeof buf, "%.4f", (double)(r.get_k() * r.get_k()));
    std::printf("regression k^2 = %s\n", buf);
    EXPECT(std::string(buf) == "609.8565");
}

static Pattern make(VertexId k, std::initializer_list<std::pair<VertexId, VertexId>> es,
                    std::vector<Edge<Empty>> &store)
{
    store.clear();
    EdgeId i = 0;
    for (auto &e : es) store.push_back(Edge<Empty>{i++, e.first, e.second});
    return Pattern(k, store.size(), store.data());
}

// tests/TestPattern2.cpp: unique relabelings / k!
static void test_patterns()
{
    std::vector<Edge<Empty>> s;
    auto chain = make(4, {{0, 1}, {1, 2}, {2, 3}}, s);
    auto c = chain.unique_relabelings();
    std::printf("4-chain %d/%d\n", c.first, c.second);
    EXPECT(c.first == 12 && c.second == 24);
    auto k4 = make(4, {{1, 0}, {2, 0}, {2, 1}, {3, 0}, {3, 1}, {3, 2}}, s);
    c = k4.unique_relabelings();
    std::printf("4-clique %d/%d\n", c.first, c.second);
    EXPECT(c.first == 1 && c.second == 24);
    auto star = make(5, {{0, 1}, {0, 2}, {0, 3}, {0, 4}}, s);
    c = star.unique_relabelings();
    std::printf("5-star %d/%d\n", c.first, c.second);
    EXPECT(c.first == 5 && c.second == 120);
    auto house = make(5, {{0, 1}, {0, 2}, {1, 2}, {1, 3}, {2, 4}, {3, 4}}, s);
    c = house.unique_relabelings();
    std::printf("5-house %d/%d\n", c.first, c.second);
    EXPECT(c.first == 60 && c.second == 120);
    std::vector<Edge<Empty>> s2;
    auto chain_rev = make(4, {{3, 2}, {2, 1}, {1, 0}}, s2);
    EXPECT(chain.is_automorphic(chain_rev));
    std::vector<Edge<Empty>> s3;
    auto other = make(4, {{0, 1}, {1, 3}, {2, 3}}, s3);
    EXPECT(!chain.is_automorphic(other));
}

// src/ZGraphInstance.cpp:15-35: estimator-weighted running mean
static void test_estimation()
{
    Estimation e;
    e.update(3, 2.0);
    e.update(1, 6.0);
    EXPECT(e.get_num_finished_estimators() == 4);
    EXPECT(std::fabs(e.get_currnet_estimation() - 3.0) < 1e-15);
    e.reset();
    EXPECT(e.get_num_finished_estimators() == 0 && e.get_currnet_estimation() == 0.0);
    // exact sums: values {0, 4, 8, 0} at log2_den 1 -> sum 24, mean 3
    zg_sums s{};
    s.sum[0] = 24;
    s.log2_den = 1;
    s.count = 4;
    s.successes = 2;
    s.sumsq = 16 + 64;
    s.mean = 3;
    e.absorb(s);
    e.update(4, 3.0);
    EXPECT(e.exact_mean() == 3.0);
    const double hw = e.half_width(0.95);
    const double expect = 1.959963984540054 * std::sqrt((80.0 / 4 - 9) * 4 / 3 / 4);
    EXPECT(std::fabs(hw - expect) < 1e-9);
    EXPECT(std::fabs(zgraph::normal_quantile_two_sided(0.95) - 1.959963984540054) < 1e-9);
}

// tests/TestSubgraph.cpp: {01, 02, 12} - {01, 02} = {12}
static void test_subgraph()
{
    SubGraph<Empty> a, b;
    a.add_edge(Edge<Empty>{4, 0, 1});
    a.add_edge(Edge<Empty>{9, 0, 2});
    a.add_edge(Edge<Empty>{2, 1, 2});
    b.add_edge(Edge<Empty>{0, 1, 0});   // either orientation matches
    b.add_edge(Edge<Empty>{1, 0, 2});
    EXPECT(a.get_min_valid_edge_id() == 10);
    EXPECT(a.get_vertices().list_size == 3 && a.get_vertices().list[0] == 0 && a.get_vertices().list[1] == 1 &&
           a.get_vertices().list[2] == 2);
    const SubGraph<Empty> d = a - b;
    std::printf("subgraph difference: %d edge(s), first %u-%u\n", d.get_edges().list_size, d.get_edges().list[0].src,
                d.get_edges().list[0].dst);
    EXPECT(d.get_edges().list_size == 1 && d.get_edges().list[0].src == 1 && d.get_edges().list[0].dst == 2);
    EXPECT(d.get_min_valid_edge_id() == 3 && d.get_vertices().list_size == 2);
}

// src/ErrorLatencyProfile.cpp:51-59
static void test_bound()
{
    const double b = ErrorProfile<Empty>::calculate_bound(0.05, 0.05, 1, 1, 1000.0, 50, 10000);
    const double expect = std::pow(0.05, -2) * std::pow(50.0, 1) * std::pow(10000.0, 1) / 1000.0 * std::log(2 / 0.05);
    std::printf("bound = %.6g (closed form %.6g)\n", b, expect);
    EXPECT(std::fabs(b / expect - 1) < 1e-12);
}

// src/Utilities.cpp:24-35 and the loader's size check (src/Graph.cpp:400-402)
static void test_files(const std::string &dir)
{
    const std::string meta = dir + "/meta_test.csv";
    {
        std::ofstream f(meta);
        f << "num_vertices,num_edges\n7115,103689\n";
    }
    VertexId n = 0;
    load_graph_meta_data(meta, n);
    EXPECT(n == 7115);
    const std::string bad = dir + "/bad.biedgelist";
    {
        std::ofstream f(bad, std::ios::binary);
        const char junk[10] = {0};
        f.write(junk, 10);   // not a multiple of the 8-byte edge unit
    }
    UniformGraphPartitioner p(10);
    RandomMapping m(1, 10);
    CSRConstructor<Empty> c;
    NumaAwareGraphDestroyer<Empty> d;
    DistGraphLoader<Empty> loader(&p, &m, &c, &d);
    Graph<Empty> g;
    int code = 0;
    try {
        loader.load_graph(bad, 10, g);
    } catch (const zgraph::Error &e) {
        code = e.code();
    }
    EXPECT(code == ZG_EINVAL);
}

// a user estimate() subclass has no device plan: calculate fails loudly
struct MyEstimator : ZGraphInstance<Empty> {
    double estimate() override { return 1.0; }
};
static void test_no_plan()
{
    MyEstimator m;
    Estimation e;
    int code = 0;
    try {
        m.calculate(10, 3, e);
    } catch (const zgraph::Error &err) {
        code = err.code();
    }
    EXPECT(code == ZG_EINVAL);
}

// plan descriptors are host logic: channel counts and conditional-call
// constants of the plans; an unknown kind is rejected
static void test_plans()
{
    zgraph::Plan tri(zgraph::Plan::of(ZG_PLAN_TRIANGLE)), motif(zgraph::Plan::of(ZG_PLAN_THREE_MOTIF)),
        fused(zgraph::Plan::of(ZG_PLAN_TRIANGLE_K4));
    EXPECT(tri.channels() == 1 && motif.channels() == 2 && fused.channels() == 2);
    EXPECT(fused.cond_calls() == 2.0 && tri.cond_calls() == 1.0);
    int code = 0;
    try {
        zgraph::Plan bad(zgraph::Plan::of(9));
    } catch (const zgraph::Error &err) {
        code = err.code();
    }
    EXPECT(code == ZG_EINVAL);
}

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    test_regression();
    test_patterns();
    test_subgraph();
    test_estimation();
    test_bound();
    test_files(dir);
    test_no_plan();
    test_plans();
    std::printf("%s\n", failures ? "FAILED" : "OK");
    return failures ? 1 : 0;
}
