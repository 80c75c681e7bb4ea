// AppCommon.hpp -- shared set-up of the application mains: read <graph>.csv
// and <graph>.biedgelist into a device graph, parse the optional arguments.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../ZGraph.hpp"
#include "../ZGraphCompat.hpp"

template <typename E>
void load_app_graph(const std::string &prefix, Graph<E> &graph)
{
    VertexId n = 0;
    load_graph_meta_data(prefix + ".csv", n);
    UniformGraphPartitioner partitioner(n);
    RandomMapping mapping(get_globally_consistent_seed(), n);
    CSRConstructor<E> csr;
    NumaAwareGraphDestroyer<E> destroyer;
    DistGraphLoader<E> loader(&partitioner, &mapping, &csr, &destroyer);
    loader.load_graph(prefix + ".biedgelist", n, graph);
    Debug::get_instance()->log("graph: ", graph.num_vertices, " vertices, ", graph.num_local_edges, " edges, max degree ",
                               graph.max_degree);
}

// argv[2]: number of estimators (default `def`), argv[3]: estimator seed
inline unsigned long long arg_estimators(int argc, char **argv, unsigned long long def)
{
    return argc > 2 ? std::strtoull(argv[2], nullptr, 10) : def;
}
inline uint64_t arg_seed(int argc, char **argv) { return argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 1; }

// one line the tests parse: "<name> <estimate> <estimators>"
inline void report(const char *name, double estimate, unsigned long long n)
{
    std::printf("%s %.17g %llu\n", name, estimate, n);
}

#define APP_MAIN_GUARD(...)                                               \
    try {                                                                 \
        __VA_ARGS__                                                       \
    } catch (const zgraph::Error &e) {                                    \
        std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());       \
        return 1;                                                         \
    }
