// Triangles and 4-cliques from the same sampled first edge, conditional sample
// and closing check per estimator id: the cached sub-pattern sharing of
// applications/ThreeMotif.cpp (src/ZGraphInstance.cpp:596-833) applied to the
// triangle -> 4-clique pair.  Each count equals what Triangle and FourClique
// report for the same estimators and seed.
#include "AppCommon.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s graph-prefix [estimators [seed]]\n", argv[0]);
        return 2;
    }
    APP_MAIN_GUARD({
        Graph<Empty> graph;
        load_app_graph(argv[1], graph);
        TriangleFourCliqueInstance<Empty> inst(&graph);
        inst.seed = arg_seed(argc, argv);
        const unsigned long long n = arg_estimators(argc, argv, 1ull << 23);
        const auto both = inst.calculate_both(n);
        Debug::get_instance()->print("estimated number of triangles: ", (unsigned long long)both.first);
        Debug::get_instance()->print("estimated number of 4-cliques: ", (unsigned long long)both.second);
        report("triangle", both.first, n);
        report("4-clique", both.second, n);
    })
    return 0;
}
