// 3-motifs: 3-chains and triangles from the same sampled wedges
// (applications/ThreeMotif.cpp:125-210, cached sub-patterns fused on the device).
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
        ThreeMotifInstance<Empty> inst(&graph);
        inst.seed = arg_seed(argc, argv);
        const unsigned long long n = arg_estimators(argc, argv, 1ull << 23);
        const auto both = inst.calculate_both(n);
        Debug::get_instance()->print("estimated number of three_chains: ", (unsigned long long)both.first);
        Debug::get_instance()->print("estimated number of triangles: ", (unsigned long long)both.second);
        report("3-chain", both.first, n);
        report("triangle", both.second, n);
    })
    return 0;
}
