// 3-chains and triangles from two independent instances, each with its own
// budget (applications/ThreeMotifWithoutCache.cpp:120-150): argv[2] the
// 3-chain budget, argv[3] the triangle budget, argv[4] the seed.
#include "AppCommon.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s graph-prefix [chain-estimators [triangle-estimators [seed]]]\n", argv[0]);
        return 2;
    }
    APP_MAIN_GUARD({
        Graph<Empty> graph;
        load_app_graph(argv[1], graph);
        const unsigned long long nc = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1ull << 23;
        const unsigned long long nt = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 1ull << 23;
        const uint64_t seed = argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 1;
        ThreeChainCountingInstance<Empty> chains(&graph);
        TriangleCoutingInstance<Empty> triangles(&graph);
        chains.seed = triangles.seed = seed;
        Estimation ec, et;
        const double c = chains.calculate(nc, 3, ec);
        const double t = triangles.calculate(nt, 3, et);
        Debug::get_instance()->print("estimated number of three_chains: ", (unsigned long long)c);
        Debug::get_instance()->print("estimated number of triangles: ", (unsigned long long)t);
        report("3-chain", ec.exact_mean(), ec.get_num_finished_estimators());
        report("triangle", et.exact_mean(), et.get_num_finished_estimators());
    })
    return 0;
}
