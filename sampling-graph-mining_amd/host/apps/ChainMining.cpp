// Paths of `order` vertices (applications/ChainMining.cpp:87-140): argv[4]
// is the order (default 4).
#include "AppCommon.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s graph-prefix [estimators [seed [order]]]\n", argv[0]);
        return 2;
    }
    APP_MAIN_GUARD({
        Graph<Empty> graph;
        load_app_graph(argv[1], graph);
        const int order = argc > 4 ? std::atoi(argv[4]) : 4;
        ChainMiningInstance<Empty> inst(&graph, order);
        inst.seed = arg_seed(argc, argv);
        Estimation est;
        const double cnt = inst.calculate(arg_estimators(argc, argv, 1ull << 23), (VertexId)order, est);
        Debug::get_instance()->print("estimated number of chains: ", (unsigned long long)cnt);
        report("chain", est.exact_mean(), est.get_num_finished_estimators());
    })
    return 0;
}
