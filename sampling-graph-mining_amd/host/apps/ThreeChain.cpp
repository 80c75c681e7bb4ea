// 3-chain (wedge) counting with a fixed budget (applications/ThreeChain.cpp:40-90).
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
        ThreeChainCountingInstance<Empty> inst(&graph);
        inst.seed = arg_seed(argc, argv);
        Estimation est;
        const double cnt = inst.calculate(arg_estimators(argc, argv, 1ull << 23), 3, est);
        Debug::get_instance()->print("estimated number of three_chains: ", (unsigned long long)cnt);
        report("3-chain", est.exact_mean(), est.get_num_finished_estimators());
    })
    return 0;
}
