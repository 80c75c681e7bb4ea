// 4-clique counting with a fixed budget (default 2^23, the budget of
// applications/FourClique.cpp:110-114), unbiased three-shape plan.
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
        FourCliqueCountingInstance<Empty> inst(&graph);
        inst.seed = arg_seed(argc, argv);
        Estimation est;
        const double cnt = inst.calculate(arg_estimators(argc, argv, 1ull << 23), 4, est);
        Debug::get_instance()->print("estimated number of four_clique: ", (unsigned long long)cnt);
        report("4-clique", est.exact_mean(), est.get_num_finished_estimators());
    })
    return 0;
}
