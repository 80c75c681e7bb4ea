// 5-vertex houses (the square 1-3-4-2 under the roof 1-0-2) through the
// generic pattern sampler (applications/FiveHouse.cpp).
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
        const VertexId ends[6][2] = {{0, 1}, {0, 2}, {1, 2}, {1, 3}, {2, 4}, {3, 4}};
        Edge<Empty> e[6];
        for (int i = 0; i < 6; ++i) e[i] = Edge<Empty>{(EdgeId)i, ends[i][0], ends[i][1]};
        Pattern house(5, 6, e);
        SamplerGenerator<Empty> inst(&graph, &house);
        inst.seed = arg_seed(argc, argv);
        Estimation est;
        const double cnt = inst.calculate(arg_estimators(argc, argv, 1ull << 23), 5, est);
        Debug::get_instance()->print("estimated number of five_houses: ", (unsigned long long)cnt);
        report("5-house", est.exact_mean(), est.get_num_finished_estimators());
    })
    return 0;
}
