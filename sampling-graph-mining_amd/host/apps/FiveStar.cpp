// 5-stars through the generic pattern sampler (applications/FiveStar.cpp):
// centre 0 joined to 1..4.
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
        Edge<Empty> e[4];
        for (int i = 0; i < 4; ++i) e[i] = Edge<Empty>{(EdgeId)i, 0u, (VertexId)(i + 1)};
        Pattern star(5, 4, e);
        SamplerGenerator<Empty> inst(&graph, &star);
        inst.seed = arg_seed(argc, argv);
        Estimation est;
        const double cnt = inst.calculate(arg_estimators(argc, argv, 1ull << 23), 5, est);
        Debug::get_instance()->print("estimated number of five_stars: ", (unsigned long long)cnt);
        report("5-star", est.exact_mean(), est.get_num_finished_estimators());
    })
    return 0;
}
