// Triangles whose three edges all carry weight >= 50
// (applications/AllPredicateWeightedTriangle.cpp: u32 weights, predicate
// at :101-103), on the filtered sub-graph.
#include "AppCommon.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s graph-prefix [estimators [seed]]\n", argv[0]);
        return 2;
    }
    APP_MAIN_GUARD({
        Graph<VertexId> graph;
        load_app_graph(argv[1], graph);
        AllPredicateWeightedTriangleCountingInstance<VertexId> inst(&graph, zgraph::at_least<VertexId>(50.f));
        inst.seed = arg_seed(argc, argv);
        Estimation est;
        const double cnt = inst.calculate(arg_estimators(argc, argv, 1ull << 23), 3, est);
        Debug::get_instance()->print("estimated number of triangles: ", (unsigned long long)cnt);
        report("all-predicate-triangle", est.exact_mean(), est.get_num_finished_estimators());
    })
    return 0;
}
