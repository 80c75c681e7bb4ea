// Triangles with at least one edge of weight >= 50: all triangles minus the
// triangles whose edges all fail the predicate
// (applications/AtLeastOnePredicateWeightedTriangle.cpp:135-152).  With a
// budget in argv[2] both terms run that many estimators; without one, each
// runs under a TimeProfile of 500 ms as the reference does.
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
        TriangleCoutingInstance<VertexId> all(&graph);
        AllPredicateWeightedTriangleCountingInstance<VertexId> none(&graph, zgraph::below<VertexId>(50.f));
        all.seed = none.seed = arg_seed(argc, argv);
        double cnt;
        if (argc > 2) {
            const unsigned long long n = arg_estimators(argc, argv, 0);
            Estimation a, b;
            all.calculate(n, 3, a);
            none.calculate(n, 3, b);
            cnt = a.exact_mean() - b.exact_mean();
            report("at-least-one-predicate-triangle", cnt, n);
        } else {
            TimeProfile<VertexId> tp;
            const double a = tp.run(all, 3, 500);
            const unsigned long long na = tp.last_estimation().get_num_finished_estimators();
            const size_t ba = tp.batches().size();
            const double b = tp.run(none, 3, 500);
            const unsigned long long nb = tp.last_estimation().get_num_finished_estimators();
            cnt = a - b;
            // each profile's estimators: ids [0, n) in tp.batches().size() batches
            report("time-profile-all", a, na);
            report("time-profile-none", b, nb);
            std::printf("time-profile-batches %zu %zu\n", ba, tp.batches().size());
            report("at-least-one-predicate-triangle", cnt, 0);
        }
        Debug::get_instance()->print("estimated number of triangles: ", (long long)cnt);
    })
    return 0;
}
