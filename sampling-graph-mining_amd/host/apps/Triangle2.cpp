// Triangles through the generic pattern sampler (applications/Triangle2.cpp
// and FourClique2.cpp: SamplerGenerator over the triangle pattern), sized by
// the regression error-latency profile (+-5 % at 95 %) or a fixed budget.
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
        Edge<Empty> e[3] = {{0, 0, 1}, {1, 0, 2}, {2, 1, 2}};
        Pattern triangle(3, 3, e);
        SamplerGenerator<Empty> inst(&graph, &triangle);
        inst.seed = arg_seed(argc, argv);
        Estimation est;
        double cnt;
        if (argc > 2) {
            cnt = inst.calculate(arg_estimators(argc, argv, 0), 3, est);
        } else {
            ErrorProfile<Empty> profile;
            cnt = profile.run(inst, 3, 0.05, 0.95, est);
        }
        Debug::get_instance()->print("estimated number of triangles: ", (unsigned long long)cnt);
        report("triangle-pattern", est.exact_mean(), est.get_num_finished_estimators());
    })
    return 0;
}
