// Triangle counting sized by the regression error-latency profile
// (the flow of applications/Triangle.cpp:75-140): ±5 % at 95 % confidence,
// or a fixed budget when argv[2] is given.
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
        TriangleCoutingInstance<Empty> inst(&graph);
        inst.seed = arg_seed(argc, argv);
        Estimation est;
        Timer::timer_start("estimation");
        double cnt;
        if (argc > 2) {
            cnt = inst.calculate(arg_estimators(argc, argv, 0), 3, est);
        } else {
            ErrorProfile<Empty> profile;
            cnt = profile.run(inst, 3, 0.05, 0.95, est);
        }
        Timer::timer_stop("estimation");
        Debug::get_instance()->print("estimated number of triangles: ", (unsigned long long)cnt);
        report("triangle", est.exact_mean(), est.get_num_finished_estimators());
    })
    return 0;
}
