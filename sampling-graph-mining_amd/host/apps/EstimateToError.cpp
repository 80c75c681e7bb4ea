// A query sized by the learned error-latency profile and judged to
// +-eps at 95 % confidence: argv[2] pattern (triangle, 4-clique, 3-chain,
// 4-chain, 4-clique-split -- the split 4-clique plan with 256 completions per
// closed triangle, sized from its pilot; 4-clique-split-auto -- the same with
// its completions chosen by two pilots; 5-house-split and 5-house-split-auto
// -- the split 5-house plan, 256 or chosen completions per closed roof; or the
// two-channel 3-motif,
// triangle+4-clique and triangle+4-clique-split, both channels to +-eps over
// one id range), argv[3] eps
// (default 0.01), argv[4] estimator seed, argv[5] the weights file (default:
// elp_weights.bin beside libzgraph_amd.so).
#include <cstring>

#include "AppCommon.hpp"

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s graph-prefix pattern [eps [seed [weights.bin]]]\n", argv[0]);
        return 2;
    }
    APP_MAIN_GUARD({
        Graph<Empty> graph;
        load_app_graph(argv[1], graph);
        const std::string pat = argv[2];
        const double eps = argc > 3 ? std::atof(argv[3]) : 0.01;
        const uint64_t seed = argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 1;
        std::string weights = argc > 5 ? argv[5] : "";
        if (weights.empty()) {
            const std::string self = argv[0];
            weights = self.substr(0, self.rfind('/') + 1) + "../../zgraph/elp_weights.bin";
        }
        if (pat == "3-motif" || pat == "triangle+4-clique" || pat == "triangle+4-clique-split") {
            std::unique_ptr<ZGraphInstance<Empty>> both;
            if (pat == "3-motif") both.reset(new ThreeMotifInstance<Empty>(&graph));
            else if (pat == "triangle+4-clique") both.reset(new TriangleFourCliqueInstance<Empty>(&graph));
            else both.reset(new TriangleFourCliqueInstance<Empty>(&graph, 0, 256));
            both->seed = seed;
            LearnedErrorProfile<Empty> profile(weights);
            Estimation e0, e1;
            const unsigned long long n = profile.run_both(*both, eps, 0.95, e0, e1);
            const char *names[2] = {pat == "3-motif" ? "3-chain" : "triangle", pat == "3-motif" ? "triangle" : "4-clique"};
            std::printf("sized %llu\n", profile.last_sized);
            const Estimation *es[2] = {&e0, &e1};
            for (int c = 0; c < 2; ++c) {
                std::printf("judged %s %d %.17g\n", names[c], zgraph::judge(*es[c], eps, 0.95) ? 1 : 0,
                            es[c]->half_width(0.95) / es[c]->exact_mean());
                report(names[c], es[c]->exact_mean(), n);
            }
            return 0;
        }
        std::unique_ptr<ZGraphInstance<Empty>> inst;
        if (pat == "triangle") inst.reset(new TriangleCoutingInstance<Empty>(&graph));
        else if (pat == "4-clique") inst.reset(new FourCliqueCountingInstance<Empty>(&graph));
        else if (pat == "3-chain") inst.reset(new ThreeChainCountingInstance<Empty>(&graph));
        else if (pat == "4-chain") inst.reset(new FourChainCountingInstance<Empty>(&graph));
        else if (pat == "4-clique-split") inst.reset(new FourCliqueSplitCountingInstance<Empty>(&graph, 256));
        else if (pat == "5-house-split") inst.reset(new FiveHouseSplitCountingInstance<Empty>(&graph, 256));
        else if (pat == "4-clique-split-auto" || pat == "5-house-split-auto") {
            const bool house = pat == "5-house-split-auto";
            const zgraph::SplitChoice c = zgraph::choose_split_completions(
                &graph, seed, house ? ZG_PLAN_HOUSE_SPLIT : ZG_PLAN_K4_SPLIT, 64, 1024, 32, 1.0, 1ull << 18,
                1ull << 30, 1ull << 60, eps, 0.95);
            std::printf("completions %d %d %.17g %.17g %llu\n", c.completions, c.hi_pilot_skipped ? 1 : 0,
                        c.v_between, c.v_within, c.pilot_ids);
            if (house) inst.reset(new FiveHouseSplitCountingInstance<Empty>(&graph, c.completions));
            else inst.reset(new FourCliqueSplitCountingInstance<Empty>(&graph, c.completions));
        }
        else throw zgraph::Error(ZG_EINVAL, "unknown pattern " + pat);
        inst->seed = seed;
        LearnedErrorProfile<Empty> profile(weights);
        Estimation est;
        const double cnt = profile.run(*inst, 0, eps, 0.95, est);
        Debug::get_instance()->print("estimated number of ", pat, ": ", (unsigned long long)cnt);
        std::printf("sized %llu\n", profile.last_sized);
        std::printf("judged %d %.17g\n", zgraph::judge(est, eps, 0.95) ? 1 : 0, est.half_width(0.95) / est.exact_mean());
        report(pat.c_str(), est.exact_mean(), est.get_num_finished_estimators());
    })
    return 0;
}
