// gpu_mc <graph.g> <motif_size>
// Same arguments and stdout lines as the reference (gpu/gpu_mc.cu:62-117):
// per connected motif "pattern = ..." and "Pattern count: N", then
// "Counting time cost: %.6lf s" summed over the motifs.  The graph is
// uploaded once and every motif's plan counted by one gm_count_multi call.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <memory>
#include <vector>

#include "cli_common.h"
#include "dataloader.h"
#include "gpu_plan.h"
#include "graph.h"
#include "motif_generator.h"
#include "schedule_IEP.h"
#include "timeinterval.h"

int main(int argc, char* argv[]) {
    if (argc < 3) {
        fprintf(stderr, "Usage: %s graph_file pattern_size\n", argv[0]);
        return 1;
    }
    if (require_device() < 1) return 1;
    TimeInterval allTime;
    auto t1 = std::chrono::system_clock::now();
    Graph* g = nullptr;
    DataLoader D;
    if (!D.fast_load(g, argv[1])) {
        fprintf(stderr, "data load failure :-(\n");
        return 1;
    }
    auto t2 = std::chrono::system_clock::now();
    fprintf(stderr, "Load data success! time: %g seconds\n",
            std::chrono::duration_cast<std::chrono::microseconds>(t2 - t1).count() / 1.0e6);
    allTime.check();
    int pattern_size = atoi(argv[2]);
    if (pattern_size < 3 || pattern_size > GM_MAX_PATTERN) {
        fprintf(stderr, "motif size must be 3..%d\n", GM_MAX_PATTERN);
        return 1;
    }
    printf("motif_size: %d\n", pattern_size);
    MotifGenerator mg(pattern_size);
    std::vector<Pattern> motifs = mg.generate();
    printf("Motifs number = %d\n", (int)motifs.size());
    try {
        std::vector<std::unique_ptr<Schedule_IEP> > scheds;
        std::vector<FlatPlan> flats(motifs.size());
        std::vector<gm_plan*> plans;
        for (size_t i = 0; i < motifs.size(); ++i) {
            bool ok = true;
            scheds.emplace_back(new Schedule_IEP(motifs[i], ok, 1, 1, true, g->v_cnt, g->e_cnt, g->tri_cnt));
            if (!ok) {
                fprintf(stderr, "pattern is invalid!\n");
                return 1;
            }
            flatten_schedule(*scheds.back(), GM_MODE_PATTERN, flats[i]);
            gm_plan* pl = nullptr;
            if (gm_plan_create(&flats[i].desc, &pl) != GM_OK) {
                fprintf(stderr, "gm_plan_create: %s\n", gm_last_error());
                return 1;
            }
            plans.push_back(pl);
        }
        gm_graph* dg = g->device_graph();
        std::vector<uint64_t> raw(plans.size());
        gm_stats st{};
        auto c0 = std::chrono::steady_clock::now();
        if (gm_count_multi(dg, plans.data(), (int32_t)plans.size(), nullptr, 0, nullptr, raw.data(), &st) != GM_OK) {
            fprintf(stderr, "gm_count_multi: %s\n", gm_last_error());
            return 1;
        }
        auto c1 = std::chrono::steady_clock::now();
        for (size_t i = 0; i < motifs.size(); ++i) {
            printf("pattern = ");
            motifs[i].print();
            printf("Pattern count: %llu\n",
                   (unsigned long long)(raw[i] / (uint64_t)scheds[i]->get_in_exclusion_optimize_redundancy()));
            gm_plan_destroy(plans[i]);
        }
        printf("Counting time cost: %.6lf s\n", std::chrono::duration<double>(c1 - c0).count());
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    allTime.print("Total time cost");
    delete g;
    return 0;
}
