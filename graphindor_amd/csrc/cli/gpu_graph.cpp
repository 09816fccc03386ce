// gpu_graph <graph.g> <pattern_size> <pattern_adj_string> [1|0 iep]
// Same arguments and stdout lines as the reference (gpu/gpu_graph.cu:53-114).
// With GRAPHSET_COST_MODEL=<file> the learned cost model proposes schedules
// among the analytic selector's candidates and a short sampled run of each,
// and of the analytic pick, decides (the analytic pick if the file is
// unusable); the choice is reported on stderr.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <memory>

#include "cli_common.h"
#include "costmodel.h"
#include "dataloader.h"
#include "graph.h"
#include "pattern.h"
#include "schedule_IEP.h"
#include "timeinterval.h"

int main(int argc, char* argv[]) {
    if (argc < 4) {
        fprintf(stderr, "Usage: %s graph_file pattern_size pattern_string <1/0 for enable iep or not>\n", argv[0]);
        return 1;
    }
    int enable_iep = 1;
    if (argc >= 5) {
        enable_iep = atoi(argv[4]);
        if (enable_iep != 0 && enable_iep != 1) {
            fprintf(stderr, "Usage: %s graph_file pattern_size pattern_string <1/0 for enable iep or not>\n", argv[0]);
            return 1;
        }
        fprintf(stderr, "Enable iep: %d\n", enable_iep);
    }
    if (require_device() < 1) return 1;
    TimeInterval allTime, tmpTime;
    auto t1 = std::chrono::system_clock::now();
    Graph* g = nullptr;
    DataLoader D;
    if (!D.fast_load(g, argv[1])) {
        fprintf(stderr, "data load failure :-(\n");
        return 0;
    }
    auto t2 = std::chrono::system_clock::now();
    fprintf(stderr, "Load data success! time: %g seconds\n",
            std::chrono::duration_cast<std::chrono::microseconds>(t2 - t1).count() / 1.0e6);
    allTime.check();
    int pattern_size = atoi(argv[2]);
    if (pattern_size < 3 || pattern_size > GM_MAX_PATTERN) {
        fprintf(stderr, "pattern size must be 3..%d\n", GM_MAX_PATTERN);
        return 1;
    }
    Pattern p(pattern_size, argv[3]);
    printf("pattern = ");
    p.print();
    fprintf(stderr, "Max intersection size %d\n", VertexSet::max_intersection_size);
    tmpTime.check();
    try {
        bool is_pattern_valid = true;
        std::unique_ptr<Schedule_IEP> owned;
        const char* model_path = getenv("GRAPHSET_COST_MODEL");
        if (model_path && *model_path) {
            std::unique_ptr<CostModel> model = CostModel::load(model_path);
            if (!model) fprintf(stderr, "cost model %s unusable: analytic schedule\n", model_path);
            ScheduleChoice ch;
            owned.reset(select_schedule_measured(p, *g, model.get(), enable_iep != 0, MeasureOptions(), &ch));
            fprintf(stderr, "Schedule: %s pick, candidate %d of %d (analytic pick %d), %.3f ms\n",
                    ch.learned ? "learned" : "analytic", ch.chosen, ch.candidates, ch.analytic, ch.select_ms);
        } else {
            owned.reset(new Schedule_IEP(p, is_pattern_valid, 1, 1, enable_iep != 0, g->v_cnt, g->e_cnt, g->tri_cnt));
        }
        if (!is_pattern_valid) {
            fprintf(stderr, "pattern is invalid!\n");
            return 1;
        }
        const Schedule_IEP& schedule_iep = *owned;
        tmpTime.print("Schedule time cost");
        tmpTime.check();
        g->device_graph();
        tmpTime.print("Prepare time cost");
        GpuRunInfo info;
        unsigned long long sum = gpu_pattern_matching(g, schedule_iep, &info);
        printf("Pattern count: %llu\n", sum);
        print_seconds("Counting time cost", info.region_s);
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    allTime.print("Total time cost");
    delete g;
    return 0;
}
