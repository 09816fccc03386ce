// gpu_clique <graph.g> <clique_size>
// Same arguments and result lines as the reference (gpu/gpu_clique.cu:573-640):
// "count %llu" and "Counting time cost: S.UUUUUU s.".  Edges are oriented by
// (degree, id) (degree_orientation_init) unless GRAPHSET_CLIQUE_ORIENTATION
// asks for the reference's reduce_edges_for_clique ("id") or for the same
// orientation with the vertices renumbered in (degree, id) order ("rank",
// degree_rank_orientation); the count is the same in every case.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>

#include "cli_common.h"
#include "dataloader.h"
#include "graph.h"
#include "pattern.h"
#include "schedule_IEP.h"
#include "timeinterval.h"

int main(int argc, char* argv[]) {
    if (argc != 3) {
        printf("usage: %s graph_file clique_size", argv[0]);
        return 0;
    }
    if (require_device() < 1) return 1;
    TimeInterval allTime, tmpTime;
    int pattern_size = atoi(argv[2]);
    if (pattern_size < 3 || pattern_size > GM_MAX_PATTERN) {
        fprintf(stderr, "clique size must be 3..%d\n", GM_MAX_PATTERN);
        return 1;
    }
    Pattern p(pattern_size);
    for (int i = 0; i < pattern_size; i++)
        for (int j = i + 1; j < pattern_size; j++) p.add_edge(i, j);
    auto t1 = std::chrono::system_clock::now();
    Graph* g = nullptr;
    DataLoader D;
    if (!D.fast_load(g, argv[1])) {
        printf("data load failure :-(\n");
        return 0;
    }
    const char* orient = getenv("GRAPHSET_CLIQUE_ORIENTATION");
    Graph* dag = nullptr;
    if (orient && strcmp(orient, "id") == 0) {
        reduce_edges_for_clique(*g);
        dag = g;
    } else if (orient && strcmp(orient, "rank") == 0) {
        degree_rank_orientation(g, dag);
    } else {
        degree_orientation_init(g, dag);
    }
    auto t2 = std::chrono::system_clock::now();
    printf("Load data success! time: %g seconds\n",
           std::chrono::duration_cast<std::chrono::microseconds>(t2 - t1).count() / 1.0e6);
    fflush(stdout);
    allTime.check();
    printf("pattern = \n");
    p.print();
    printf("max intersection size %d\n", VertexSet::max_intersection_size);
    try {
        Schedule_IEP schedule_iep(p.get_adj_mat_ptr(), pattern_size);
        schedule_iep.set_in_exclusion_optimize_redundancy(1);
        printf("basic prefix %d, total prefix %d\n", schedule_iep.get_basic_prefix_num(),
               schedule_iep.get_total_prefix_num());
        tmpTime.check();
        dag->device_graph();
        tmpTime.print("Prepare time cost");
        GpuRunInfo info;
        unsigned long long sum = gpu_clique_matching(dag, schedule_iep, &info);
        printf("count %llu\n", sum);
        print_seconds("Counting time cost", info.region_s);
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    allTime.print("Total time cost");
    if (dag != g) delete dag;
    delete g;
    return 0;
}
