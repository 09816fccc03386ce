// Graph and the device drivers (reference src/graph.cpp; gpu/gpu_graph.cu:22-51,
// gpu/gpu_mc.cu:29-59, gpu/gpu_clique.cu:366-570).
#include "graph.h"

#include <algorithm>
#include <chrono>
#include <vector>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "generator.h"
#include "gpu_plan.h"
#include "graphindor_amd.h"
#include "motif_generator.h"
#include "timeinterval.h"

int VertexSet::max_intersection_size = -1;

namespace {

void check(int rc, const char* what) {
    if (rc != GM_OK) throw std::runtime_error(std::string(what) + ": " + gm_last_error());
}

unsigned long long run_device(Graph* g, const Schedule_IEP& s, int mode, GpuRunInfo* info) {
    FlatPlan fp;
    flatten_schedule(s, mode, fp);
    gm_plan* plan = nullptr;
    check(gm_plan_create(&fp.desc, &plan), "gm_plan_create");
    gm_graph* dg = g->device_graph();
    uint64_t raw = 0;
    gm_stats st{};
    auto t0 = std::chrono::steady_clock::now();
    int rc = gm_count(dg, plan, nullptr, 0, nullptr, &raw, &st);
    auto t1 = std::chrono::steady_clock::now();
    gm_plan_destroy(plan);
    check(rc, "gm_count");
    if (info) {
        info->raw_sum = raw;
        info->kernel_ms = st.kernel_ms;
        info->region_s = std::chrono::duration<double>(t1 - t0).count();
        info->alg_bytes = st.alg_bytes;
    }
    // Same unsigned division as gpu/gpu_graph.cu:44.
    return raw / (unsigned long long)s.get_in_exclusion_optimize_redundancy();
}

}  // namespace

void flatten_schedule(const Schedule_IEP& s, int mode, FlatPlan& fp) {
    const int n = s.get_size();
    const int P = n * (n - 1) / 2 + 1;
    fp.adj.assign(s.get_adj_mat_ptr(), s.get_adj_mat_ptr() + n * n);
    fp.father.assign(s.get_father_prefix_id_ptr(), s.get_father_prefix_id_ptr() + P);
    fp.last.assign(s.get_last_ptr(), s.get_last_ptr() + n);
    fp.next.assign(s.get_next_ptr(), s.get_next_ptr() + P);
    fp.brk.assign(s.get_break_size_ptr(), s.get_break_size_ptr() + P);
    fp.loopset.assign(s.get_loop_set_prefix_id_ptr(), s.get_loop_set_prefix_id_ptr() + n);
    fp.rlast.assign(s.get_restrict_last_ptr(), s.get_restrict_last_ptr() + n);
    fp.rnext.assign(s.get_restrict_next_ptr(), s.get_restrict_next_ptr() + P);
    fp.rindex.assign(s.get_restrict_index_ptr(), s.get_restrict_index_ptr() + P);
    fp.only.resize(P);
    for (int i = 0; i < P; ++i) fp.only[i] = s.get_prefix_only_need_size(i) ? 1 : 0;
    const size_t I = s.in_exclusion_optimize_vertex_id.size();
    const size_t A = s.in_exclusion_optimize_coef.size();
    fp.iep_id.assign(s.in_exclusion_optimize_vertex_id.begin(), s.in_exclusion_optimize_vertex_id.end());
    fp.iep_vflag.resize(I);
    fp.iep_vcoef.resize(I);
    for (size_t i = 0; i < I; ++i) {
        fp.iep_vflag[i] = s.in_exclusion_optimize_vertex_flag[i] ? 1 : 0;
        fp.iep_vcoef[i] = s.in_exclusion_optimize_vertex_coef[i];
    }
    fp.iep_coef.assign(s.in_exclusion_optimize_coef.begin(), s.in_exclusion_optimize_coef.end());
    fp.iep_flag.resize(A);
    fp.iep_pos.resize(A);
    for (size_t i = 0; i < A; ++i) {
        fp.iep_flag[i] = s.in_exclusion_optimize_flag[i] ? 1 : 0;
        fp.iep_pos[i] = s.in_exclusion_optimize_ans_pos[i];
    }
    gm_plan_desc& d = fp.desc;
    d.size = n;
    d.total_prefix_num = s.get_total_prefix_num();
    d.basic_prefix_num = s.get_basic_prefix_num();
    d.total_restrict_num = s.get_total_restrict_num();
    d.in_exclusion_optimize_num = s.get_in_exclusion_optimize_num();
    d.is_vertex_induced = s.is_vertex_induced ? 1 : 0;
    d.mode = mode;
    d.adj_mat = fp.adj.data();
    d.father_prefix_id = fp.father.data();
    d.last = fp.last.data();
    d.next = fp.next.data();
    d.break_size = fp.brk.data();
    d.loop_set_prefix_id = fp.loopset.data();
    d.restrict_last = fp.rlast.data();
    d.restrict_next = fp.rnext.data();
    d.restrict_index = fp.rindex.data();
    d.only_need_size = fp.only.data();
    d.iep_vertex_num = (int32_t)I;
    d.iep_vertex_id = fp.iep_id.data();
    d.iep_vertex_flag = fp.iep_vflag.data();
    d.iep_vertex_coef = fp.iep_vcoef.data();
    d.iep_array_size = (int32_t)A;
    d.iep_coef = fp.iep_coef.data();
    d.iep_flag = fp.iep_flag.data();
    d.iep_ans_pos = fp.iep_pos.data();
}

Graph::~Graph() {
    invalidate_device();
    delete[] edge;
    delete[] vertex;
    delete[] edge_from;
}

gm_graph* Graph::device_graph() {
    if (!dev_) check(gm_graph_create(vertex, edge, v_cnt, e_cnt, VertexSet::max_intersection_size, &dev_), "gm_graph_create");
    return dev_;
}

void Graph::invalidate_device() {
    if (dev_) gm_graph_destroy(dev_);
    dev_ = nullptr;
}

void Graph::build_reverse_edges() {
    delete[] edge_from;
    edge_from = new v_index_t[e_cnt > 0 ? e_cnt : 1];
    for (v_index_t u = 0; u < v_cnt; ++u)
        for (e_index_t e = vertex[u]; e < vertex[u + 1]; ++e) edge_from[e] = u;
}

long long Graph::pattern_matching(const Schedule_IEP& schedule, bool /*clique*/) {
    return (long long)run_device(this, schedule, GM_MODE_PATTERN, nullptr);
}

long long Graph::triangle_counting() { return count_triangles(*this); }

void Graph::motif_counting(int pattern_size) {
    double total_counting_time = 0;
    TimeInterval allTime, tmpTime;
    MotifGenerator mg(pattern_size);
    std::vector<Pattern> motifs = mg.generate();
    for (size_t i = 0; i < motifs.size(); ++i) {
        Pattern p = motifs[i];
        printf("pattern = \n");
        p.print();
        printf("max intersection size %d\n", VertexSet::max_intersection_size);
        bool is_pattern_valid;
        Schedule_IEP schedule_iep(p, is_pattern_valid, 1, 1, true, v_cnt, e_cnt, tri_cnt);
        if (!is_pattern_valid) {
            printf("pattern is invalid!\n");
            continue;
        }
        tmpTime.check();
        long long ans = pattern_matching(schedule_iep);
        total_counting_time += tmpTime.print("Pattern counting time: ");
        printf("ans: %lld\n", ans);
    }
    printf("Counting time cost: %.6lf s\n", total_counting_time);
    allTime.print("Total time cost");
}

// Keep N(v) ∩ [0, v) in place (src/graph.cpp:710-735).
void reduce_edges_for_clique(Graph& g) {
    printf("Trying to reduce edge. the pattern is a clique.\n");
    printf("start erasing edge\n");
    fflush(stdout);
    e_index_t kept = 0;
    for (v_index_t v = 0; v < g.v_cnt; ++v) {
        e_index_t begin = kept;
        for (e_index_t e = g.vertex[v]; e < g.vertex[v + 1] && e < g.e_cnt; ++e)
            if (g.edge[e] < v) g.edge[kept++] = g.edge[e];
        g.vertex[v] = begin;
    }
    g.vertex[g.v_cnt] = kept;
    g.e_cnt = kept;
    printf("newe: %ld\n", (long)g.e_cnt);
    printf("Finish reduce.\n");
    g.invalidate_device();
}

void degree_orientation_init(Graph* original_g, Graph*& g) {
    const v_index_t n = original_g->v_cnt;
    g = new Graph();
    g->v_cnt = n;
    g->tri_cnt = original_g->tri_cnt;
    g->vertex = new e_index_t[(size_t)n + 1];
    std::vector<e_index_t> deg(n);
    for (v_index_t u = 0; u < n; ++u) deg[u] = original_g->vertex[u + 1] - original_g->vertex[u];
    auto forward = [&](v_index_t u, v_index_t v) { return deg[u] < deg[v] || (deg[u] == deg[v] && u < v); };
    e_index_t total = 0;
    for (v_index_t u = 0; u < n; ++u) {
        g->vertex[u] = total;
        for (e_index_t i = original_g->vertex[u]; i < original_g->vertex[u + 1]; ++i)
            if (forward(u, original_g->edge[i])) ++total;
    }
    g->vertex[n] = total;
    g->e_cnt = total;
    g->edge = new v_index_t[total > 0 ? (size_t)total : 1];
    e_index_t max_out = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(max : max_out)
    for (v_index_t u = 0; u < n; ++u) {
        e_index_t p = g->vertex[u];
        for (e_index_t i = original_g->vertex[u]; i < original_g->vertex[u + 1]; ++i)
            if (forward(u, original_g->edge[i])) g->edge[p++] = original_g->edge[i];
        max_out = std::max(max_out, p - g->vertex[u]);
    }
    printf("Maximum degree after orientation: %ld\n", (long)max_out);
}

// degree_orientation_init with the vertices renumbered by their (degree, id)
// rank, the order the text loader's oriented_type 2 sorts by
// (src/dataloader.cpp:279-293): vertex r of the result is the r-th vertex in
// that order, and each edge is kept once, from its lower rank to its higher.
// So every row holds only larger ids (N(r) is a subset of (r, n)), the order
// in which a row lists its neighbours is also their rank order, and the local
// graph of any row is upper-triangular in list positions -- what the bitmap
// clique engine exploits.  Counts are those of the original graph: the
// renumbering is an isomorphism.
void degree_rank_orientation(Graph* original_g, Graph*& g) {
    const v_index_t n = original_g->v_cnt;
    std::vector<e_index_t> deg(n);
    e_index_t max_deg = 0;
    for (v_index_t u = 0; u < n; ++u) {
        deg[u] = original_g->vertex[u + 1] - original_g->vertex[u];
        max_deg = std::max(max_deg, deg[u]);
    }
    // Counting sort by degree, stable in id: by_rank[r] = old id of rank r.
    std::vector<e_index_t> start((size_t)max_deg + 2, 0);
    for (v_index_t u = 0; u < n; ++u) ++start[(size_t)deg[u] + 1];
    for (size_t d = 1; d < start.size(); ++d) start[d] += start[d - 1];
    std::vector<v_index_t> by_rank(n), rank(n);
    for (v_index_t u = 0; u < n; ++u) {
        const v_index_t r = (v_index_t)start[(size_t)deg[u]]++;
        by_rank[r] = u;
        rank[u] = r;
    }
    g = new Graph();
    g->v_cnt = n;
    g->tri_cnt = original_g->tri_cnt;
    g->vertex = new e_index_t[(size_t)n + 1];
    e_index_t total = 0;
    for (v_index_t r = 0; r < n; ++r) {
        g->vertex[r] = total;
        const v_index_t u = by_rank[r];
        for (e_index_t i = original_g->vertex[u]; i < original_g->vertex[u + 1]; ++i)
            if (rank[original_g->edge[i]] > r) ++total;
    }
    g->vertex[n] = total;
    g->e_cnt = total;
    g->edge = new v_index_t[total > 0 ? (size_t)total : 1];
    e_index_t max_out = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(max : max_out)
    for (v_index_t r = 0; r < n; ++r) {
        const v_index_t u = by_rank[r];
        e_index_t p = g->vertex[r];
        for (e_index_t i = original_g->vertex[u]; i < original_g->vertex[u + 1]; ++i) {
            const v_index_t w = rank[original_g->edge[i]];
            if (w > r) g->edge[p++] = w;
        }
        std::sort(g->edge + g->vertex[r], g->edge + p);
        max_out = std::max(max_out, p - g->vertex[r]);
    }
    printf("Maximum degree after orientation: %ld\n", (long)max_out);
}

// The vertices renumbered in (degree, id) order -- ascending for
// oriented_type 2, descending for oriented_type 1 of the text loader
// (src/dataloader.cpp:279-293) -- with every edge kept in both directions.
// Pattern counts are those of the original graph (an isomorphism); what
// changes is which endpoint the schedule's restrictions (emb[j] < emb[i])
// let through first, and so the work of a count.
void degree_relabel(Graph* original_g, Graph*& g, bool ascending) {
    const v_index_t n = original_g->v_cnt;
    std::vector<v_index_t> order(n), rank(n);
    for (v_index_t u = 0; u < n; ++u) order[u] = u;
    auto deg = [&](v_index_t u) { return original_g->vertex[u + 1] - original_g->vertex[u]; };
    std::stable_sort(order.begin(), order.end(), [&](v_index_t a, v_index_t b) {
        return ascending ? deg(a) < deg(b) : deg(a) > deg(b);
    });
    for (v_index_t r = 0; r < n; ++r) rank[order[r]] = r;
    g = new Graph();
    g->v_cnt = n;
    g->e_cnt = original_g->e_cnt;
    g->tri_cnt = original_g->tri_cnt;
    g->vertex = new e_index_t[(size_t)n + 1];
    g->edge = new v_index_t[g->e_cnt > 0 ? (size_t)g->e_cnt : 1];
    g->vertex[0] = 0;
    for (v_index_t r = 0; r < n; ++r) g->vertex[r + 1] = g->vertex[r] + deg(order[r]);
#pragma omp parallel for schedule(dynamic, 1024)
    for (v_index_t r = 0; r < n; ++r) {
        const v_index_t u = order[r];
        e_index_t p = g->vertex[r];
        for (e_index_t i = original_g->vertex[u]; i < original_g->vertex[u + 1]; ++i) g->edge[p++] = rank[original_g->edge[i]];
        std::sort(g->edge + g->vertex[r], g->edge + p);
    }
}

unsigned long long gpu_pattern_matching(Graph* g, const Schedule_IEP& s, GpuRunInfo* info) {
    return run_device(g, s, GM_MODE_PATTERN, info);
}

unsigned long long gpu_clique_matching(Graph* g, const Schedule_IEP& s, GpuRunInfo* info) {
    return run_device(g, s, GM_MODE_CLIQUE, info);
}
