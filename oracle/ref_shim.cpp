// C-callable shim over the reference GraphSet CPU API (include/graph.h,
// include/schedule_IEP.h, include/pattern.h, include/dataloader.h).
//
// Test infrastructure only: it is compiled together with the reference's own
// sources (src/*.cpp of the reference tree) by oracle/build_ref.sh into
// oracle/_ref/libgsref.so, and is used in this container to (1) dump the exact
// schedule the reference's Schedule_IEP constructor builds and (2) count with
// the reference's Graph::pattern_matching.  Nothing in the product links it.
#include "graph.h"
#include "dataloader.h"
#include "pattern.h"
#include "schedule_IEP.h"
#include "vertex_set.h"
#include "motif_generator.h"

#include <cstdint>
#include <cstring>
#include <vector>

namespace {

Graph* make_graph(int32_t v_cnt, int64_t e_cnt, const int64_t* vertex, const int32_t* edge,
                  int64_t tri_cnt) {
    Graph* g = new Graph();
    g->v_cnt = v_cnt;
    g->e_cnt = e_cnt;
    g->tri_cnt = tri_cnt;
    g->vertex = new e_index_t[v_cnt + 1];
    g->edge = new v_index_t[e_cnt > 0 ? e_cnt : 1];
    std::memcpy(g->vertex, vertex, sizeof(int64_t) * (v_cnt + 1));
    if (e_cnt > 0) std::memcpy(g->edge, edge, sizeof(int32_t) * e_cnt);
    return g;
}

// Flat dump layout (all int32):
//  [0] size  [1] total_prefix_num  [2] basic_prefix_num  [3] total_restrict_num
//  [4] in_exclusion_optimize_num  [5] is_vertex_induced  [6] redundancy (low 31 bits)
//  [7] iep_vertex_num  [8] iep_array_size  [9] n_restrict_pairs
//  then: adj_mat[n*n], father_prefix_id[P], last[n], next[P], break_size[P],
//        loop_set_prefix_id[n], restrict_last[n], restrict_next[P], restrict_index[P],
//        only_need_size[P], iep_vertex_id[I], iep_vertex_flag[I], iep_vertex_coef[I],
//        iep_coef[A], iep_flag[A], iep_ans_pos[A], restrict_pair[2*R]
// where P = n(n-1)/2+1.
int dump(const Schedule_IEP& s, int32_t* out, int cap) {
    std::vector<int32_t> v;
    int n = s.get_size();
    int P = n * (n - 1) / 2 + 1;
    int I = (int)s.in_exclusion_optimize_vertex_id.size();
    int A = (int)s.in_exclusion_optimize_coef.size();
    v.push_back(n);
    v.push_back(s.get_total_prefix_num());
    v.push_back(s.get_basic_prefix_num());
    v.push_back(s.get_total_restrict_num());
    v.push_back(s.get_in_exclusion_optimize_num());
    v.push_back(s.is_vertex_induced ? 1 : 0);
    v.push_back((int32_t)s.get_in_exclusion_optimize_redundancy());
    v.push_back(I);
    v.push_back(A);
    v.push_back((int)s.restrict_pair.size());
    for (int i = 0; i < n * n; ++i) v.push_back(s.get_adj_mat_ptr()[i]);
    for (int i = 0; i < P; ++i) v.push_back(s.get_father_prefix_id_ptr()[i]);
    for (int i = 0; i < n; ++i) v.push_back(s.get_last_ptr()[i]);
    for (int i = 0; i < P; ++i) v.push_back(s.get_next_ptr()[i]);
    for (int i = 0; i < P; ++i) v.push_back(s.get_break_size_ptr()[i]);
    for (int i = 0; i < n; ++i) v.push_back(s.get_loop_set_prefix_id_ptr()[i]);
    for (int i = 0; i < n; ++i) v.push_back(s.get_restrict_last_ptr()[i]);
    for (int i = 0; i < P; ++i) v.push_back(s.get_restrict_next_ptr()[i]);
    for (int i = 0; i < P; ++i) v.push_back(s.get_restrict_index_ptr()[i]);
    for (int i = 0; i < P; ++i) v.push_back(s.get_prefix_only_need_size(i) ? 1 : 0);
    for (int i = 0; i < I; ++i) v.push_back(s.in_exclusion_optimize_vertex_id[i]);
    for (int i = 0; i < I; ++i) v.push_back(s.in_exclusion_optimize_vertex_flag[i] ? 1 : 0);
    for (int i = 0; i < I; ++i) v.push_back(s.in_exclusion_optimize_vertex_coef[i]);
    for (int i = 0; i < A; ++i) v.push_back(s.in_exclusion_optimize_coef[i]);
    for (int i = 0; i < A; ++i) v.push_back(s.in_exclusion_optimize_flag[i] ? 1 : 0);
    for (int i = 0; i < A; ++i) v.push_back(s.in_exclusion_optimize_ans_pos[i]);
    for (auto& p : s.restrict_pair) { v.push_back(p.first); v.push_back(p.second); }
    if ((int)v.size() > cap) return -(int)v.size();
    std::memcpy(out, v.data(), v.size() * sizeof(int32_t));
    return (int)v.size();
}

}  // namespace

#define GSREF_API __attribute__((visibility("default")))

extern "C" {

// Schedule_IEP(pattern, valid, perf, restricts, iep, v_cnt, e_cnt, tri_cnt, vertex_induced)
// (reference include/schedule_IEP.h:15-18); returns dump length or <0.
GSREF_API int gsref_schedule_dump(int n, const char* adj, int perf, int restricts, int iep, int32_t v_cnt,
                        int64_t e_cnt, int64_t tri_cnt, int vertex_induced, int32_t* out, int cap) {
    try {
        Pattern p(n, adj);
        bool ok = true;
        Schedule_IEP s(p, ok, perf, restricts, iep != 0, v_cnt, e_cnt, tri_cnt, vertex_induced != 0);
        if (!ok) return -1;
        return dump(s, out, cap);
    } catch (...) {
        return -2;
    }
}

// Schedule_IEP(adj_mat, size) two-argument form (include/schedule_IEP.h:26).
GSREF_API int gsref_schedule_dump_plain(int n, const char* adj, int32_t* out, int cap) {
    try {
        Pattern p(n, adj);
        Schedule_IEP s(p.get_adj_mat_ptr(), n);
        return dump(s, out, cap);
    } catch (...) {
        return -2;
    }
}

GSREF_API void gsref_set_max_intersection_size(int32_t m) { VertexSet::max_intersection_size = m; }

// pm_test semantics (src/pm_test.cpp:14-47): Schedule_IEP(p, ok, 1, 1, iep, v, e, tri)
// then Graph::pattern_matching.  Returns the count.
GSREF_API long long gsref_pattern_matching(int32_t v_cnt, int64_t e_cnt, const int64_t* vertex,
                                 const int32_t* edge, int32_t max_isect, int n, const char* adj,
                                 int iep, int64_t tri_cnt) {
    VertexSet::max_intersection_size = max_isect;
    Graph* g = make_graph(v_cnt, e_cnt, vertex, edge, tri_cnt);
    Pattern p(n, adj);
    bool ok = true;
    Schedule_IEP s(p, ok, 1, 1, iep != 0, g->v_cnt, g->e_cnt, tri_cnt);
    VertexSet::max_intersection_size = std::max<int32_t>(VertexSet::max_intersection_size, max_isect);
    long long ans = g->pattern_matching(s);
    delete g;
    return ans;
}

// clique_test semantics (src/clique_test.cpp:14-76): K_k pattern, reduce_edges_for_clique,
// Schedule_IEP(p, ok, 1, 1, true, v, e) on the reduced graph, pattern_matching(s, true).
GSREF_API long long gsref_clique(int32_t v_cnt, int64_t e_cnt, const int64_t* vertex, const int32_t* edge,
                       int32_t max_isect, int k) {
    VertexSet::max_intersection_size = max_isect;
    Graph* g = make_graph(v_cnt, e_cnt, vertex, edge, 0);
    Pattern p(k);
    for (int i = 0; i < k; ++i)
        for (int j = i + 1; j < k; ++j) p.add_edge(i, j);
    reduce_edges_for_clique(*g);
    bool ok = true;
    Schedule_IEP s(p, ok, 1, 1, true, g->v_cnt, g->e_cnt);
    VertexSet::max_intersection_size = std::max<int32_t>(VertexSet::max_intersection_size, max_isect);
    long long ans = g->pattern_matching(s, true);
    delete g;
    return ans;
}

// MotifGenerator(k).generate() (src/motif_generator.cpp:33-44): writes the patterns'
// row-major adjacency strings (k*k chars each, no separators); returns their count.
GSREF_API int gsref_motifs(int k, char* out, int cap) {
    MotifGenerator mg(k);
    std::vector<Pattern> v = mg.generate();
    int need = (int)v.size() * k * k;
    if (need > cap) return -need;
    int pos = 0;
    for (auto& p : v)
        for (int i = 0; i < k * k; ++i) out[pos++] = p.get_adj_mat_ptr()[i] ? '1' : '0';
    return (int)v.size();
}

}  // extern "C"
