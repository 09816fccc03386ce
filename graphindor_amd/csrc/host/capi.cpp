// Plain-C entry points of the host library for scripting (tests, bench.py):
// graph generation and .g I/O, the schedule compiler, plan flattening and
// the device drivers.  Errors return negative values; gs_last_error() holds
// the message.
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>

#include "autotune.h"
#include "costmodel.h"
#include "dataloader.h"
#include "generator.h"
#include "gpu_plan.h"
#include "graph.h"
#include "graphindor_amd.h"
#include "motif_generator.h"
#include "schedule_IEP.h"

#define GS_API extern "C" __attribute__((visibility("default")))

namespace {
thread_local std::string g_err;

struct GsSchedule {
    Schedule_IEP* s;
    FlatPlan flat;
    int flat_mode = -1;
};

template <class F>
int guarded(F f) {
    try {
        return f();
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    } catch (...) {
        g_err = "unknown exception";
        return -1;
    }
}
}  // namespace

GS_API const char* gs_last_error() { return g_err.c_str(); }

// ---- graphs ---------------------------------------------------------------

GS_API void* gs_graph_load(const char* path) {
    DataLoader d;
    Graph* g = nullptr;
    if (!d.fast_load(g, path)) {
        g_err = std::string("cannot load ") + path;
        return nullptr;
    }
    return g;
}

GS_API int gs_graph_save(void* g, const char* path, int32_t max_isect) {
    return DataLoader::dump(*(Graph*)g, path, (uint32_t)max_isect) ? 0 : -1;
}

GS_API void* gs_graph_from_csr(int32_t n, int64_t nnz, const int64_t* vertex, const int32_t* edge, int64_t tri_cnt) {
    Graph* g = new Graph();
    g->v_cnt = n;
    g->e_cnt = nnz;
    g->tri_cnt = tri_cnt;
    g->vertex = new e_index_t[(size_t)n + 1];
    g->edge = new v_index_t[nnz > 0 ? (size_t)nnz : 1];
    std::memcpy(g->vertex, vertex, sizeof(int64_t) * ((size_t)n + 1));
    if (nnz > 0) std::memcpy(g->edge, edge, sizeof(int32_t) * (size_t)nnz);
    return g;
}

GS_API void* gs_graph_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, int permute) {
    int32_t mi = 0;
    Graph* g = generate_rmat(scale, edge_factor, a, b, c, seed, permute != 0, &mi);
    return g;
}

GS_API void* gs_graph_er(int32_t n, double p, uint64_t seed) {
    int32_t mi = 0;
    return generate_erdos_renyi(n, p, seed, &mi);
}

GS_API void* gs_graph_batch(int count, uint64_t seed_base, double log2_min, double log2_max, double deg_min,
                            double deg_max, int64_t* vertex_off, int64_t* edge_off) {
    return generate_batch(count, seed_base, log2_min, log2_max, deg_min, deg_max, vertex_off, edge_off);
}

GS_API void gs_graph_free(void* g) { delete (Graph*)g; }

// out: v_cnt, e_cnt, tri_cnt, second-largest degree
GS_API void gs_graph_info(void* gp, int64_t* out) {
    Graph* g = (Graph*)gp;
    out[0] = g->v_cnt;
    out[1] = g->e_cnt;
    out[2] = g->tri_cnt;
    out[3] = second_largest_degree(*g);
}

GS_API void gs_graph_csr(void* gp, const int64_t** vertex, const int32_t** edge) {
    Graph* g = (Graph*)gp;
    *vertex = g->vertex;
    *edge = g->edge;
}

GS_API void gs_graph_set_tri_cnt(void* gp, int64_t t) { ((Graph*)gp)->tri_cnt = t; }

GS_API int64_t gs_graph_count_triangles(void* gp) { return count_triangles(*(Graph*)gp); }

GS_API void gs_graph_reduce_for_clique(void* gp) { reduce_edges_for_clique(*(Graph*)gp); }

GS_API void* gs_graph_orient_degree(void* gp) {
    Graph* out = nullptr;
    degree_orientation_init((Graph*)gp, out);
    fflush(stdout);
    return out;
}

GS_API void* gs_graph_orient_degree_rank(void* gp) {
    Graph* out = nullptr;
    degree_rank_orientation((Graph*)gp, out);
    fflush(stdout);
    return out;
}

GS_API void* gs_graph_relabel_degree(void* gp, int ascending) {
    Graph* out = nullptr;
    degree_relabel((Graph*)gp, out, ascending != 0);
    return out;
}

GS_API void gs_set_max_intersection_size(int32_t m) { VertexSet::max_intersection_size = m; }

// ---- schedules -------------------------------------------------------------

GS_API void gs_set_restrict_validation(int mode) {
    Schedule_IEP::set_restrict_validation(mode == 1 ? RestrictValidation::Reference : RestrictValidation::Exact);
}

GS_API void* gs_schedule_new(int n, const char* adj, int perf, int restricts, int iep, int32_t v_cnt, int64_t e_cnt,
                             int64_t tri_cnt, int vertex_induced) {
    void* out = nullptr;
    guarded([&] {
        Pattern p(n, adj);
        bool ok = true;
        Schedule_IEP* s = new Schedule_IEP(p, ok, perf, restricts, iep != 0, v_cnt, e_cnt, tri_cnt, vertex_induced != 0);
        out = new GsSchedule{s, FlatPlan(), -1};
        return 0;
    });
    return out;
}

// The two-argument form: order as given, no restrictions, no IEP.
GS_API void* gs_schedule_new_plain(int n, const char* adj) {
    void* out = nullptr;
    guarded([&] {
        Pattern p(n, adj);
        out = new GsSchedule{new Schedule_IEP(p.get_adj_mat_ptr(), n), FlatPlan(), -1};
        return 0;
    });
    return out;
}

// Candidates of the analytic selector.  Record per candidate in `out`:
// use_iep, npairs, order[n], pairs[2 * npairs]; costs[i] = analytic cost.
// Returns the number of candidates, or -1 (error) / -2 (buffer too small).
GS_API int gs_schedule_candidates(int n, const char* adj, int iep, int32_t v_cnt, int64_t e_cnt, int64_t tri_cnt,
                                  int32_t* out, int cap, double* costs, int cost_cap) {
    int rc = -1;
    guarded([&] {
        Pattern p(n, adj);
        std::vector<ScheduleCandidate> cs = Schedule_IEP::enumerate_candidates(p, iep != 0, v_cnt, e_cnt, tri_cnt);
        int pos = 0;
        if ((int)cs.size() > cost_cap) {
            rc = -2;
            return 0;
        }
        for (size_t i = 0; i < cs.size(); ++i) {
            const ScheduleCandidate& c = cs[i];
            int need = 2 + n + 2 * (int)c.pairs.size();
            if (pos + need > cap) {
                rc = -2;
                return 0;
            }
            out[pos++] = c.use_iep ? 1 : 0;
            out[pos++] = (int32_t)c.pairs.size();
            for (int v : c.order) out[pos++] = v;
            for (const auto& pr : c.pairs) {
                out[pos++] = pr.first;
                out[pos++] = pr.second;
            }
            costs[i] = c.analytic_cost;
        }
        rc = (int)cs.size();
        return 0;
    });
    return rc;
}

GS_API void* gs_schedule_from_candidate(int n, const char* adj, const int32_t* order, int npairs, const int32_t* pairs,
                                        int iep) {
    void* out = nullptr;
    guarded([&] {
        Pattern p(n, adj);
        ScheduleCandidate c;
        c.order.assign(order, order + n);
        for (int i = 0; i < npairs; ++i) c.pairs.push_back(std::make_pair(pairs[2 * i], pairs[2 * i + 1]));
        c.use_iep = iep != 0;
        out = new GsSchedule{new Schedule_IEP(p, c), FlatPlan(), -1};
        return 0;
    });
    return out;
}

GS_API void gs_schedule_free(void* h) {
    GsSchedule* g = (GsSchedule*)h;
    if (!g) return;
    delete g->s;
    delete g;
}

GS_API void gs_schedule_set_redundancy(void* h, int64_t r) { ((GsSchedule*)h)->s->set_in_exclusion_optimize_redundancy(r); }
GS_API int64_t gs_schedule_redundancy(void* h) { return ((GsSchedule*)h)->s->get_in_exclusion_optimize_redundancy(); }
GS_API int gs_schedule_multiplicity(void* h) { return ((GsSchedule*)h)->s->get_multiplicity(); }
GS_API void gs_schedule_print(void* h) { ((GsSchedule*)h)->s->print_schedule(); fflush(stdout); }

// Count on K_m with the CPU engine semantics (schedule-compiler helper).
GS_API int64_t gs_schedule_count_complete(void* h, int m) { return ((GsSchedule*)h)->s->count_on_complete_graph(m); }

GS_API int gs_restrictions_valid(int n, const char* adj, int npairs, const int32_t* pairs) {
    Pattern p(n, adj);
    std::vector<std::pair<int, int> > v;
    for (int i = 0; i < npairs; ++i) v.push_back(std::make_pair(pairs[2 * i], pairs[2 * i + 1]));
    return Schedule_IEP::restrictions_exactly_valid(p.get_adj_mat_ptr(), n, v) ? 1 : 0;
}

GS_API int gs_iep_head_multiplicity(int n, const char* adj, int npairs, const int32_t* pairs, int k) {
    Pattern p(n, adj);
    std::vector<std::pair<int, int> > v;
    for (int i = 0; i < npairs; ++i) v.push_back(std::make_pair(pairs[2 * i], pairs[2 * i + 1]));
    return Schedule_IEP::iep_head_multiplicity(p.get_adj_mat_ptr(), n, v, k);
}

// Same flat layout as the reference shim's dump (oracle/ref_shim.cpp).
GS_API int gs_schedule_dump(void* h, int32_t* out, int cap) {
    const Schedule_IEP& s = *((GsSchedule*)h)->s;
    std::vector<int32_t> v;
    int n = s.get_size();
    int P = n * (n - 1) / 2 + 1;
    int I = (int)s.in_exclusion_optimize_vertex_id.size();
    int A = (int)s.in_exclusion_optimize_coef.size();
    int head[] = {n, s.get_total_prefix_num(), s.get_basic_prefix_num(), s.get_total_restrict_num(),
                  s.get_in_exclusion_optimize_num(), s.is_vertex_induced ? 1 : 0,
                  (int)s.get_in_exclusion_optimize_redundancy(), I, A, (int)s.restrict_pair.size()};
    v.insert(v.end(), head, head + 10);
    v.insert(v.end(), s.get_adj_mat_ptr(), s.get_adj_mat_ptr() + n * n);
    v.insert(v.end(), s.get_father_prefix_id_ptr(), s.get_father_prefix_id_ptr() + P);
    v.insert(v.end(), s.get_last_ptr(), s.get_last_ptr() + n);
    v.insert(v.end(), s.get_next_ptr(), s.get_next_ptr() + P);
    v.insert(v.end(), s.get_break_size_ptr(), s.get_break_size_ptr() + P);
    v.insert(v.end(), s.get_loop_set_prefix_id_ptr(), s.get_loop_set_prefix_id_ptr() + n);
    v.insert(v.end(), s.get_restrict_last_ptr(), s.get_restrict_last_ptr() + n);
    v.insert(v.end(), s.get_restrict_next_ptr(), s.get_restrict_next_ptr() + P);
    v.insert(v.end(), s.get_restrict_index_ptr(), s.get_restrict_index_ptr() + P);
    for (int i = 0; i < P; ++i) v.push_back(s.get_prefix_only_need_size(i) ? 1 : 0);
    for (int i = 0; i < I; ++i) v.push_back(s.in_exclusion_optimize_vertex_id[i]);
    for (int i = 0; i < I; ++i) v.push_back(s.in_exclusion_optimize_vertex_flag[i] ? 1 : 0);
    for (int i = 0; i < I; ++i) v.push_back(s.in_exclusion_optimize_vertex_coef[i]);
    for (int i = 0; i < A; ++i) v.push_back(s.in_exclusion_optimize_coef[i]);
    for (int i = 0; i < A; ++i) v.push_back(s.in_exclusion_optimize_flag[i] ? 1 : 0);
    for (int i = 0; i < A; ++i) v.push_back(s.in_exclusion_optimize_ans_pos[i]);
    for (const auto& p : s.restrict_pair) {
        v.push_back(p.first);
        v.push_back(p.second);
    }
    if ((int)v.size() > cap) return -(int)v.size();
    std::memcpy(out, v.data(), v.size() * sizeof(int32_t));
    return (int)v.size();
}

// Flattened plan for `mode`; the pointers stay valid while the schedule lives.
GS_API const gm_plan_desc* gs_schedule_plan_desc(void* h, int mode) {
    GsSchedule* g = (GsSchedule*)h;
    if (g->flat_mode != mode) {
        flatten_schedule(*g->s, mode, g->flat);
        g->flat_mode = mode;
    }
    return &g->flat.desc;
}

// ---- device drivers ----------------------------------------------------------

// Count on the GPU; returns 0 and fills count (after /redundancy), raw sum
// and kernel time, or -1.
GS_API int gs_gpu_count(void* gp, void* h, int mode, uint64_t* count, uint64_t* raw, double* kernel_ms) {
    return guarded([&] {
        GpuRunInfo info;
        Graph* g = (Graph*)gp;
        const Schedule_IEP& s = *((GsSchedule*)h)->s;
        *count = (mode == GM_MODE_CLIQUE) ? gpu_clique_matching(g, s, &info) : gpu_pattern_matching(g, s, &info);
        if (raw) *raw = info.raw_sum;
        if (kernel_ms) *kernel_ms = info.kernel_ms;
        return 0;
    });
}

// ---- motifs ------------------------------------------------------------------

GS_API int gs_motifs(int k, char* out, int cap) {
    MotifGenerator mg(k);
    std::vector<Pattern> v = mg.generate();
    int need = (int)v.size() * k * k;
    if (need > cap) return -need;
    int pos = 0;
    for (const auto& p : v)
        for (int i = 0; i < k * k; ++i) out[pos++] = p.get_adj_mat_ptr()[i] ? '1' : '0';
    return (int)v.size();
}

// ---- autotuner ---------------------------------------------------------------

// Fills out[0..5] = waves_per_cu, chunk_edges, lane_tail, sample_ms*1000, from_cache, tried.
GS_API int gs_autotune_batch(const gm_graph* g, const int64_t* edge_offsets, int32_t ng, const gm_plan* const* plans,
                             int32_t np, const char* key, const char* cache_path, int32_t root_stride, int64_t* out) {
    return guarded([&] {
        TuneChoice c = autotune_batch(g, edge_offsets, ng, plans, np, key, cache_path, root_stride);
        out[0] = c.waves_per_cu;
        out[1] = c.chunk_edges;
        out[2] = c.lane_tail;
        out[3] = (int64_t)(c.sample_ms * 1000.0);
        out[4] = c.from_cache;
        out[5] = c.tried;
        return 0;
    });
}

GS_API int gs_degree_class(const int64_t* row_ptr, int32_t n, char* out, int cap) {
    std::string s = degree_class(row_ptr, n);
    if ((int)s.size() + 1 > cap) return -1;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return 0;
}

// ---- learned schedule cost model -------------------------------------------

// Features of every candidate (the order of gs_schedule_candidates with
// v_cnt = n, e_cnt = row_ptr[n]); rows of kFeatures floats.  Returns the
// number of candidates, -1 on error, -2 if `cap_rows` is too small.
GS_API int gs_encode_candidates(int n, const char* adj, int iep, const int64_t* row_ptr, int32_t nv, int64_t tri_cnt,
                                float* out, int cap_rows) {
    int rc = -1;
    guarded([&] {
        Pattern p(n, adj);
        GraphStats st = graph_stats(row_ptr, nv, tri_cnt);
        std::vector<ScheduleCandidate> cs = Schedule_IEP::enumerate_candidates(p, iep != 0, nv, st.nnz, tri_cnt);
        if ((int)cs.size() > cap_rows) {
            rc = -2;
            return 0;
        }
        std::vector<float> x = encode_candidates(p, cs, st);
        std::memcpy(out, x.data(), sizeof(float) * x.size());
        rc = (int)cs.size();
        return 0;
    });
    return rc;
}

GS_API int gs_feature_count() { return kFeatures; }

GS_API void* gs_costmodel_new(int hidden, uint64_t seed) {
    void* out = nullptr;
    guarded([&] {
        out = new CostModel(hidden, seed);
        return 0;
    });
    return out;
}

GS_API void* gs_costmodel_load(const char* path) {
    void* out = nullptr;
    guarded([&] {
        std::unique_ptr<CostModel> m = CostModel::load(path);
        if (!m) throw std::runtime_error(std::string("no usable cost model at ") + path);
        out = m.release();
        return 0;
    });
    return out;
}

GS_API void gs_costmodel_free(void* h) { delete (CostModel*)h; }

GS_API int gs_costmodel_save(void* h, const char* path) {
    return guarded([&] {
        ((CostModel*)h)->save(path);
        return 0;
    });
}

GS_API int gs_costmodel_train(void* h, const float* x, const float* t, const int32_t* goff, int32_t groups,
                              int32_t steps, float lr, float wd, int fit_norm, float* loss) {
    return guarded([&] {
        CostModel* m = (CostModel*)h;
        const int64_t rows = goff[groups];
        std::vector<float> xv(x, x + rows * kFeatures), tv(t, t + rows);
        std::vector<int32_t> go(goff, goff + groups + 1);
        if (fit_norm) m->fit_normalisation(xv, rows);
        *loss = m->train(xv, tv, go, steps, lr, wd);
        return 0;
    });
}

GS_API void gs_costmodel_set_margin(void* h, float m) { ((CostModel*)h)->set_margin(m); }
GS_API float gs_costmodel_margin(void* h) { return ((CostModel*)h)->margin(); }

GS_API int gs_costmodel_predict(void* h, const float* x, int32_t rows, float* y) {
    return guarded([&] {
        std::vector<float> xv(x, x + (int64_t)rows * kFeatures);
        std::vector<float> out = ((CostModel*)h)->predict(xv, rows);
        std::memcpy(y, out.data(), sizeof(float) * out.size());
        return 0;
    });
}

// The model's top_k proposals and the analytic pick, timed on a root sample of
// graph `gp` (sample_roots roots, time_limit_ms each); the fastest is kept.
// info = {candidates, chosen, analytic, learned}, *select_ms = time spent.
GS_API void* gs_schedule_select_measured(int n, const char* adj, int iep, void* gp, void* model, int top_k,
                                         int sample_roots, int64_t root_stride, int time_limit_ms, int lane_tail,
                                         int hub_degree, int waves_per_cu, int32_t* info, double* select_ms) {
    void* out = nullptr;
    guarded([&] {
        Pattern p(n, adj);
        MeasureOptions opt;
        opt.top_k = top_k;
        opt.sample_roots = sample_roots;
        opt.root_stride = root_stride;
        opt.time_limit_ms = time_limit_ms;
        opt.tuning.lane_tail = lane_tail;
        opt.tuning.hub_degree = hub_degree;
        opt.tuning.waves_per_cu = waves_per_cu;
        ScheduleChoice ch;
        Schedule_IEP* s = select_schedule_measured(p, *(Graph*)gp, (CostModel*)model, iep != 0, opt, &ch);
        out = new GsSchedule{s, FlatPlan(), -1};
        if (info) {
            info[0] = ch.candidates;
            info[1] = ch.chosen;
            info[2] = ch.analytic;
            info[3] = ch.learned ? 1 : 0;
        }
        if (select_ms) *select_ms = ch.select_ms;
        return 0;
    });
    return out;
}

// Schedule chosen by `model` (null: the analytic pick) among the candidates;
// info = {candidates, chosen, analytic pick, learned}, *select_ms = time spent.
GS_API void* gs_schedule_select(int n, const char* adj, int iep, const int64_t* row_ptr, int32_t nv, int64_t tri_cnt,
                                void* model, int32_t* info, double* select_ms) {
    void* out = nullptr;
    guarded([&] {
        Pattern p(n, adj);
        GraphStats st = graph_stats(row_ptr, nv, tri_cnt);
        ScheduleChoice ch;
        Schedule_IEP* s = select_schedule(p, st, (CostModel*)model, iep != 0, &ch);
        out = new GsSchedule{s, FlatPlan(), -1};
        if (info) {
            info[0] = ch.candidates;
            info[1] = ch.chosen;
            info[2] = ch.analytic;
            info[3] = ch.learned ? 1 : 0;
        }
        if (select_ms) *select_ms = ch.select_ms;
        return 0;
    });
    return out;
}
