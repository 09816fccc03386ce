// Schedule compiler (reference src/schedule_IEP.cpp).  The order search, the
// FP64 analytic cost model, the restriction generator and the prefix/IEP
// table construction follow the reference step for step -- candidate order,
// strict-< tie breaks and floating-point evaluation order included -- so that
// the "plain analytic schedule" is the reference's.  Build with
// -ffp-contract=off: the cost model must not be contracted into FMAs.
#include "schedule_IEP.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace {

RestrictValidation g_validation_mode = [] {
    const char* e = std::getenv("GRAPHSET_RESTRICT_VALIDATION");
    return (e && std::strcmp(e, "reference") == 0) ? RestrictValidation::Reference
                                                    : RestrictValidation::Exact;
}();

int prefix_table_len(int n) { return n * (n - 1) / 2 + 1; }

// ---------------------------------------------------------------------------
// Enumeration on a complete graph K_m, following the CPU engine
// Graph::pattern_matching / pattern_matching_aggressive_func
// (src/graph.cpp:305-355, 429-566) and VertexSet (src/vertex_set.cpp).
// Used only by the schedule compiler itself.
// ---------------------------------------------------------------------------
class CompleteGraphCounter {
public:
    CompleteGraphCounter(const Schedule_IEP& s, int m) : s_(s), m_(m) {
        adj_.resize(m);
        for (int v = 0; v < m; ++v)
            for (int u = 0; u < m; ++u)
                if (u != v) adj_[v].push_back(u);
        int sets = s.get_total_prefix_num() + s.get_size() + 2;
        data_.assign(sets, std::vector<int>());
        view_ptr_.assign(sets, nullptr);
        view_len_.assign(sets, 0);
        ans_.assign(s.in_exclusion_optimize_vertex_id.size() + 1, 0);
    }

    long long run() {
        long long total = 0;
        for (int v0 = 0; v0 < m_; ++v0) {
            for (int p = s_.get_last(0); p != -1; p = s_.get_next(p)) build(p, adj_[v0].data(), (int)adj_[v0].size());
            emb_.assign(1, v0);
            recurse(1, total);
        }
        return total / s_.get_in_exclusion_optimize_redundancy();
    }

private:
    void build(int p, const int* in, int n) {
        int f = s_.get_father_prefix_id(p);
        if (f == -1) {
            view_ptr_[p] = in;
            view_len_[p] = n;
            return;
        }
        std::vector<int>& out = data_[p];
        out.clear();
        const int* a = view_ptr_[f];
        int na = view_len_[f];
        int i = 0, j = 0;
        while (i < na && j < n) {
            if (a[i] < in[j]) ++i;
            else if (a[i] > in[j]) ++j;
            else { out.push_back(a[i]); ++i; ++j; }
        }
        view_ptr_[p] = out.data();
        view_len_[p] = (int)out.size();
    }

    bool in_emb(int v) const { return std::find(emb_.begin(), emb_.end(), v) != emb_.end(); }

    // |set[0:limit) \ emb| (src/vertex_set.cpp:264-277).
    int subtraction_size(const int* set, int limit) const {
        int r = limit;
        for (int x : emb_)
            if (std::binary_search(set, set + limit, x)) --r;
        return r;
    }

    int restriction_bound(int depth) const {
        int b = m_;
        for (int i = s_.get_restrict_last(depth); i != -1; i = s_.get_restrict_next(i))
            b = std::min(b, emb_[s_.get_restrict_index(i)]);
        return b;
    }

    void recurse(int depth, long long& total) {
        int loop_id = s_.get_loop_set_prefix_id(depth);
        const int* loop = view_ptr_[loop_id];
        int loop_n = view_len_[loop_id];
        if (loop_n <= 0) return;
        std::vector<int> induced;
        if (s_.is_vertex_induced) {
            // remove_anti_edge_vertices (src/graph.cpp:195-232): drop candidates
            // adjacent to an embedded vertex the pattern says is not adjacent.
            // In K_m a candidate is adjacent to every vertex but itself.
            const int* am = s_.get_adj_mat_ptr();
            int n = s_.get_size();
            for (int i = 0; i < loop_n; ++i) {
                bool keep = true;
                for (int u = 0; u < depth && keep; ++u)
                    if (am[u * n + depth] == 0 && loop[i] != emb_[u]) keep = false;
                if (keep) induced.push_back(loop[i]);
            }
            loop = induced.data();
            loop_n = (int)induced.size();
        }
        int n = s_.get_size();
        if (depth == n - s_.get_in_exclusion_optimize_num()) {
            for (size_t i = 0; i < s_.in_exclusion_optimize_vertex_id.size(); ++i) {
                int p = s_.in_exclusion_optimize_vertex_id[i];
                if (s_.in_exclusion_optimize_vertex_flag[i])
                    ans_[i] = view_len_[p] - s_.in_exclusion_optimize_vertex_coef[i];
                else
                    ans_[i] = subtraction_size(view_ptr_[p], view_len_[p]);
            }
            int last_pos = -1;
            long long val = 0;
            for (size_t pos = 0; pos < s_.in_exclusion_optimize_coef.size(); ++pos) {
                if ((int)pos == last_pos + 1) val = ans_[s_.in_exclusion_optimize_ans_pos[pos]];
                else if (val != 0) val = val * ans_[s_.in_exclusion_optimize_ans_pos[pos]];
                if (s_.in_exclusion_optimize_flag[pos]) {
                    last_pos = (int)pos;
                    total += val * s_.in_exclusion_optimize_coef[pos];
                }
            }
            return;
        }
        if (depth == n - 1) {
            if (s_.get_total_restrict_num() > 0) {
                int bound = restriction_bound(depth);
                int lim = (int)(std::lower_bound(loop, loop + loop_n, bound) - loop);
                if (lim > 0) total += subtraction_size(loop, lim);
            } else {
                total += subtraction_size(loop, loop_n);
            }
            return;
        }
        int bound = restriction_bound(depth);
        std::vector<int> cand(loop, loop + loop_n);  // the loop set is rebuilt below
        for (int v : cand) {
            if (bound <= v) break;
            if (in_emb(v)) continue;
            bool prune = false;
            for (int p = s_.get_last(depth); p != -1; p = s_.get_next(p)) {
                build(p, adj_[v].data(), (int)adj_[v].size());
                if (view_len_[p] == s_.break_size[p]) {
                    prune = true;
                    break;
                }
            }
            if (prune) continue;
            emb_.push_back(v);
            recurse(depth + 1, total);
            emb_.pop_back();
        }
    }

    const Schedule_IEP& s_;
    int m_;
    std::vector<std::vector<int> > adj_;
    std::vector<std::vector<int> > data_;
    std::vector<const int*> view_ptr_;
    std::vector<int> view_len_;
    std::vector<int> emb_;
    std::vector<int> ans_;
};

}  // namespace

void Schedule_IEP::set_restrict_validation(RestrictValidation mode) { g_validation_mode = mode; }
RestrictValidation Schedule_IEP::get_restrict_validation() { return g_validation_mode; }

long long Schedule_IEP::count_on_complete_graph(int m) const {
    if (m > 8) throw std::logic_error("count_on_complete_graph: m > 8");
    return CompleteGraphCounter(*this, m).run();
}

void Schedule_IEP::allocate_tables() {
    int P = prefix_table_len(size);
    adj_mat = new int[size * size];
    father_prefix_id = new int[P];
    last = new int[size];
    next = new int[P];
    break_size = new int[P];
    loop_set_prefix_id = new int[size];
    prefix = new Prefix[P];
    restrict_last = new int[size];
    restrict_next = new int[P];
    restrict_index = new int[P];
    std::fill(father_prefix_id, father_prefix_id + P, -1);
    std::fill(next, next + P, -1);
    std::fill(break_size, break_size + P, -1);
    std::fill(restrict_next, restrict_next + P, -1);
    std::fill(restrict_index, restrict_index + P, -1);
    std::fill(last, last + size, -1);
    std::fill(loop_set_prefix_id, loop_set_prefix_id + size, -1);
    std::fill(restrict_last, restrict_last + size, -1);
    total_prefix_num = basic_prefix_num = total_restrict_num = 0;
    in_exclusion_optimize_num = 0;
    k_val = 0;
    in_exclusion_optimize_redundancy = 1;
}

// src/schedule_IEP.cpp:12-247.
Schedule_IEP::Schedule_IEP(const Pattern& pattern, bool& is_pattern_valid, int performance_modeling_type,
                           int restricts_type, bool use_in_exclusion_optimize, int v_cnt, int64_t e_cnt,
                           long long tri_cnt, bool vertex_induced) {
    is_vertex_induced = vertex_induced;
    if (vertex_induced) use_in_exclusion_optimize = false;
    if (performance_modeling_type == 1 && tri_cnt == -1)
        throw std::logic_error("Fatal: Can not use performance modeling if not have triangle number of this dataset.\n");

    is_pattern_valid = true;
    size = pattern.get_size();
    allocate_tables();
    const int* pattern_adj = pattern.get_adj_mat_ptr();
    std::memcpy(adj_mat, pattern_adj, sizeof(int) * size * size);

    std::vector<std::pair<int, int> > best_pairs;
    if (performance_modeling_type != 0) {
        std::vector<std::vector<int> > candidates = all_permutations(size);
        remove_invalid_permutation(candidates);
        if (performance_modeling_type == 1) {
            // Keep only the orders with the deepest IEP (src/schedule_IEP.cpp:82-96).
            int deepest = 0;
            for (const auto& vec : candidates) deepest = std::max(deepest, get_vec_optimize_num(vec));
            k_val = deepest;
            std::vector<std::vector<int> > kept;
            for (const auto& vec : candidates)
                if (get_vec_optimize_num(vec) == deepest) kept.push_back(vec);
            candidates.swap(kept);
        }
        std::vector<int> best_order(size);
        double min_val = 1e18;
        bool have_best = false;
        for (const auto& vec : candidates) {
            copy_adj_mat_from(vec, pattern_adj);
            setup_optimization_info(use_in_exclusion_optimize);
            std::vector<std::vector<std::pair<int, int> > > sets;
            if (restricts_type == 1) {
                restricts_generate(adj_mat, sets);
            } else {
                Schedule_IEP plain(adj_mat, size);
                std::vector<std::pair<int, int> > pairs;
                plain.GraphZero_aggressive_optimize(pairs);
                sets.push_back(pairs);
            }
            if (sets.empty()) {
                // Exact mode: an order without a valid restriction set is not
                // a candidate (an empty set would count every automorphic copy).
                if (g_validation_mode == RestrictValidation::Exact) continue;
                sets.push_back(std::vector<std::pair<int, int> >());
            }
            for (const auto& pairs : sets) {
                double val = 1e18;
                if (performance_modeling_type == 1) val = new_estimate_schedule_restrict(pairs, v_cnt, e_cnt, tri_cnt);
                if (!have_best || val < min_val) {
                    have_best = true;
                    min_val = val;
                    best_order = vec;
                    best_pairs = pairs;
                }
            }
        }
        if (!have_best && use_in_exclusion_optimize) {
            // Exact mode found no order whose restrictions survive the IEP
            // tail: choose again without the IEP.
            use_in_exclusion_optimize = false;
            for (const auto& vec : candidates) {
                copy_adj_mat_from(vec, pattern_adj);
                setup_optimization_info(false);
                std::vector<std::vector<std::pair<int, int> > > sets;
                restricts_generate(adj_mat, sets);
                for (const auto& pairs : sets) {
                    double val = 1e18;
                    if (performance_modeling_type == 1) val = new_estimate_schedule_restrict(pairs, v_cnt, e_cnt, tri_cnt);
                    if (!have_best || val < min_val) {
                        have_best = true;
                        min_val = val;
                        best_order = vec;
                        best_pairs = pairs;
                    }
                }
            }
        }
        if (!have_best) throw std::runtime_error("Schedule_IEP: no valid restriction set for any vertex order");
        copy_adj_mat_from(best_order, pattern_adj);
    } else {
        std::vector<std::vector<std::pair<int, int> > > sets;
        if (g_validation_mode == RestrictValidation::Exact) setup_optimization_info(use_in_exclusion_optimize);
        if (restricts_type == 1) {
            restricts_generate(adj_mat, sets);
            if (sets.empty() && g_validation_mode == RestrictValidation::Exact && use_in_exclusion_optimize) {
                use_in_exclusion_optimize = false;
                setup_optimization_info(false);
                restricts_generate(adj_mat, sets);
            }
        } else if (restricts_type != 0) {
            std::vector<std::pair<int, int> > pairs;
            GraphZero_aggressive_optimize(pairs);
            sets.push_back(pairs);
        }
        bool have_best = false;
        double min_val = 0;
        for (const auto& pairs : sets) {
            // The reference evaluates this before any IEP depth is set (the
            // member is still unset there); the depth is 0 here.
            double val = new_estimate_schedule_restrict(pairs, v_cnt, e_cnt, tri_cnt);
            if (!have_best || val < min_val) {
                have_best = true;
                min_val = val;
                best_pairs = pairs;
            }
        }
    }

    if (!check_connectivity()) {
        is_pattern_valid = false;
        throw std::runtime_error("pattern is not connected");
    }
    setup_optimization_info(use_in_exclusion_optimize);
    if (restricts_type != 0) add_restrict(best_pairs);
    set_in_exclusion_optimize_redundancy();
}

// src/schedule_IEP.cpp:249-291.
Schedule_IEP::Schedule_IEP(const int* _adj_mat, int _size) {
    size = _size;
    allocate_tables();
    std::memcpy(adj_mat, _adj_mat, sizeof(int) * size * size);
    is_vertex_induced = true;
    if (!check_connectivity()) throw std::runtime_error("pattern is not connected");
    build_loop_invariant();
    set_in_exclusion_optimize_redundancy();
}

Schedule_IEP::Schedule_IEP(const Pattern& pattern, const ScheduleCandidate& c) {
    is_vertex_induced = false;
    size = pattern.get_size();
    allocate_tables();
    if ((int)c.order.size() != size) throw std::invalid_argument("Schedule_IEP: candidate order has the wrong length");
    std::vector<int> seen(size, 0);
    for (int v : c.order) {
        if (v < 0 || v >= size || seen[v]++) throw std::invalid_argument("Schedule_IEP: candidate order is not a permutation");
    }
    for (const auto& pr : c.pairs)
        if (pr.first < 0 || pr.first >= size || pr.second < 0 || pr.second >= size || pr.first == pr.second)
            throw std::invalid_argument("Schedule_IEP: candidate restriction out of range");
    copy_adj_mat_from(c.order, pattern.get_adj_mat_ptr());
    if (!check_connectivity()) throw std::invalid_argument("Schedule_IEP: candidate order is not connected");
    setup_optimization_info(c.use_iep);
    if (!restrictions_exactly_valid(adj_mat, size, c.pairs) ||
        (in_exclusion_optimize_num >= 2 && iep_head_multiplicity(adj_mat, size, c.pairs, in_exclusion_optimize_num) <= 0))
        throw std::invalid_argument("Schedule_IEP: candidate restriction set does not break every automorphism once");
    add_restrict(c.pairs);
    set_in_exclusion_optimize_redundancy();
}

std::vector<ScheduleCandidate> Schedule_IEP::enumerate_candidates(const Pattern& pattern, bool use_iep, int v_cnt,
                                                                  int64_t e_cnt, long long tri_cnt) {
    if (tri_cnt < 0) throw std::logic_error("enumerate_candidates: the triangle count is required");
    const int n = pattern.get_size();
    const int* src = pattern.get_adj_mat_ptr();
    // Scratch whose tables each order rebuilds.  Its constructor wants an
    // order in which every vertex touches an earlier one: a BFS order.
    std::vector<int> bfs(1, 0), seen(n, 0);
    seen[0] = 1;
    for (size_t h = 0; h < bfs.size(); ++h)
        for (int v = 0; v < n; ++v)
            if (!seen[v] && src[bfs[h] * n + v]) {
                seen[v] = 1;
                bfs.push_back(v);
            }
    if ((int)bfs.size() != n) throw std::runtime_error("pattern is not connected");
    std::vector<int> start(n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) start[i * n + j] = src[bfs[i] * n + bfs[j]];
    Schedule_IEP s(start.data(), n);
    // Orders below are relabellings of the pattern as given.
    std::memcpy(s.adj_mat, src, sizeof(int) * n * n);
    std::vector<std::vector<int> > orders = all_permutations(n);
    s.remove_invalid_permutation(orders);
    int deepest = 0;
    for (const auto& vec : orders) deepest = std::max(deepest, s.get_vec_optimize_num(vec));
    std::vector<std::vector<int> > kept;
    for (const auto& vec : orders)
        if (s.get_vec_optimize_num(vec) == deepest) kept.push_back(vec);
    std::vector<ScheduleCandidate> out;
    for (int pass = 0; pass < 2 && out.empty(); ++pass) {
        const bool iep = pass == 0 && use_iep;
        if (pass == 1 && !use_iep) break;
        for (const auto& vec : kept) {
            s.copy_adj_mat_from(vec, src);
            s.setup_optimization_info(iep);
            std::vector<std::vector<std::pair<int, int> > > sets;
            s.restricts_generate(s.adj_mat, sets);
            for (const auto& pairs : sets) {
                ScheduleCandidate c;
                c.order = vec;
                c.pairs = pairs;
                c.use_iep = iep;
                c.analytic_cost = s.new_estimate_schedule_restrict(pairs, v_cnt, e_cnt, tri_cnt);
                out.push_back(c);
            }
        }
    }
    if (out.empty()) throw std::runtime_error("enumerate_candidates: no valid restriction set for any vertex order");
    return out;
}

Schedule_IEP::~Schedule_IEP() {
    delete[] adj_mat;
    delete[] father_prefix_id;
    delete[] last;
    delete[] next;
    delete[] break_size;
    delete[] loop_set_prefix_id;
    delete[] prefix_target;
    delete[] prefix;
    delete[] restrict_last;
    delete[] restrict_next;
    delete[] restrict_index;
}

// Analytic cost of the current order with `pairs` (src/schedule_IEP.cpp:307-382).
double Schedule_IEP::new_estimate_schedule_restrict(const std::vector<std::pair<int, int> >& pairs, int v_cnt,
                                                    e_index_t e_cnt, long long tri_cnt) {
    int max_degree = get_max_degree();
    std::vector<double> p_size(max_degree), pp_size(max_degree);
    double p0 = e_cnt * 1.0 / v_cnt / v_cnt;
    double p1 = tri_cnt * 1.0 * v_cnt / e_cnt / e_cnt;
    p_size[0] = v_cnt;
    for (int i = 1; i < max_degree; ++i) p_size[i] = p_size[i - 1] * p0;
    pp_size[0] = 1;
    for (int i = 1; i < max_degree; ++i) pp_size[i] = pp_size[i - 1] * p1;

    std::vector<std::pair<int, int> > rs = pairs;
    std::sort(rs.begin(), rs.end());
    int nr = (int)rs.size();
    // sum[i]: fraction of the n! relative orders meeting restrictions 0..i,
    // then turned into conditional fractions.
    std::vector<double> sum(nr, 0.0);
    std::vector<int> perm(size);
    for (int i = 0; i < size; ++i) perm[i] = i;
    do {
        for (int i = 0; i < nr; ++i) {
            if (perm[rs[i].first] > perm[rs[i].second]) sum[i] += 1;
            else break;
        }
    } while (std::next_permutation(perm.begin(), perm.end()));
    double total = 1;
    for (int i = 2; i <= size; ++i) total *= i;
    for (int i = 0; i < nr; ++i) sum[i] = sum[i] / total;
    for (int i = nr - 1; i > 0; --i) sum[i] /= sum[i - 1];

    std::vector<std::vector<int> > invariant_size(size);
    double val = 1;
    for (int i = size - in_exclusion_optimize_num - 1; i >= 0; --i) {
        int cnt_forward = 0;
        for (int j = 0; j < i; ++j)
            if (adj_mat[INDEX(j, i, size)]) ++cnt_forward;
        int c = cnt_forward;
        for (int j = i - 1; j >= 0; --j)
            if (adj_mat[INDEX(j, i, size)]) invariant_size[j].push_back(c--);
        for (int x : invariant_size[i])
            if (x > 1) val += p_size[1] * pp_size[x - 2] * std::log2(p_size[1]);
        for (int j = 0; j < nr; ++j)
            if (rs[j].second == i) val *= sum[j];
        if (i) val *= p_size[1] * pp_size[cnt_forward - 1];
        else val *= p_size[0];
    }
    return val;
}

// Relabel: new vertex k is original vertex vec[k] (src/schedule_IEP.cpp:384-391).
void Schedule_IEP::copy_adj_mat_from(const std::vector<int>& vec, const int* src) {
    std::vector<int> rank(size);
    for (int i = 0; i < size; ++i) rank[vec[i]] = i;
    for (int i = 0; i < size; ++i)
        for (int j = 0; j < size; ++j) adj_mat[INDEX(rank[i], rank[j], size)] = src[INDEX(i, j, size)];
}

// src/schedule_IEP.cpp:420-450.  The reference resets only the first
// n(n-1)/2 entries of the prefix-indexed tables; Reference mode keeps that.
void Schedule_IEP::setup_optimization_info(bool use_in_exclusion_optimize) {
    int P = prefix_table_len(size);
    int reset = (g_validation_mode == RestrictValidation::Reference) ? P - 1 : P;
    std::fill(father_prefix_id, father_prefix_id + reset, -1);
    std::fill(next, next + reset, -1);
    std::fill(break_size, break_size + reset, -1);
    std::fill(restrict_next, restrict_next + reset, -1);
    std::fill(last, last + size, -1);
    std::fill(restrict_last, restrict_last + size, -1);
    if (g_validation_mode == RestrictValidation::Exact) {
        // Fresh tables for every order: no flags or IEP rows carried over.
        for (int i = 0; i < P; ++i) {
            prefix[i].set_has_child(false);
            prefix[i].set_only_need_size(false);
        }
        in_exclusion_optimize_vertex_flag.clear();
        in_exclusion_optimize_vertex_coef.clear();
        in_exclusion_optimize_ans_pos.clear();
    }
    total_prefix_num = 0;
    total_restrict_num = 0;
    in_exclusion_optimize_num = 0;
    if (use_in_exclusion_optimize) {
        std::vector<int> identity(size);
        for (int i = 0; i < size; ++i) identity[i] = i;
        int k = get_vec_optimize_num(identity);
        if (k > 1) {
            in_exclusion_optimize_num = k;
            init_in_exclusion_optimize();
        }
    }
    build_loop_invariant(in_exclusion_optimize_num);
}

// Every vertex after the first must touch an earlier one (src/schedule_IEP.cpp:452-467).
bool Schedule_IEP::check_connectivity() const {
    for (int i = 1; i < size; ++i) {
        bool ok = false;
        for (int j = 0; j < i && !ok; ++j) ok = adj_mat[INDEX(i, j, size)] != 0;
        if (!ok) return false;
    }
    return true;
}

int Schedule_IEP::get_max_degree() const {
    int best = 0;
    for (int i = 0; i < size; ++i) {
        int d = 0;
        for (int j = 0; j < size; ++j) d += adj_mat[INDEX(i, j, size)];
        best = std::max(best, d);
    }
    return best;
}

// Prefix trie, break sizes and IEP tables (src/schedule_IEP.cpp:480-627).
void Schedule_IEP::build_loop_invariant(int iep_num) {
    std::vector<int> data;
    data.reserve(size);
    loop_set_prefix_id[0] = -1;
    for (int i = 1; i < size; ++i) {
        data.clear();
        for (int j = 0; j < i; ++j)
            if (adj_mat[INDEX(i, j, size)]) data.push_back(j);
        loop_set_prefix_id[i] = find_father_prefix((int)data.size(), data.data());
    }
    basic_prefix_num = total_prefix_num;

    // break_size: earlier vertices outside the prefix that are adjacent to
    // all of it; they are always members, so |set| == break_size means no
    // new vertex can follow.
    for (int p = 0; p < basic_prefix_num; ++p) {
        const int* d = prefix[p].get_data_ptr();
        int ds = prefix[p].get_size();
        int count = 0;
        for (int node = 0; node < d[ds - 1]; ++node) {
            bool member = false, all_adjacent = true;
            for (int i = 0; i < ds; ++i) {
                if (d[i] == node) { member = true; break; }
                if (!adj_mat[INDEX(node, d[i], size)]) { all_adjacent = false; break; }
            }
            if (!member && all_adjacent) ++count;
        }
        break_size[p] = count;
    }

    if (iep_num > 0) {
        in_exclusion_optimize_vertex_id.clear();
        in_exclusion_optimize_coef.clear();
        in_exclusion_optimize_flag.clear();
        int head = size - iep_num;  // vertices [head, size) form the IEP tail
        for (size_t g = 0; g < in_exclusion_optimize_group.size(); ++g) {
            const std::vector<std::vector<int> >& blocks = in_exclusion_optimize_group[g];
            long long gval = in_exclusion_optimize_val[g];
            for (size_t b = 0; b < blocks.size(); ++b) {
                data.clear();
                for (int node = 0; node < head; ++node)
                    for (int member : blocks[b])
                        if (adj_mat[INDEX(node, member + head, size)]) {
                            data.push_back(node);
                            break;
                        }
                int pid = find_father_prefix((int)data.size(), data.data());
                int pos = -1;
                for (size_t i = 0; i < in_exclusion_optimize_vertex_id.size(); ++i)
                    if (in_exclusion_optimize_vertex_id[i] == pid) { pos = (int)i; break; }
                if (pos == -1) {
                    in_exclusion_optimize_vertex_id.push_back(pid);
                    pos = (int)in_exclusion_optimize_vertex_id.size() - 1;
                    // flag: every non-tail vertex is in the prefix or adjacent to
                    // all of it, so |S \ emb| == |S| - (such outside vertices).
                    int in_set = 0, full_connect = 0;
                    for (int i = 0; i < head; ++i) {
                        int cnt = 0;
                        bool hit = false;
                        for (int x : data) {
                            if (x == i) { hit = true; break; }
                            if (adj_mat[INDEX(i, x, size)]) ++cnt;
                        }
                        if (hit) ++in_set;
                        else if (cnt == (int)data.size()) ++full_connect;
                    }
                    in_exclusion_optimize_vertex_flag.push_back(in_set + full_connect == head);
                    in_exclusion_optimize_vertex_coef.push_back(full_connect);
                }
                in_exclusion_optimize_ans_pos.push_back(pos);
                bool closes = (b + 1 == blocks.size());
                in_exclusion_optimize_coef.push_back(closes ? (int)gval : 0);
                in_exclusion_optimize_flag.push_back(closes);
            }
        }
        for (size_t i = 0; i < in_exclusion_optimize_vertex_flag.size(); ++i)
            if (in_exclusion_optimize_vertex_flag[i] && i < in_exclusion_optimize_vertex_id.size()) {
                int pid = in_exclusion_optimize_vertex_id[i];
                if (!prefix[pid].get_has_child()) prefix[pid].set_only_need_size(true);
            }
    }

    // Reverse each per-vertex prefix list so prefixes are built in creation
    // order (fathers before children).
    for (int i = 0; i < size; ++i) {
        if (last[i] == -1) continue;
        std::vector<int> chain;
        for (int id = last[i]; id != -1; id = next[id]) chain.push_back(id);
        last[i] = chain.back();
        for (size_t k = chain.size() - 1; k > 0; --k) next[chain[k]] = chain[k - 1];
        next[chain[0]] = -1;
    }
}

// Find or create the prefix `data` (sorted), creating its ancestors first
// (src/schedule_IEP.cpp:629-648).
int Schedule_IEP::find_father_prefix(int data_size, const int* data) {
    if (data_size == 0) return -1;
    int tail = data[data_size - 1];
    for (int p = last[tail]; p != -1; p = next[p])
        if (prefix[p].equal(data_size, data)) return p;
    int father = find_father_prefix(data_size - 1, data);
    if (father >= 0) prefix[father].set_has_child(true);
    int id = total_prefix_num++;
    father_prefix_id[id] = father;
    next[id] = last[tail];
    last[tail] = id;
    prefix[id].init(data_size, data);
    return id;
}

// Drop pairs implied by two others, then thread the restriction lists by the
// later vertex (src/schedule_IEP.cpp:650-684).
void Schedule_IEP::add_restrict(const std::vector<std::pair<int, int> >& restricts) {
    restrict_pair = restricts;
    for (size_t i = 0; i < restrict_pair.size();) {
        bool implied = false;
        for (size_t j = 0; j < restrict_pair.size() && !implied; ++j) {
            if (i == j || restrict_pair[j].first != restrict_pair[i].first) continue;
            for (size_t k = 0; k < restrict_pair.size(); ++k)
                if (k != i && k != j && restrict_pair[k].second == restrict_pair[i].second &&
                    restrict_pair[j].second == restrict_pair[k].first) {
                    implied = true;
                    break;
                }
        }
        if (implied) restrict_pair.erase(restrict_pair.begin() + i);
        else ++i;
    }
    int reset = prefix_table_len(size) - 1;
    std::fill(restrict_last, restrict_last + size, -1);
    std::fill(restrict_next, restrict_next + reset, -1);
    total_restrict_num = 0;
    for (const auto& p : restrict_pair) {
        restrict_index[total_restrict_num] = p.first;
        restrict_next[total_restrict_num] = restrict_last[p.second];
        restrict_last[p.second] = total_restrict_num;
        ++total_restrict_num;
    }
}

int Schedule_IEP::get_multiplicity() const { return (int)get_isomorphism_vec().size(); }

namespace {

// Removes every permutation group made of one transposition plus fixed
// points, recording the transposition as a restriction
// (src/schedule_IEP.cpp:701-728 and 797-822).
void take_lone_transpositions(std::vector<std::vector<int> >& iso, std::vector<std::vector<std::vector<int> > >& groups,
                              std::vector<std::pair<int, int> >& pairs) {
    for (size_t i = 0; i < groups.size();) {
        int twos = 0;
        std::pair<int, int> found;
        for (const std::vector<int>& cyc : groups[i]) {
            if (cyc.size() == 2) {
                ++twos;
                found = std::make_pair(cyc[0], cyc[1]);
            } else if (cyc.size() != 1) {
                twos = -1;
                break;
            }
        }
        if (twos == 1) {
            groups.erase(groups.begin() + i);
            iso.erase(iso.begin() + i);
            pairs.push_back(found);
        } else {
            ++i;
        }
    }
}

// Removes the isomorphisms whose image of the restrictions would close a
// cycle with the restriction DAG (src/schedule_IEP.cpp:734-749, 837-851).
void drop_conflicting(const Pattern& base_dag, std::vector<std::vector<int> >& iso,
                      std::vector<std::vector<std::vector<int> > >& groups, const std::vector<std::pair<int, int> >& pairs) {
    for (size_t i = 0; i < iso.size();) {
        Pattern test(base_dag);
        for (const auto& p : pairs) test.add_ordered_edge(iso[i][p.first], iso[i][p.second]);
        if (!test.is_dag()) {
            groups.erase(groups.begin() + i);
            iso.erase(iso.begin() + i);
        } else {
            ++i;
        }
    }
}

void all_pairs_dfs(const Pattern& base_dag, std::vector<std::vector<int> > iso,
                   std::vector<std::vector<std::vector<int> > > groups, const std::vector<std::pair<int, int> >& pairs,
                   std::vector<std::vector<std::pair<int, int> > >& out) {
    drop_conflicting(base_dag, iso, groups, pairs);
    if (iso.size() == 1) {
        out.push_back(pairs);
        return;
    }
    // Branch on every transposition of the first group that has one
    // (src/schedule_IEP.cpp:858-887).
    for (size_t i = 0; i < groups.size(); ++i) {
        bool branched = false;
        for (const std::vector<int>& cyc : groups[i]) {
            if (cyc.size() != 2) continue;
            branched = true;
            std::vector<std::vector<std::vector<int> > > g2 = groups;
            std::vector<std::vector<int> > iso2 = iso;
            std::vector<std::pair<int, int> > pairs2 = pairs;
            Pattern dag2 = base_dag;
            g2.erase(g2.begin() + i);
            iso2.erase(iso2.begin() + i);
            pairs2.push_back(std::make_pair(cyc[0], cyc[1]));
            dag2.add_ordered_edge(cyc[0], cyc[1]);
            all_pairs_dfs(dag2, iso2, g2, pairs2, out);
        }
        if (branched) break;
    }
}

}  // namespace

// One valid restriction set, greedily (src/schedule_IEP.cpp:691-783).
void Schedule_IEP::aggressive_optimize(std::vector<std::pair<int, int> >& ordered_pairs) const {
    std::vector<std::vector<int> > iso = get_isomorphism_vec();
    std::vector<std::vector<std::vector<int> > > groups;
    for (const auto& v : iso) groups.push_back(calc_permutation_group(v, size));
    ordered_pairs.clear();
    take_lone_transpositions(iso, groups, ordered_pairs);
    Pattern base_dag(size);
    for (const auto& p : ordered_pairs) base_dag.add_ordered_edge(p.first, p.second);
    bool changed = true;
    while (changed && iso.size() != 1) {
        drop_conflicting(base_dag, iso, groups, ordered_pairs);
        changed = false;
        for (size_t i = 0; i < groups.size(); ++i) {
            const std::vector<int>* two = nullptr;
            for (const std::vector<int>& cyc : groups[i])
                if (cyc.size() == 2) { two = &cyc; break; }
            if (two) {
                std::pair<int, int> found((*two)[0], (*two)[1]);
                groups.erase(groups.begin() + i);
                iso.erase(iso.begin() + i);
                ordered_pairs.push_back(found);
                base_dag.add_ordered_edge(found.first, found.second);
                changed = true;
                break;
            }
        }
    }
}

// Every restriction set reachable by branching on transpositions
// (src/schedule_IEP.cpp:785-888); validity is checked by the caller.
void Schedule_IEP::aggressive_optimize_get_all_pairs(std::vector<std::vector<std::pair<int, int> > >& out) {
    std::vector<std::vector<int> > iso = get_isomorphism_vec();
    std::vector<std::vector<std::vector<int> > > groups;
    for (const auto& v : iso) groups.push_back(calc_permutation_group(v, size));
    out.clear();
    std::vector<std::pair<int, int> > pairs;
    take_lone_transpositions(iso, groups, pairs);
    Pattern base_dag(size);
    for (const auto& p : pairs) base_dag.add_ordered_edge(p.first, p.second);
    all_pairs_dfs(base_dag, iso, groups, pairs, out);
}

// GraphZero's restrictions via a stabiliser chain (src/schedule_IEP.cpp:890-940).
void Schedule_IEP::GraphZero_aggressive_optimize(std::vector<std::pair<int, int> >& ordered_pairs) const {
    std::vector<std::vector<int> > aut;
    GraphZero_get_automorphisms(aut);
    std::vector<std::pair<int, int> > L;
    for (int v = 0; v < size; ++v) {
        std::vector<std::vector<int> > stab;
        for (const auto& x : aut) {
            if (x[v] == v) {
                stab.push_back(x);
                continue;
            }
            std::pair<int, int> cur(std::min(v, x[v]), std::max(v, x[v]));
            if (std::find(L.begin(), L.end(), cur) == L.end()) L.push_back(cur);
        }
        aut.swap(stab);
    }
    ordered_pairs.clear();
    for (const auto& l : L) {
        bool fresh = true;
        for (auto& o : ordered_pairs)
            if (l.second == o.second) {
                fresh = false;
                if (l.first > o.first) o.first = l.first;
                break;
            }
        if (fresh) ordered_pairs.push_back(l);
    }
}

void Schedule_IEP::GraphZero_get_automorphisms(std::vector<std::vector<int> >& aut) const {
    aut.clear();
    for (const auto& p : all_permutations(size)) {
        bool same = true;
        for (int i = 0; i < size && same; ++i)
            for (int j = 0; j < size; ++j)
                if (adj_mat[INDEX(i, j, size)] != adj_mat[INDEX(p[i], p[j], size)]) { same = false; break; }
        if (same) aut.push_back(p);
    }
}

// Permutations mapping every edge onto an edge, in lexicographic order
// (src/schedule_IEP.cpp:966-996).
std::vector<std::vector<int> > Schedule_IEP::get_isomorphism_vec() const {
    std::vector<std::vector<int> > out;
    for (const auto& v : all_permutations(size)) {
        bool keeps = true;
        for (int i = 0; i < size && keeps; ++i)
            for (int j = i + 1; j < size; ++j)
                if (adj_mat[INDEX(i, j, size)] != 0 && adj_mat[INDEX(v[i], v[j], size)] == 0) { keeps = false; break; }
        if (keeps) out.push_back(v);
    }
    return out;
}

// Cycle decomposition, each cycle starting at its smallest unvisited element
// (src/schedule_IEP.cpp:1016-1040).
std::vector<std::vector<int> > Schedule_IEP::calc_permutation_group(const std::vector<int> vec, int size) {
    std::vector<char> seen(size, 0);
    std::vector<std::vector<int> > cycles;
    for (int i = 0; i < (int)vec.size(); ++i) {
        if (seen[i]) continue;
        std::vector<int> cyc(1, i);
        seen[i] = 1;
        for (int x = vec[i]; !seen[x]; x = vec[x]) {
            seen[x] = 1;
            cyc.push_back(x);
        }
        cycles.push_back(cyc);
    }
    return cycles;
}

// IEP coefficients per set partition of the tail (src/schedule_IEP.cpp:1400-1496).
// For a block of b vertices the signed count of connected spanning
// subgraphs of K_b is tallied by enumerating edge subsets with union-find.
void Schedule_IEP::init_in_exclusion_optimize() {
    int k = in_exclusion_optimize_num;
    std::vector<int> val(2 * k, 0);
    for (int n = 1; n <= k; ++n) {
        if (n == 1) {
            val[0] = 1;
            continue;
        }
        std::vector<std::pair<int, int> > edges;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) edges.push_back(std::make_pair(i, j));
        int m = (int)edges.size();
        DisjointSetUnion dsu(n);
        for (int s = 0; s < (1 << m); ++s) {
            dsu.init();
            int bits = 0;
            for (int i = 0; i < m; ++i)
                if (s & (1 << i)) {
                    ++bits;
                    dsu.merge(edges[i].first, edges[i].second);
                }
            if (dsu.get_set_size() == 1) ++val[2 * n - 2 + (bits & 1)];
        }
    }
    in_exclusion_optimize_group.clear();
    in_exclusion_optimize_val.clear();
    std::vector<int> id(k);
    get_in_exclusion_optimize_group(0, id.data(), 0, val.data());
}

// Restricted-growth enumeration of set partitions, "new block" first
// (src/schedule_IEP.cpp:1453-1496).
void Schedule_IEP::get_in_exclusion_optimize_group(int depth, int* id, int id_cnt, int* ival) {
    int k = in_exclusion_optimize_num;
    if (depth == k) {
        std::vector<int> bsize(id_cnt, 0);
        for (int i = 0; i < k; ++i) ++bsize[id[i]];
        int even = ival[bsize[0] * 2 - 2], odd = ival[bsize[0] * 2 - 1];
        for (int i = 1; i < id_cnt; ++i) {
            int e = even, o = odd;
            even = e * ival[bsize[i] * 2 - 2] + o * ival[bsize[i] * 2 - 1];
            odd = e * ival[bsize[i] * 2 - 1] + o * ival[bsize[i] * 2 - 2];
        }
        std::vector<std::vector<int> > blocks(id_cnt);
        for (int j = 0; j < k; ++j) blocks[id[j]].push_back(j);
        in_exclusion_optimize_group.push_back(blocks);
        in_exclusion_optimize_val.push_back(even - odd);
        return;
    }
    id[depth] = id_cnt;
    get_in_exclusion_optimize_group(depth + 1, id, id_cnt + 1, ival);
    for (int i = 0; i < id_cnt; ++i) {
        id[depth] = i;
        get_in_exclusion_optimize_group(depth + 1, id, id_cnt, ival);
    }
}

void Schedule_IEP::print_schedule() const {
    printf("Schedule_IEP:\n");
    for (int i = 0; i < size; ++i) {
        for (int j = 0; j < size; ++j) printf("%d", adj_mat[INDEX(i, j, size)]);
        puts("");
    }
    printf("loop_set_prefix_ids:");
    for (int i = 0; i < size; ++i) printf(" %d", loop_set_prefix_id[i]);
    printf("\nlast:");
    for (int i = 0; i < size; ++i) printf(" %d", last[i]);
    printf("\nnext:");
    for (int i = 0; i < size * (size - 1) / 2; ++i) printf(" %d", next[i]);
    printf("\nfather_prefix_id:");
    for (int i = 0; i < size; ++i) printf(" %d", father_prefix_id[i]);
    printf("\nrestrictions:");
    for (const auto& p : restrict_pair) printf(" (%d, %d)", p.first, p.second);
    printf("\nbreak_size:");
    for (int i = 0; i < size * (size - 1) / 2; i++) printf(" %d", break_size[i]);
    printf("\n");
}

// Restriction candidates for the order `cur_adj_mat`
// (src/schedule_IEP.cpp:1737-1760), filtered per the validation mode.
void Schedule_IEP::restricts_generate(const int* cur_adj_mat,
                                      std::vector<std::vector<std::pair<int, int> > >& restricts) {
    Schedule_IEP plain(cur_adj_mat, get_size());
    plain.aggressive_optimize_get_all_pairs(restricts);
    int n = plain.get_size();
    if (g_validation_mode == RestrictValidation::Reference) {
        long long want = plain.count_on_complete_graph(n + 1) / plain.get_multiplicity();
        for (size_t i = 0; i < restricts.size();) {
            Schedule_IEP cur(plain.get_adj_mat_ptr(), n);
            cur.add_restrict(restricts[i]);
            if (cur.count_on_complete_graph(n + 1) != want) restricts.erase(restricts.begin() + i);
            else ++i;
        }
    } else {
        // With an IEP tail of k vertices the terminal ignores the pairs whose
        // later vertex is in the tail; the head pairs must then leave every
        // occurrence the same multiplicity, which the redundancy divides out.
        int k = in_exclusion_optimize_num;
        for (size_t i = 0; i < restricts.size();) {
            bool keep = restrictions_exactly_valid(cur_adj_mat, n, restricts[i]) &&
                        (k < 2 || iep_head_multiplicity(cur_adj_mat, n, restricts[i], k) > 0);
            if (!keep) restricts.erase(restricts.begin() + i);
            else ++i;
        }
    }
}

// Number of automorphic relabellings of an occurrence that satisfy the pairs
// whose later vertex lies before the IEP tail, if it is the same for every
// relative order of the n embedded ids; otherwise -1.
int Schedule_IEP::iep_head_multiplicity(const int* adj, int n, const std::vector<std::pair<int, int> >& pairs, int k) {
    std::vector<std::pair<int, int> > head;
    for (const auto& p : pairs)
        if (p.second < n - k) head.push_back(p);
    std::vector<std::vector<int> > aut;
    for (const auto& p : all_permutations(n)) {
        bool same = true;
        for (int i = 0; i < n && same; ++i)
            for (int j = 0; j < n; ++j)
                if (adj[INDEX(i, j, n)] != adj[INDEX(p[i], p[j], n)]) { same = false; break; }
        if (same) aut.push_back(p);
    }
    int mult = -1;
    for (const auto& rank : all_permutations(n)) {
        int hits = 0;
        for (const auto& s : aut) {
            bool ok = true;
            for (const auto& pr : head)
                if (!(rank[s[pr.second]] < rank[s[pr.first]])) { ok = false; break; }
            hits += ok;
        }
        if (mult == -1) mult = hits;
        else if (hits != mult) return -1;
    }
    return mult;
}

// Valid iff for every relative order of the n embedded ids exactly one
// automorphic relabelling satisfies all pairs (a, b) : emb[b] < emb[a].
bool Schedule_IEP::restrictions_exactly_valid(const int* adj, int n, const std::vector<std::pair<int, int> >& pairs) {
    std::vector<std::vector<int> > aut;
    for (const auto& p : all_permutations(n)) {
        bool same = true;
        for (int i = 0; i < n && same; ++i)
            for (int j = 0; j < n; ++j)
                if (adj[INDEX(i, j, n)] != adj[INDEX(p[i], p[j], n)]) { same = false; break; }
        if (same) aut.push_back(p);
    }
    for (const auto& rank : all_permutations(n)) {
        int hits = 0;
        for (const auto& s : aut) {
            bool ok = true;
            for (const auto& pr : pairs)
                if (!(rank[s[pr.second]] < rank[s[pr.first]])) { ok = false; break; }
            hits += ok;
        }
        if (hits != 1) return false;
    }
    return true;
}

// Deepest IEP tail of order `vec`, or -1 if the order is disconnected
// (src/schedule_IEP.cpp:1762-1789).
int Schedule_IEP::get_vec_optimize_num(const std::vector<int>& vec) {
    for (int i = 1; i < size; ++i) {
        bool linked = false;
        for (int j = 0; j < i && !linked; ++j) linked = adj_mat[INDEX(vec[i], vec[j], size)] != 0;
        if (!linked) return -1;
    }
    for (int k = 2; k <= size - 2; ++k) {
        for (int i = size - k + 1; i < size; ++i)
            if (adj_mat[INDEX(vec[size - k], vec[i], size)]) return k - 1;
    }
    return size - 2;
}

void Schedule_IEP::remove_invalid_permutation(std::vector<std::vector<int> >& cands) {
    for (size_t c = 0; c < cands.size();) {
        const auto& vec = cands[c];
        bool ok = true;
        for (int x = 1; x < size && ok; ++x) {
            bool linked = false;
            for (int y = 0; y < x && !linked; ++y) linked = adj_mat[INDEX(vec[x], vec[y], size)] != 0;
            ok = linked;
        }
        if (ok) ++c;
        else cands.erase(cands.begin() + c);
    }
}

int Schedule_IEP::get_in_exclusion_optimize_num_when_not_optimize() {
    std::vector<int> identity(size);
    for (int i = 0; i < size; ++i) identity[i] = i;
    return get_vec_optimize_num(identity);
}

// Divisor for IEP double counting: count on K_n with and without the IEP
// tail (src/schedule_IEP.cpp:2050-2068).
void Schedule_IEP::set_in_exclusion_optimize_redundancy() {
    int k = get_in_exclusion_optimize_num();
    in_exclusion_optimize_redundancy = 1;
    if (k <= 1) return;
    long long with_iep = count_on_complete_graph(size);
    set_in_exclusion_optimize_num(0);
    long long without = count_on_complete_graph(size);
    set_in_exclusion_optimize_num(k);
    if (without == 0) throw std::logic_error("Schedule_IEP: pattern has no embedding in K_n");
    if (g_validation_mode == RestrictValidation::Exact && with_iep % without != 0)
        throw std::logic_error("Schedule_IEP: IEP multiplicity is not an integer for this restriction set");
    in_exclusion_optimize_redundancy = with_iep / without;
}
