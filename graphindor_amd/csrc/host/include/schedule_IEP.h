// Schedule_IEP: the schedule compiler.  Chooses the vertex order of a
// pattern, its symmetry-breaking restrictions and the inclusion-exclusion
// (IEP) depth, and builds the loop invariants (prefix trie, break sizes, IEP
// tables) the enumeration kernels follow.
//
// Public interface identical to the reference's (include/schedule_IEP.h:11-143)
// so that callers written against it compile unchanged.
//
// Restriction validation.  The reference validates candidate restriction sets
// by counting with a vertex-induced schedule on the complete graph K_{n+1}
// (src/schedule_IEP.cpp:1737-1760), which accepts every set for a non-clique
// pattern (its induced count there is 0 with or without restrictions) and can
// therefore pick a set that drops embeddings (the 5-cycle does).  Two modes:
//   RestrictValidation::Exact (default): keep a set iff, for every relative
//     order of the n embedded vertex ids, exactly one automorphism image
//     satisfies it -- the definition of a valid symmetry breaking -- and,
//     with an IEP tail of k vertices (whose pairs the IEP terminal cannot
//     apply), iff the pairs on the first n-k vertices leave every relative
//     order the same number of satisfying images, so that the K_n
//     redundancy ratio is exact.  Orders without such a set are skipped (the
//     reference would fall back to no restrictions at all).  The IEP tables
//     and prefix flags are rebuilt from scratch for every order.
//   RestrictValidation::Reference: the reference's check, and its table
//     reuse across candidate orders (IEP vertex flags/coefs/positions and
//     Prefix flags accumulate, src/schedule_IEP.cpp:517-605), reproduced
//     bit for bit; used to pin this compiler against the reference.
// The mode is process-wide; the environment variable
// GRAPHSET_RESTRICT_VALIDATION=reference selects Reference at start-up.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "disjoint_set_union.h"
#include "pattern.h"
#include "prefix.h"
#include "types.h"

enum class RestrictValidation { Exact = 0, Reference = 1 };

// One candidate schedule of a pattern, as the analytic selector sees it: a
// vertex order (new vertex k is pattern vertex order[k]), a restriction set
// over the reordered vertices (emb[second] < emb[first]), whether the IEP
// tail is used, and the analytic cost of the pair.
struct ScheduleCandidate {
    std::vector<int> order;
    std::vector<std::pair<int, int> > pairs;
    bool use_iep = true;
    double analytic_cost = 0;
};

class Schedule_IEP {
public:
    // performance_modeling_type: 0 = keep the given order, 1 = analytic model
    // (src/schedule_IEP.cpp:307-382); restricts_type: 0 none, 1 generated +
    // validated, 2 GraphZero's (src/schedule_IEP.cpp:890-940).
    Schedule_IEP(const Pattern& pattern, bool& is_pattern_valid, int performance_modeling_type,
                 int restricts_type, bool use_in_exclusion_optimize, int v_cnt, int64_t e_cnt,
                 long long tri_cnt = 0, bool vertex_induced = false);
    // Order as given, no restrictions, no IEP, vertex-induced
    // (src/schedule_IEP.cpp:249-291).
    Schedule_IEP(const int* _adj_mat, int _size);
    // Exactly the given candidate (order, restrictions, IEP on/off).  The
    // restriction set is validated exactly; an invalid one throws
    // std::invalid_argument.  Lets a selector other than the analytic model
    // choose among the candidates of enumerate_candidates().
    Schedule_IEP(const Pattern& pattern, const ScheduleCandidate& candidate);
    // Every (order, restriction set) the analytic selector of the first
    // constructor weighs with performance_modeling_type = 1 and
    // restricts_type = 1 (deepest-IEP connected orders, each exactly valid
    // restriction set; the orders without IEP when none survives), with its
    // analytic cost, in the order the selector visits them.  Exact mode.
    static std::vector<ScheduleCandidate> enumerate_candidates(const Pattern& pattern, bool use_in_exclusion_optimize,
                                                               int v_cnt, int64_t e_cnt, long long tri_cnt);
    Schedule_IEP(const Schedule_IEP&) = delete;
    Schedule_IEP& operator=(const Schedule_IEP&) = delete;
    ~Schedule_IEP();

    inline int get_total_prefix_num() const { return total_prefix_num; }
    inline int get_basic_prefix_num() const { return basic_prefix_num; }
    inline int get_father_prefix_id(int prefix_id) const { return father_prefix_id[prefix_id]; }
    inline int* get_father_prefix_id_ptr() const { return father_prefix_id; }
    inline int get_loop_set_prefix_id(int loop) const { return loop_set_prefix_id[loop]; }
    inline int* get_loop_set_prefix_id_ptr() const { return loop_set_prefix_id; }
    inline bool get_prefix_only_need_size(int prefix_id) const { return prefix[prefix_id].get_only_need_size(); }
    inline int get_size() const { return size; }
    inline int get_last(int i) const { return last[i]; }
    inline int* get_last_ptr() const { return last; }
    inline int get_next(int i) const { return next[i]; }
    inline int* get_next_ptr() const { return next; }
    inline int* get_break_size_ptr() const { return break_size; }
    inline int get_prefix_target(int i) const { return prefix_target[i]; }
    inline int* get_prefix_target_ptr() const { return prefix_target; }
    inline int get_in_exclusion_optimize_num() const { return in_exclusion_optimize_num; }
    int get_in_exclusion_optimize_num_when_not_optimize();
    void add_restrict(const std::vector<std::pair<int, int> >& restricts);
    inline int get_total_restrict_num() const { return total_restrict_num; }
    inline int get_restrict_last(int i) const { return restrict_last[i]; }
    inline int* get_restrict_last_ptr() const { return restrict_last; }
    inline int get_restrict_next(int i) const { return restrict_next[i]; }
    inline int* get_restrict_next_ptr() const { return restrict_next; }
    inline int get_restrict_index(int i) const { return restrict_index[i]; }
    inline int* get_restrict_index_ptr() const { return restrict_index; }
    inline int get_k_val() const { return k_val; }
    int get_max_degree() const;
    int get_multiplicity() const;
    void aggressive_optimize(std::vector<std::pair<int, int> >& ordered_pairs) const;
    void aggressive_optimize_get_all_pairs(std::vector<std::vector<std::pair<int, int> > >& ordered_pairs_vector);
    void restricts_generate(const int* cur_adj_mat, std::vector<std::vector<std::pair<int, int> > >& restricts);
    void GraphZero_aggressive_optimize(std::vector<std::pair<int, int> >& ordered_pairs) const;
    void GraphZero_get_automorphisms(std::vector<std::vector<int> >& Aut) const;
    std::vector<std::vector<int> > get_isomorphism_vec() const;
    static std::vector<std::vector<int> > calc_permutation_group(const std::vector<int> vec, int size);
    inline const int* get_adj_mat_ptr() const { return adj_mat; }

    inline void set_in_exclusion_optimize_redundancy(long long redundancy) { in_exclusion_optimize_redundancy = redundancy; }
    inline long long get_in_exclusion_optimize_redundancy() const { return in_exclusion_optimize_redundancy; }

    void print_schedule() const;

    // Count this schedule's embeddings on the complete graph K_m with the
    // CPU engine semantics of Graph::pattern_matching (src/graph.cpp:305-566).
    // m <= 8; the schedule compiler's own use (restriction validation and
    // the IEP redundancy), not a counting path for user graphs.
    long long count_on_complete_graph(int m) const;

    // True iff the restriction set breaks every automorphism exactly once.
    static bool restrictions_exactly_valid(const int* adj_mat, int size,
                                           const std::vector<std::pair<int, int> >& pairs);
    // Multiplicity left by the restrictions on the first size-k vertices when
    // it is the same for every relative order, else -1.
    static int iep_head_multiplicity(const int* adj_mat, int size, const std::vector<std::pair<int, int> >& pairs,
                                     int k);

    static void set_restrict_validation(RestrictValidation mode);
    static RestrictValidation get_restrict_validation();

    std::vector<std::pair<int, int> > restrict_pair;

    std::vector<int> in_exclusion_optimize_vertex_id;
    std::vector<bool> in_exclusion_optimize_vertex_flag;
    std::vector<int> in_exclusion_optimize_vertex_coef;

    std::vector<int> in_exclusion_optimize_coef;
    std::vector<bool> in_exclusion_optimize_flag;
    std::vector<int> in_exclusion_optimize_ans_pos;

    int* break_size = nullptr;

    bool is_vertex_induced;

private:
    void allocate_tables();
    bool check_connectivity() const;
    void setup_optimization_info(bool use_in_exclusion_optimize = true);
    void copy_adj_mat_from(const std::vector<int>& vec, const int* src_adj_mat);
    void build_loop_invariant(int in_exclusion_optimize_num = 0);
    int find_father_prefix(int data_size, const int* data);
    double new_estimate_schedule_restrict(const std::vector<std::pair<int, int> >& restrictions, int v_cnt,
                                          e_index_t e_cnt, long long tri_cnt);
    void get_in_exclusion_optimize_group(int depth, int* id, int id_cnt, int* in_exclusion_val);
    void init_in_exclusion_optimize();
    int get_vec_optimize_num(const std::vector<int>& vec);
    void remove_invalid_permutation(std::vector<std::vector<int> >& candidate_permutations);
    inline void set_in_exclusion_optimize_num(int num) { in_exclusion_optimize_num = num; }
    void set_in_exclusion_optimize_redundancy();

    int* adj_mat = nullptr;
    int* father_prefix_id = nullptr;
    int* last = nullptr;
    int* next = nullptr;
    int* loop_set_prefix_id = nullptr;
    int* prefix_target = nullptr;
    Prefix* prefix = nullptr;
    int* restrict_last = nullptr;
    int* restrict_next = nullptr;
    int* restrict_index = nullptr;
    int size;
    int total_prefix_num;
    int basic_prefix_num;
    int total_restrict_num;
    int in_exclusion_optimize_num;
    int k_val;
    long long in_exclusion_optimize_redundancy;

    std::vector<std::vector<std::vector<int> > > in_exclusion_optimize_group;
    std::vector<int> in_exclusion_optimize_val;
};
