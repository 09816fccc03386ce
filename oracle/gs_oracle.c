/*
 * gs_oracle.c -- CPU oracle for the pattern-mining hot path.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load this file's library (oracle/liboracle.so); the
 * product never does.  It restates, in plain C with OpenMP, the counting
 * rules the device kernels must reproduce bit for bit:
 *
 *   oracle_count_vertex_rooted  Graph::pattern_matching + pattern_matching_aggressive_func
 *                               (src/graph.cpp:305-355, 429-566), VertexSet
 *                               (src/vertex_set.cpp:83-159, 203-215, 256-277).
 *   oracle_count_edge_rooted    the reference GPU driver's rule: root = a directed
 *                               edge (gpu/src/gpu_pattern_matching.cuh:14-104), DFS of
 *                               gpu/function/pattern_matching.cuh:110-288, set ops of
 *                               gpu/component/gpu_vertex_set.cuh:67-248; it also
 *                               tallies the algorithmic byte model (below).
 *   oracle_count_clique         gpu_clique's rule (gpu/gpu_clique.cu:215-364).
 *   oracle_brute_force          schedule-free: injective edge-preserving maps of the
 *                               pattern (optionally also non-edge-preserving for
 *                               vertex-induced) divided by |Aut|; pins everything above.
 *
 * Pinning (see tests/): wiki-vote triangles 608389 and 4-cliques 2077903
 * (include/dataloader.h:32 and an independent brute force); closed forms on
 * complete graphs; brute force on seeded small graphs; the four rules agree
 * with each other; and, in this container, the reference's own
 * Graph::pattern_matching built from source (oracle/_ref).
 *
 * Algorithmic byte model (per launch, summed over the enumeration):
 *   root edge processed          8 (endpoints) + 16 (row pointers)
 *   loop element consumed        4; +8 row pointers when it is expanded
 *   materialised C = A ∩ B       4 (|A| + |B| + |C|)
 *   size-only |A ∩ B|            4 (|A| + |B|)
 *   subtraction / membership     4 |emb|
 * Size-only prefixes are derived from the plan structure (no children, not a
 * loop set before the IEP depth, every IEP row reading them has flag 1), the
 * same rule the device uses.
 */
#include <limits.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "../include/graphindor_amd.h"

typedef struct {
    int32_t n;
    int64_t nnz;
    const int64_t* vertex;
    const int32_t* edge;
} or_graph;

/* ------------------------------------------------------------------------ */

typedef struct {
    const int32_t* ptr;
    int32_t len;
} or_set;

typedef struct {
    const or_graph* g;
    const gm_plan_desc* p;
    int32_t* store;   /* P + n buffers of cap ints */
    int64_t cap;
    or_set* sets;     /* P + n (slot P+depth: vertex-induced scratch) */
    uint32_t emb[GM_MAX_PATTERN + 1];
    int emb_n;
    int64_t ans[GM_MAX_IEP_VERTEX];
    uint64_t sum;
    uint64_t bytes;
    uint8_t size_only[GM_MAX_PREFIX];
    int gpu_rule;     /* 1: edge-rooted device rule, 0: CPU engine rule */
    int32_t d1_stride, d1_offset; /* CPU rule: expand only depth-1 candidates i = offset mod stride */
    int32_t d2_lo, d2_hi;         /* device rule: walk only the depth-2 candidates [d2_lo, d2_hi) */
    int tail_depth;   /* device rule: builds at this depth are charged as size-only */
} or_ctx;

static int64_t max_degree(const or_graph* g) {
    int64_t m = 0;
    for (int32_t v = 0; v < g->n; ++v) {
        int64_t d = g->vertex[v + 1] - g->vertex[v];
        if (d > m) m = d;
    }
    return m;
}

/* Size-only rule shared with the device. */
static void derive_size_only(const gm_plan_desc* p, uint8_t* out) {
    int P = p->total_prefix_num;
    int tail = p->size - p->in_exclusion_optimize_num;
    for (int i = 0; i < P; ++i) out[i] = 0;
    if (p->mode == GM_MODE_CLIQUE) {
        /* only the size of the last loop set is read */
        int q = p->loop_set_prefix_id[p->size - 1], used = 0;
        for (int c = 0; c < P; ++c) used |= (p->father_prefix_id[c] == q);
        for (int d = 1; d < p->size - 1; ++d) used |= (p->loop_set_prefix_id[d] == q);
        if (q >= 0 && p->father_prefix_id[q] >= 0 && !used) out[q] = 1;
        return;
    }
    if (p->in_exclusion_optimize_num == 0) return;
    for (int i = 0; i < p->iep_vertex_num; ++i) out[p->iep_vertex_id[i]] = 1;
    for (int i = 0; i < p->iep_vertex_num; ++i)
        if (!p->iep_vertex_flag[i]) out[p->iep_vertex_id[i]] = 0;
    for (int q = 0; q < P; ++q)
        if (p->father_prefix_id[q] >= 0) out[p->father_prefix_id[q]] = 0;
    for (int d = 1; d < tail; ++d)
        if (p->loop_set_prefix_id[d] >= 0) out[p->loop_set_prefix_id[d]] = 0;
}

static int in_emb(const or_ctx* c, uint32_t v) {
    for (int i = 0; i < c->emb_n; ++i)
        if (c->emb[i] == v) return 1;
    return 0;
}

static int bsearch_u32(const int32_t* a, int32_t n, uint32_t x) {
    int32_t lo = 0, hi = n - 1;
    while (lo <= hi) {
        int32_t mid = lo + ((hi - lo) >> 1);
        uint32_t y = (uint32_t)a[mid];
        if (y == x) return 1;
        if (y < x) lo = mid + 1;
        else hi = mid - 1;
    }
    return 0;
}

/* First index with a[i] >= x (unsigned compare, as the device's lower_bound). */
static int32_t lower_bound_u32(const int32_t* a, int32_t n, uint32_t x) {
    int32_t lo = 0, hi = n;
    while (lo < hi) {
        int32_t mid = lo + ((hi - lo) >> 1);
        if ((uint32_t)a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* |set[0:limit) \ emb| */
static int32_t subtraction_size(or_ctx* c, const int32_t* set, int32_t limit) {
    int32_t r = limit;
    for (int i = 0; i < c->emb_n; ++i)
        if (bsearch_u32(set, limit, c->emb[i])) --r;
    c->bytes += 4ull * (uint64_t)c->emb_n;
    return r;
}

/* Depth whose builds the device counts lane-parallel (pattern_tail in the
 * engine): the last candidate depth, >= 2, at most 4 prefixes built there,
 * at most 8 IEP rows, not vertex-induced; -1 otherwise. */
static int tail_depth_of(const gm_plan_desc* p) {
    if (p->mode != GM_MODE_PATTERN || p->is_vertex_induced) return -1;
    int k = p->in_exclusion_optimize_num;
    int D = k >= 2 ? p->size - k - 1 : (k == 0 ? p->size - 2 : -1);
    if (D < 2) return -1;
    int np = 0;
    for (int q = p->last[D]; q != -1; q = p->next[q]) ++np;
    if (np > 4) return -1;
    if (k >= 2 && p->iep_vertex_num > 8) return -1;
    return D;
}

/* Sorted intersection of a[0:na) and b[0:nb) into out (NULL: count only).
 * Lists of comparable length are merged; when one is at least 32 times
 * longer, each element of the shorter is found in it by galloping from the
 * previous hit (hub neighbour lists), which gives the same elements. */
static int32_t isect(const int32_t* a, int32_t na, const int32_t* b, int32_t nb, int32_t* out) {
    int32_t k = 0;
    if ((int64_t)na * 32 <= nb || (int64_t)nb * 32 <= na) {
        const int32_t* s = na <= nb ? a : b;
        const int32_t* l = na <= nb ? b : a;
        const int32_t ns = na <= nb ? na : nb, nl = na <= nb ? nb : na;
        int32_t pos = 0;
        for (int32_t i = 0; i < ns && pos < nl; ++i) {
            const int32_t x = s[i];
            int32_t step = 1, hi = pos;
            while (hi < nl && l[hi] < x) { pos = hi + 1; hi += step; step <<= 1; }
            if (hi > nl) hi = nl;
            while (pos < hi) {  /* first index in [pos, hi) with l[.] >= x */
                int32_t mid = pos + ((hi - pos) >> 1);
                if (l[mid] < x) pos = mid + 1;
                else hi = mid;
            }
            if (pos < nl && l[pos] == x) {
                if (out) out[k] = x;
                ++k;
                ++pos;
            }
        }
        return k;
    }
    int32_t i = 0, j = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) ++i;
        else if (a[i] > b[j]) ++j;
        else { if (out) out[k] = a[i]; ++k; ++i; ++j; }
    }
    return k;
}

static void build(or_ctx* c, int pid, const int32_t* in, int32_t n_in) {
    int f = c->p->father_prefix_id[pid];
    if (f == -1) {
        c->sets[pid].ptr = in;
        c->sets[pid].len = n_in;
        return;
    }
    const int32_t* a = c->sets[f].ptr;
    int32_t na = c->sets[f].len;
    int32_t k;
    if (c->gpu_rule && c->size_only[pid]) {
        k = isect(a, na, in, n_in, NULL);
        c->sets[pid].ptr = NULL;
        c->sets[pid].len = k;
        c->bytes += 4ull * (uint64_t)(na + n_in);
        return;
    }
    int32_t* out = c->store + (int64_t)pid * c->cap;
    k = isect(a, na, in, n_in, out);
    c->sets[pid].ptr = out;
    c->sets[pid].len = k;
    c->bytes += 4ull * (uint64_t)(na + n_in);
    if (!(c->gpu_rule && c->emb_n == c->tail_depth)) c->bytes += 4ull * (uint64_t)k;
}

static void iep_terminal(or_ctx* c) {
    const gm_plan_desc* p = c->p;
    for (int i = 0; i < p->iep_vertex_num; ++i) {
        or_set s = c->sets[p->iep_vertex_id[i]];
        if (p->iep_vertex_flag[i]) c->ans[i] = (int32_t)(s.len - p->iep_vertex_coef[i]);
        else c->ans[i] = subtraction_size(c, s.ptr, s.len);
    }
    int last_pos = -1;
    int64_t val = 0;
    for (int pos = 0; pos < p->iep_array_size; ++pos) {
        if (pos == last_pos + 1) val = c->ans[p->iep_ans_pos[pos]];
        else if (val != 0) val = (int64_t)((uint64_t)val * (uint64_t)c->ans[p->iep_ans_pos[pos]]);
        if (p->iep_flag[pos]) {
            last_pos = pos;
            c->sum += (uint64_t)val * (uint64_t)(int64_t)p->iep_coef[pos];
        }
    }
}

/* remove_anti_edge_vertices (src/graph.cpp:195-232,
 * gpu/function/pattern_matching.cuh:145-206) into slot P + depth. */
static or_set remove_anti_edge(or_ctx* c, or_set in, int depth) {
    const gm_plan_desc* p = c->p;
    int32_t* out = c->store + (int64_t)(p->total_prefix_num + depth) * c->cap;
    int32_t k = 0;
    for (int32_t i = 0; i < in.len; ++i) {
        int keep = 1;
        for (int u = 0; u < depth && keep; ++u) {
            if (p->adj_mat[u * p->size + depth] != 0) continue;
            uint32_t v = c->emb[u];
            int64_t l = c->g->vertex[v], r = c->g->vertex[v + 1];
            if (bsearch_u32(c->g->edge + l, (int32_t)(r - l), (uint32_t)in.ptr[i])) keep = 0;
        }
        if (keep) out[k++] = in.ptr[i];
    }
    or_set o = {out, k};
    return o;
}

static uint32_t restriction_bound(const or_ctx* c, int depth, uint32_t init) {
    uint32_t b = init;
    for (int i = c->p->restrict_last[depth]; i != -1; i = c->p->restrict_next[i]) {
        uint32_t e = c->emb[c->p->restrict_index[i]];
        if (e < b) b = e;
    }
    return b;
}

static void dfs(or_ctx* c, int depth) {
    const gm_plan_desc* p = c->p;
    int n = p->size, k = p->in_exclusion_optimize_num;
    or_set loop = c->sets[p->loop_set_prefix_id[depth]];
    if (!c->gpu_rule) {
        /* CPU engine: empty loop set returns before anything else. */
        if (loop.len <= 0) return;
        if (p->is_vertex_induced) loop = remove_anti_edge(c, loop, depth);
        if (depth == n - k) { iep_terminal(c); return; }
    } else {
        if (depth == n - k) { iep_terminal(c); return; }
        if (p->is_vertex_induced) loop = remove_anti_edge(c, loop, depth);
    }
    if (depth == n - 1 && k == 0) {
        if (c->gpu_rule) {
            uint32_t b = restriction_bound(c, depth, 0xffffffffu);
            int32_t lim = lower_bound_u32(loop.ptr, loop.len, b);
            c->sum += (uint64_t)(int64_t)subtraction_size(c, loop.ptr, lim);
        } else if (p->total_restrict_num > 0) {
            uint32_t b = restriction_bound(c, depth, (uint32_t)c->g->n);
            int32_t lim = lower_bound_u32(loop.ptr, loop.len, b);
            if (lim > 0) c->sum += (uint64_t)(int64_t)subtraction_size(c, loop.ptr, lim);
        } else {
            c->sum += (uint64_t)(int64_t)subtraction_size(c, loop.ptr, loop.len);
        }
        return;
    }
    uint32_t bound = restriction_bound(c, depth, c->gpu_rule ? 0xffffffffu : (uint32_t)c->g->n);
    const int sampled = depth == 1 && c->d1_stride > 1;
    int32_t i0 = 0, i1 = loop.len;
    if (depth == 2) {
        i0 = c->d2_lo;
        if (c->d2_hi < i1) i1 = c->d2_hi;
    }
    for (int32_t i = i0; i < i1; ++i) {
        uint32_t v = (uint32_t)loop.ptr[i];
        if (bound <= v) break;
        if (sampled && i % c->d1_stride != c->d1_offset) continue;
        c->bytes += 4ull;
        if (in_emb(c, v)) continue;
        int64_t l = c->g->vertex[v], r = c->g->vertex[v + 1];
        c->bytes += 8ull;
        int prune = 0;
        for (int pid = p->last[depth]; pid != -1; pid = p->next[pid]) {
            build(c, pid, c->g->edge + l, (int32_t)(r - l));
            if (c->sets[pid].len == p->break_size[pid]) { prune = 1; break; }
        }
        if (prune) continue;
        c->emb[c->emb_n++] = v;
        dfs(c, depth + 1);
        --c->emb_n;
    }
}

static or_ctx* ctx_new(const or_graph* g, const gm_plan_desc* p, int gpu_rule) {
    or_ctx* c = (or_ctx*)calloc(1, sizeof(or_ctx));
    c->g = g;
    c->p = p;
    c->cap = max_degree(g) + 1;
    c->store = (int32_t*)malloc(sizeof(int32_t) * (size_t)c->cap * (size_t)(p->total_prefix_num + p->size));
    c->sets = (or_set*)calloc((size_t)(p->total_prefix_num + p->size), sizeof(or_set));
    c->gpu_rule = gpu_rule;
    c->tail_depth = gpu_rule ? tail_depth_of(p) : -1;
    c->d2_lo = 0;
    c->d2_hi = INT32_MAX;
    derive_size_only(p, c->size_only);
    return c;
}

static void ctx_free(or_ctx* c) {
    free(c->store);
    free(c->sets);
    free(c);
}

static int plan_ok(const gm_plan_desc* p) {
    return p && p->size >= 2 && p->size <= GM_MAX_PATTERN && p->total_prefix_num <= GM_MAX_PREFIX &&
           p->iep_vertex_num <= GM_MAX_IEP_VERTEX && p->iep_array_size <= GM_MAX_IEP;
}

/* Raw sum (before the redundancy division) with the CPU engine's rule,
 * expanding only the depth-1 candidates with index = offset mod stride of
 * every root (stride 1: everything).  Summed over all offsets this is the
 * full count; one offset is an even sample of the work, hub roots included. */
uint64_t oracle_count_vertex_rooted_strided(const or_graph* g, const gm_plan_desc* p, int threads, int32_t stride,
                                            int32_t offset) {
    if (!plan_ok(p) || stride < 1) return 0;
    uint64_t total = 0;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel reduction(+ : total)
    {
        or_ctx* c = ctx_new(g, p, 0);
        c->d1_stride = stride;
        c->d1_offset = offset % stride;
#pragma omp for schedule(dynamic)
        for (int32_t v0 = 0; v0 < g->n; ++v0) {
            const int32_t* nb = g->edge + g->vertex[v0];
            int32_t d = (int32_t)(g->vertex[v0 + 1] - g->vertex[v0]);
            for (int pid = p->last[0]; pid != -1; pid = p->next[pid]) build(c, pid, nb, d);
            c->emb[0] = (uint32_t)v0;
            c->emb_n = 1;
            dfs(c, 1);
        }
        total += c->sum;
        ctx_free(c);
    }
    return total;
}

uint64_t oracle_count_vertex_rooted(const or_graph* g, const gm_plan_desc* p, int threads) {
    return oracle_count_vertex_rooted_strided(g, p, threads, 1, 0);
}

/* Root edge index -> source vertex, for a contiguous range of edges. */
static int32_t edge_source(const or_graph* g, int64_t e) {
    int32_t lo = 0, hi = g->n - 1;
    while (lo < hi) {
        int32_t mid = lo + ((hi - lo + 1) >> 1);
        if (g->vertex[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

/* Is depth 2 a candidate loop under the device rule (and so splittable)? */
static int depth2_loops(const gm_plan_desc* p) {
    const int n = p->size, k = p->in_exclusion_optimize_num;
    if (p->mode == GM_MODE_CLIQUE) return n >= 4;
    return n - k > 2 && !(k == 0 && n == 3);
}

/* Depth-0/1 builds of root e (bytes charged); returns the length of its
 * depth-2 loop set, or -1 when the root is skipped or ends at depth 1. */
static int32_t edge_root_prefixes(or_ctx* c, const or_graph* g, const gm_plan_desc* p, int32_t v0, int64_t e) {
    uint32_t v1 = (uint32_t)g->edge[e];
    if (p->restrict_last[1] != -1 && (uint32_t)v0 <= v1) return -1;
    c->bytes += 24ull;
    c->emb_n = 0;
    const int32_t* n0 = g->edge + g->vertex[v0];
    int32_t d0 = (int32_t)(g->vertex[v0 + 1] - g->vertex[v0]);
    for (int pid = p->last[0]; pid != -1; pid = p->next[pid]) build(c, pid, n0, d0);
    const int32_t* n1 = g->edge + g->vertex[v1];
    int32_t d1 = (int32_t)(g->vertex[v1 + 1] - g->vertex[v1]);
    for (int pid = p->last[1]; pid != -1; pid = p->next[pid]) {
        build(c, pid, n1, d1);
        if (c->sets[pid].len == 0 && pid < p->basic_prefix_num) return -1;
    }
    c->emb[0] = (uint32_t)v0;
    c->emb[1] = v1;
    c->emb_n = 2;
    return c->sets[p->loop_set_prefix_id[2]].len;
}

/* One root edge e = (v0, edge[e]) under the device rule: its sum and bytes go
 * to c->sum and c->bytes. */
static void edge_root(or_ctx* c, const or_graph* g, const gm_plan_desc* p, int32_t v0, int64_t e) {
    uint32_t v1 = (uint32_t)g->edge[e];
    if (p->restrict_last[1] != -1 && (uint32_t)v0 <= v1) return;
    c->bytes += 24ull;
    c->emb_n = 0; /* root-level builds (depths 0 and 1) */
    const int32_t* n0 = g->edge + g->vertex[v0];
    int32_t d0 = (int32_t)(g->vertex[v0 + 1] - g->vertex[v0]);
    for (int pid = p->last[0]; pid != -1; pid = p->next[pid]) build(c, pid, n0, d0);
    const int32_t* n1 = g->edge + g->vertex[v1];
    int32_t d1 = (int32_t)(g->vertex[v1 + 1] - g->vertex[v1]);
    for (int pid = p->last[1]; pid != -1; pid = p->next[pid]) {
        build(c, pid, n1, d1);
        if (c->sets[pid].len == 0 && pid < p->basic_prefix_num) return;
    }
    c->emb[0] = (uint32_t)v0;
    c->emb[1] = v1;
    c->emb_n = 2;
    if (p->mode == GM_MODE_CLIQUE) {
        extern void oracle_clique_dfs(or_ctx*, int);
        oracle_clique_dfs(c, 2);
    } else {
        dfs(c, 2);
    }
}

/* Raw sum with the device rule over the root edges e with e % root_stride
 * == 0 (the device's sampling rule, gm_tuning.root_stride; 1 counts every
 * edge); *alg_bytes receives the byte model of those roots. */
uint64_t oracle_count_edge_rooted_strided(const or_graph* g, const gm_plan_desc* p, int threads, int64_t root_stride,
                                          uint64_t* alg_bytes) {
    if (!plan_ok(p) || root_stride < 1) return 0;
    uint64_t total = 0, bytes = 0;
    const int64_t nroots = (g->nnz + root_stride - 1) / root_stride;
    if (threads > 0) omp_set_num_threads(threads);
    /* A root's depth-2 candidates are split into up to 64 pieces of at least
     * 64 candidates, so that the threads share a hub root's work; a piece
     * rebuilds the root's prefixes, and only the first is charged for them. */
    int32_t* pieces = (int32_t*)calloc((size_t)(nroots > 0 ? nroots : 1), sizeof(int32_t));
    const int split = depth2_loops(p);
    if (split) {
#pragma omp parallel
        {
            or_ctx* c = ctx_new(g, p, 1);
#pragma omp for schedule(dynamic, 64)
            for (int64_t ri = 0; ri < nroots; ++ri) {
                const int64_t e = ri * root_stride;
                int32_t L = edge_root_prefixes(c, g, p, edge_source(g, e), e);
                int32_t k = L <= 64 ? 1 : (L + 63) / 64;
                pieces[ri] = k > 64 ? 64 : k;
            }
            ctx_free(c);
        }
    } else {
        for (int64_t ri = 0; ri < nroots; ++ri) pieces[ri] = 1;
    }
    int64_t* first = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nroots + 1));
    first[0] = 0;
    for (int64_t ri = 0; ri < nroots; ++ri) first[ri + 1] = first[ri] + pieces[ri];
    const int64_t nitems = first[nroots];
#pragma omp parallel reduction(+ : total, bytes)
    {
        or_ctx* c = ctx_new(g, p, 1);
#pragma omp for schedule(dynamic)
        for (int64_t it = 0; it < nitems; ++it) {
            int64_t lo = 0, hi = nroots;  /* root of item it: largest ri with first[ri] <= it */
            while (hi - lo > 1) {
                int64_t mid = (lo + hi) / 2;
                if (first[mid] <= it) lo = mid;
                else hi = mid;
            }
            const int64_t ri = lo, e = ri * root_stride;
            const int32_t piece = (int32_t)(it - first[ri]), np_ = pieces[ri];
            const int32_t v0 = edge_source(g, e);
            if (np_ == 1) {
                edge_root(c, g, p, v0, e);
                continue;
            }
            const uint64_t b0 = c->bytes;
            const int32_t L = edge_root_prefixes(c, g, p, v0, e);
            if (piece > 0) c->bytes = b0;
            if (L < 0) continue;
            c->d2_lo = (int32_t)((int64_t)L * piece / np_);
            c->d2_hi = (int32_t)((int64_t)L * (piece + 1) / np_);
            if (p->mode == GM_MODE_CLIQUE) {
                extern void oracle_clique_dfs(or_ctx*, int);
                oracle_clique_dfs(c, 2);
            } else {
                dfs(c, 2);
            }
            c->d2_lo = 0;
            c->d2_hi = INT32_MAX;
        }
        total += c->sum;
        bytes += c->bytes;
        ctx_free(c);
    }
    free(first);
    free(pieces);
    if (alg_bytes) *alg_bytes = bytes;
    return total;
}

/* A batch of ng graphs held as one CSR with global vertex ids, graph i owning
 * the root edges [gedge[i], gedge[i+1]) (the layout gm_count_batch takes):
 * sums[i] receives graph i's raw sum over its root edges e with
 * e % root_stride == 0 -- absolute edge indices, as on the device, so the
 * sample does not depend on how the batch is cut. */
void oracle_count_batch_strided(const or_graph* g, const int64_t* gedge, int32_t ng, const gm_plan_desc* p,
                                int threads, int64_t root_stride, uint64_t* sums) {
    for (int32_t i = 0; i < ng; ++i) sums[i] = 0;
    if (!plan_ok(p) || root_stride < 1 || ng < 1) return;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
    {
        or_ctx* c = ctx_new(g, p, 1);
#pragma omp for schedule(dynamic, 1)
        for (int32_t gi = 0; gi < ng; ++gi) {
            const int64_t lo = (gedge[gi] + root_stride - 1) / root_stride * root_stride;
            c->sum = 0;
            if (lo < gedge[gi + 1]) {
                int32_t v0 = edge_source(g, lo);
                for (int64_t e = lo; e < gedge[gi + 1]; e += root_stride) {
                    while (g->vertex[v0 + 1] <= e) ++v0;
                    edge_root(c, g, p, v0, e);
                }
            }
            sums[gi] = c->sum;
        }
        ctx_free(c);
    }
}

/* Raw sum with the device rule; *alg_bytes receives the byte model. */
uint64_t oracle_count_edge_rooted(const or_graph* g, const gm_plan_desc* p, int threads, uint64_t* alg_bytes) {
    return oracle_count_edge_rooted_strided(g, p, threads, 1, alg_bytes);
}

/* gpu_clique's DFS: no restrictions, no membership test, the last level
 * adds |loop set| (gpu/gpu_clique.cu:215-258). */
void oracle_clique_dfs(or_ctx* c, int depth) {
    const gm_plan_desc* p = c->p;
    or_set loop = c->sets[p->loop_set_prefix_id[depth]];
    if (depth == p->size - 1) {
        c->sum += (uint64_t)loop.len;
        return;
    }
    int32_t i0 = 0, i1 = loop.len;
    if (depth == 2) {
        i0 = c->d2_lo;
        if (c->d2_hi < i1) i1 = c->d2_hi;
    }
    for (int32_t i = i0; i < i1; ++i) {
        uint32_t v = (uint32_t)loop.ptr[i];
        int64_t l = c->g->vertex[v], r = c->g->vertex[v + 1];
        c->bytes += 12ull;
        int prune = 0;
        for (int pid = p->last[depth]; pid != -1; pid = p->next[pid]) {
            build(c, pid, c->g->edge + l, (int32_t)(r - l));
            if (c->sets[pid].len == p->break_size[pid]) { prune = 1; break; }
        }
        if (prune) continue;
        c->emb[c->emb_n++] = v;
        oracle_clique_dfs(c, depth + 1);
        --c->emb_n;
    }
}

uint64_t oracle_count_clique(const or_graph* g, const gm_plan_desc* p, int threads, uint64_t* alg_bytes) {
    gm_plan_desc q = *p;
    q.mode = GM_MODE_CLIQUE;
    return oracle_count_edge_rooted(g, &q, threads, alg_bytes);
}

/* ------------------------------------------------------------------------ */
/* Schedule-free brute force. */

typedef struct {
    const or_graph* g;
    int n;
    const int32_t* adj;
    int induced;
    uint32_t m[GM_MAX_PATTERN + 1];
    uint64_t count;
} bf_ctx;

static int has_edge(const or_graph* g, uint32_t a, uint32_t b) {
    int64_t l = g->vertex[a], r = g->vertex[a + 1];
    return bsearch_u32(g->edge + l, (int32_t)(r - l), b);
}

static void bf_rec(bf_ctx* c, int d) {
    if (d == c->n) {
        ++c->count;
        return;
    }
    if (d >= GM_MAX_PATTERN) return; /* unreachable: n <= GM_MAX_PATTERN */
    int anchor = -1;
    for (int j = 0; j < d; ++j)
        if (c->adj[j * c->n + d]) { anchor = j; break; }
    const int32_t* cand;
    int64_t nc;
    if (anchor >= 0) {
        uint32_t a = c->m[anchor];
        cand = c->g->edge + c->g->vertex[a];
        nc = c->g->vertex[a + 1] - c->g->vertex[a];
    } else {
        cand = NULL;
        nc = c->g->n;
    }
    for (int64_t i = 0; i < nc; ++i) {
        uint32_t v = cand ? (uint32_t)cand[i] : (uint32_t)i;
        int ok = 1;
        for (int j = 0; j < d && ok; ++j) {
            if (c->m[j] == v) ok = 0;
            else if (c->adj[j * c->n + d]) ok = has_edge(c->g, c->m[j], v);
            else if (c->induced) ok = !has_edge(c->g, c->m[j], v);
        }
        if (!ok) continue;
        c->m[d] = v;
        bf_rec(c, d + 1);
    }
}

static uint64_t automorphisms(int n, const int32_t* adj) {
    int perm[GM_MAX_PATTERN];
    for (int i = 0; i < n; ++i) perm[i] = i;
    uint64_t count = 0;
    for (;;) {
        int same = 1;
        for (int i = 0; i < n && same; ++i)
            for (int j = 0; j < n; ++j)
                if ((adj[i * n + j] != 0) != (adj[perm[i] * n + perm[j]] != 0)) { same = 0; break; }
        count += (uint64_t)same;
        /* next permutation */
        int i = n - 2;
        while (i >= 0 && perm[i] >= perm[i + 1]) --i;
        if (i < 0) break;
        int j = n - 1;
        while (perm[j] <= perm[i]) --j;
        int t = perm[i]; perm[i] = perm[j]; perm[j] = t;
        for (int a = i + 1, b = n - 1; a < b; ++a, --b) { t = perm[a]; perm[a] = perm[b]; perm[b] = t; }
    }
    return count;
}

/* Occurrences of the pattern: injective maps preserving its edges (and, if
 * `induced`, its non-edges) divided by |Aut|. */
uint64_t oracle_brute_force(const or_graph* g, int n, const int32_t* adj, int induced, int threads) {
    if (n < 1 || n > GM_MAX_PATTERN) return 0;
    uint64_t total = 0;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic) reduction(+ : total)
    for (int32_t v0 = 0; v0 < g->n; ++v0) {
        bf_ctx c;
        memset(&c, 0, sizeof(c));
        c.g = g;
        c.n = n;
        c.adj = adj;
        c.induced = induced;
        c.m[0] = (uint32_t)v0;
        bf_rec(&c, 1);
        total += c.count;
    }
    return total / automorphisms(n, adj);
}

/* The clique rule split the way the device splits it: roots v0 of out-degree
 * d <= dmax go to the local-bitmap engine, whose byte model is
 * 16 + 4 d + sum over u in N(v0) of (8 + 4 |N(u)|) for roots with d >= k-1;
 * the edges of the other roots take the list rule with its own model.  The
 * count is the clique rule's either way. */
uint64_t oracle_count_clique_hybrid(const or_graph* g, const gm_plan_desc* p, int threads, int32_t dmax,
                                    uint64_t* alg_bytes) {
    gm_plan_desc q = *p;
    q.mode = GM_MODE_CLIQUE;
    int k = p->size;
    uint64_t total = 0, bytes = 0;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel reduction(+ : total, bytes)
    {
        or_ctx* c = ctx_new(g, &q, 1);
#pragma omp for schedule(dynamic, 64)
        for (int32_t v0 = 0; v0 < g->n; ++v0) {
            int64_t l0 = g->vertex[v0], r0 = g->vertex[v0 + 1];
            int32_t d0 = (int32_t)(r0 - l0);
            int small = d0 <= dmax;
            if (small && d0 < k - 1) continue;
            uint64_t b0 = c->bytes;
            for (int64_t e = l0; e < r0; ++e) {
                uint32_t v1 = (uint32_t)g->edge[e];
                c->bytes += 24ull;
                const int32_t* n0 = g->edge + l0;
                for (int pid = q.last[0]; pid != -1; pid = q.next[pid]) build(c, pid, n0, d0);
                const int32_t* n1 = g->edge + g->vertex[v1];
                int32_t d1 = (int32_t)(g->vertex[v1 + 1] - g->vertex[v1]);
                int zero = 0;
                for (int pid = q.last[1]; pid != -1; pid = q.next[pid]) {
                    build(c, pid, n1, d1);
                    if (c->sets[pid].len == 0 && pid < q.basic_prefix_num) { zero = 1; break; }
                }
                if (zero) continue;
                c->emb[0] = (uint32_t)v0;
                c->emb[1] = v1;
                c->emb_n = 2;
                oracle_clique_dfs(c, 2);
            }
            if (small) {
                c->bytes = b0 + 16ull + 4ull * (uint64_t)d0;
                for (int64_t e = l0; e < r0; ++e) {
                    uint32_t u = (uint32_t)g->edge[e];
                    c->bytes += 8ull + 4ull * (uint64_t)(g->vertex[u + 1] - g->vertex[u]);
                }
            }
        }
        total += c->sum;
        bytes += c->bytes;
        ctx_free(c);
    }
    if (alg_bytes) *alg_bytes = bytes;
    return total;
}
