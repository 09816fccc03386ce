/*
 * graphindor_amd.h -- C ABI of the MI355X (gfx950) pattern-mining engine.
 *
 * This is the boundary between the host C++ library (Graph / Pattern /
 * Schedule_IEP and the gpu_graph / gpu_clique / gpu_mc command lines) and
 * the HIP kernels in libgm_hip.so.  Signatures carry plain pointers, sizes
 * and POD structs only.
 *
 * What it replaces in the reference (sth1997/GraphIndor):
 *   - the file-local device driver `pattern_matching(Graph*, const Schedule_IEP&)`
 *     of gpu/gpu_graph.cu:22-51 and gpu/gpu_mc.cu:29-59, and
 *     `pattern_matching_init` of gpu/gpu_clique.cu:366-570;
 *   - `PatternMatchingDeviceContext::init/destroy`
 *     (gpu/function/pattern_matching.cuh:18-104): CSR upload, scratch, counters;
 *   - `GPUSchedule::create_from_schedule` (gpu/component/gpu_schedule.cuh:89-190):
 *     the flattened schedule ("plan"), whose fields gm_plan_desc mirrors.
 *
 * Conventions: every int-returning entry point returns GM_OK (0) on success
 * and a negative GM_ERR_* code otherwise; it never calls exit().  The message
 * of the last failure on the calling thread is gm_last_error().  Handles are
 * immutable after creation; one handle may be used from one host thread at a
 * time.  The library owns all device memory; callers own their host buffers.
 * The division by the IEP redundancy stays with the caller, as in the
 * reference (gpu/gpu_graph.cu:44, src/graph.cpp:354).
 */
#ifndef GRAPHINDOR_AMD_H
#define GRAPHINDOR_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GM_OK 0
#define GM_ERR_INVALID (-1)     /* bad argument or malformed plan/graph */
#define GM_ERR_HIP (-2)         /* HIP runtime error (message says which) */
#define GM_ERR_OOM (-3)         /* device allocation failed */
#define GM_ERR_UNSUPPORTED (-4) /* pattern too large, no device, ... */
#define GM_ERR_TIMEOUT (-5)     /* gm_tuning.time_limit_ms passed; the sums are incomplete */

/* Largest pattern the kernels enumerate; the reference GPU DFS stops at
 * MAX_DEPTH = 7 (gpu/function/pattern_matching.cuh:7). */
#define GM_MAX_PATTERN 7
/* Prefix-table length for GM_MAX_PATTERN: n(n-1)/2 + 1
 * (src/schedule_IEP.cpp:38-39). */
#define GM_MAX_PREFIX (GM_MAX_PATTERN * (GM_MAX_PATTERN - 1) / 2 + 1)
/* IEP table bounds for an IEP tail of k <= GM_MAX_PATTERN - 2 = 5 vertices
 * (src/schedule_IEP.cpp:1762-1789 allows k up to n-2).  The vertex table
 * has one row per non-empty subset of the tail: at most 2^5 - 1 = 31.  The
 * answer array has one entry per block of every set partition of the tail:
 * sum of block counts over the partitions of 5 = Bell(6) - Bell(5) = 151. */
#define GM_MAX_IEP_VERTEX 32
#define GM_MAX_IEP 152

/* Counting semantics of a plan. */
#define GM_MODE_PATTERN 0 /* gpu_graph / gpu_mc: restrictions, IEP, embedding subtraction
                             (gpu/src/gpu_pattern_matching.cuh:14-104) */
#define GM_MODE_CLIQUE 1  /* gpu_clique: graph already reduced to one direction per edge;
                             the last level adds |loop set| (gpu/gpu_clique.cu:215-258) */

typedef struct gm_graph gm_graph;
typedef struct gm_plan gm_plan;

/* Flattened schedule; field-for-field the GPUSchedule of
 * gpu/component/gpu_schedule.cuh:41-84.  Arrays indexed by prefix id have
 * n(n-1)/2+1 entries, arrays indexed by pattern vertex have n entries. */
typedef struct gm_plan_desc {
    int32_t size;                      /* pattern vertices n, 3 <= n <= GM_MAX_PATTERN */
    int32_t total_prefix_num;
    int32_t basic_prefix_num;
    int32_t total_restrict_num;
    int32_t in_exclusion_optimize_num; /* IEP depth k (0 = off) */
    int32_t is_vertex_induced;
    int32_t mode;                      /* GM_MODE_* */
    const int32_t* adj_mat;            /* n*n row-major 0/1 */
    const int32_t* father_prefix_id;   /* [P] -1 = root (aliases a neighbour list) */
    const int32_t* last;               /* [n] head of prefixes built when vertex i is fixed */
    const int32_t* next;               /* [P] next prefix in that list, -1 ends */
    const int32_t* break_size;         /* [P] prune when |prefix| == break_size */
    const int32_t* loop_set_prefix_id; /* [n] candidate set of vertex i */
    const int32_t* restrict_last;      /* [n] head of restrictions checked at depth i */
    const int32_t* restrict_next;      /* [P] */
    const int32_t* restrict_index;     /* [P] emb[i] must be < emb[restrict_index] */
    const uint8_t* only_need_size;     /* [P] size-only intersection */
    int32_t iep_vertex_num;            /* I <= GM_MAX_IEP_VERTEX */
    const int32_t* iep_vertex_id;      /* [I] prefix whose size feeds the IEP */
    const uint8_t* iep_vertex_flag;    /* [I] 1: ans = |S| - coef; 0: ans = |S \ emb| */
    const int32_t* iep_vertex_coef;    /* [I] */
    int32_t iep_array_size;            /* A <= GM_MAX_IEP */
    const int32_t* iep_coef;           /* [A] group coefficient on the closing entry */
    const uint8_t* iep_flag;           /* [A] 1 closes a product group */
    const int32_t* iep_ans_pos;        /* [A] index into ans[] */
} gm_plan_desc;

/* Launch tuning; zero fields mean "library default". */
typedef struct gm_tuning {
    int32_t waves_per_cu;      /* resident waves per CU of the persistent grid (0 = all that fit).
                                  The pattern rule has two builds: one holding its DFS state in
                                  registers, 8 waves per CU, and one limited to 128 VGPRs with some
                                  state spilled to scratch, 16 per CU; 1..8 selects the first, more
                                  than 8 the second, 0 the second for launches of at least 16384
                                  root edges and the first below (gm_count_static always the first) */
    int32_t chunk_edges;       /* most root edges one work-steal claims (claims shrink towards 1 as the
                                  work runs out; 0 = 64) */
    int32_t lds_set_cap;       /* GM_MODE_PATTERN: the last candidate level stages its father sets in
                                  LDS when they hold at most this many elements in all
                                  (0 = default 2048, the per-wave arena; -1 = never) */
    int32_t bitmap_max_degree; /* GM_MODE_CLIQUE: roots of out-degree up to this use the
                                  local-bitmap engine (0 = default 1024, -1 = off) */
    int32_t lane_tail;         /* GM_MODE_PATTERN: last candidate depth counted one candidate per lane
                                  when its intersections walk at most this many elements, wave-wide
                                  otherwise (0 = default 64, -1 = off: every depth wave-wide) */
    int32_t root_stride;       /* GM_MODE_PATTERN: count only the root edges e with e % root_stride == 0
                                  (a uniform sample for timing, identical however the range is split;
                                  0 or 1 = every edge) */
    int32_t slice_len;         /* GM_MODE_PATTERN: once no root is left to claim, a wave whose current
                                  root still has more than 4*slice_len depth-2 candidates keeps
                                  slice_len of them and donates the rest, in slices of slice_len, to
                                  idle waves, which claim them from a queue (0 = default 8, -1 = off) */
    int32_t bitmap_block_degree; /* GM_MODE_CLIQUE: bitmap roots of out-degree above this whose rows fit
                                  in one block's LDS are counted by a whole block each (0 = default
                                  64, -1 = off: every bitmap root by one wave) */
    int32_t time_limit_ms;     /* GM_MODE_PATTERN / list engine: waves stop once they have run this long
                                  and the call fails with GM_ERR_TIMEOUT (0 = no limit; for timing
                                  harnesses that must not wait on a slow candidate) */
    int32_t hub_degree;        /* GM_MODE_PATTERN: vertices of degree at least this carry their
                                  neighbour set as a bitmap over vertex ids, so that a lookup in a hub's
                                  list is one load (0 or -1 = none, the default: whether they pay
                                  depends on the pattern, so a tuner measures it; the bitmaps, n/8
                                  bytes per hub, at most 4 GiB, are built on the graph at first use) */
} gm_tuning;

/* Per-call measurements, filled when the pointer is non-null. */
typedef struct gm_stats {
    double kernel_ms;         /* HIP-event time of the whole counting region on the stream */
    uint64_t alg_bytes;       /* algorithmic bytes (only when GM_COUNT_ALG_BYTES is set) */
    uint64_t root_tasks;      /* root edges the grid claimed */
    int32_t grid_waves;       /* persistent waves launched */
    int32_t flags_used;
    double list_ms;           /* HIP-event time inside the list-engine launches (enumerate kernel) */
    double bitmap_ms;         /* HIP-event time inside the local-bitmap clique engine launches */
    uint64_t bitmap_alg_bytes;/* the bitmap engine's share of alg_bytes */
    double bitmap_block_ms;   /* of bitmap_ms: the block-per-root launches (large roots) */
    uint64_t bitmap_block_alg_bytes; /* of bitmap_alg_bytes: the block-per-root launches */
    /* Depth-2 and depth-3 slice donation (tuning.slice_len > 0), list engine only: */
    uint64_t slices;          /* slices pushed to the queue */
    uint64_t donations;       /* donations (each pushes one or more slices) */
    uint64_t queue_retries;   /* queue-head claims lost to another wave */
    double slice_wave_ms;     /* wall time waves spent on donated slices, summed over waves */
    double idle_wave_ms;      /* wall time waves spent waiting for slices, summed over waves */
    int64_t hub_vertices;     /* hub bitmaps the pattern rule used (gm_tuning.hub_degree) */
    uint64_t deep_donations;  /* of `donations`, those of depth-3 slices: the candidates below one
                                 depth-2 candidate, donated while waves wait on an empty queue */
} gm_stats;

/* gm_count flags */
#define GM_COUNT_ALG_BYTES 1      /* also tally the algorithmic byte model (slower) */
#define GM_COUNT_DEEP_DONATION 4  /* reserved: depth-3 slice donation (slices of the depth-3 loop
                                     set below one depth-2 candidate).  Switched off -- gm_count and
                                     gm_batch_create refuse it with GM_ERR_UNSUPPORTED -- because a
                                     31-root 5-cycle sample of R-MAT 2^22 that takes 8 s without it
                                     did not finish in 180 s with it */
#define GM_COUNT_EAGER_DONATION 2 /* donate depth-2 slices whenever a set is long, not only once the
                                     roots run out (for tests and tuning; counts are unchanged) */

/* Upload a CSR graph (row_ptr[n+1] int64, col[nnz] int32, rows sorted) to
 * the current device.  max_isect bounds every materialised intersection
 * (the .g header's max_intersection_size, src/dataloader.cpp:100; the
 * reference sizes scratch from it, gpu/function/pattern_matching.cuh:25).
 * Replaces the CSR part of PatternMatchingDeviceContext::init
 * (gpu/function/pattern_matching.cuh:27-61). */
int gm_graph_create(const int64_t* row_ptr, const int32_t* col, int32_t n, int64_t nnz,
                    int32_t max_isect, gm_graph** out);
void gm_graph_destroy(gm_graph* g);

/* Copy a flattened schedule to the device.  Replaces
 * GPUSchedule::create_from_schedule (gpu/component/gpu_schedule.cuh:89-190). */
int gm_plan_create(const gm_plan_desc* desc, gm_plan** out);
void gm_plan_destroy(gm_plan* p);

/* Count embeddings: raw sum before the division by the IEP redundancy.
 * Enqueues on `stream` (a hipStream_t, NULL = default stream) and
 * synchronises it before returning.  Replaces the launch + sync + 8-byte
 * readback of gpu/gpu_graph.cu:38-42 (gpu/gpu_clique.cu:522-529 for
 * GM_MODE_CLIQUE). */
int gm_count(const gm_graph* g, const gm_plan* p, const gm_tuning* tuning, int32_t flags,
             void* stream, uint64_t* raw_sum, gm_stats* stats);

/* gm_count without allocating: no reference counterpart; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Every device buffer the
 * count needs (counters, the sum cell, the list engine's per-wave prefix
 * slots) is carved from one static per-device workspace of
 * GM_STATIC_WORKSPACE_BYTES; the list engine runs as many waves as have room
 * for their prefix slots there (fewer than gm_count on graphs with long
 * lists), slice donation is off, and stats->kernel_ms is host wall time
 * (list_ms, bitmap_ms are 0).  Sums equal gm_count's.  Fails with
 * GM_ERR_UNSUPPORTED when one block of waves does not fit (second-largest
 * degree x prefixes too large).  Calls on one device share the workspace:
 * they are serialised within a process, and calls from several processes
 * on one GPU race. */
#define GM_STATIC_WORKSPACE_BYTES (64u << 20)
int gm_count_static(const gm_graph* g, const gm_plan* p, const gm_tuning* tuning, int32_t flags,
                    void* stream, uint64_t* raw_sum, gm_stats* stats);

/* Count several plans over one graph (e.g. every connected k-motif) in one
 * persistent launch; raw_sums[np].  Replaces the per-motif init + launch
 * loop of gpu/gpu_mc.cu:96-115. */
int gm_count_multi(const gm_graph* g, const gm_plan* const* plans, int32_t np,
                   const gm_tuning* tuning, int32_t flags, void* stream, uint64_t* raw_sums,
                   gm_stats* stats);

/* Count np plans over a batch of ng graphs held as one CSR with globally
 * numbered vertices (graph i owns the root edges [graph_edge_offsets[i],
 * graph_edge_offsets[i+1]); offsets[0] = 0, offsets[ng] = nnz), in one
 * persistent launch per counting rule.  raw_sums[i * np + j] is graph i's raw
 * sum for plan j.  Replaces running gpu_mc once per graph (gpu/gpu_mc.cu:62-117). */
int gm_count_batch(const gm_graph* g, const int64_t* graph_edge_offsets, int32_t ng,
                   const gm_plan* const* plans, int32_t np, const gm_tuning* tuning, int32_t flags,
                   void* stream, uint64_t* raw_sums, gm_stats* stats);

/* The same batch count captured once into a hipGraph (counter resets by
 * kernels, then the fused persistent launch) and replayed by gm_batch_run.
 * Pattern-rule plans only; the tuning (root_stride included) is fixed at
 * creation.  The graph and plans must outlive the batch; the
 * batch owns its scratch, so runs of different batches do not share state. */
typedef struct gm_batch gm_batch;
int gm_batch_create(const gm_graph* g, const int64_t* graph_edge_offsets, int32_t ng,
                    const gm_plan* const* plans, int32_t np, const gm_tuning* tuning, int32_t flags,
                    gm_batch** out);
int gm_batch_run(gm_batch* b, void* stream, uint64_t* raw_sums, gm_stats* stats);
void gm_batch_destroy(gm_batch* b);

/* ---- Learned schedule cost model ------------------------------------------
 * No counterpart in the reference: the schedule compiler's analytic model
 * (new_estimate_schedule_restrict, src/schedule_IEP.cpp:307-382) ranks the
 * candidate (order, restriction set) pairs by formula; this FP32 MLP ranks
 * them from measured run times instead, with the analytic model as the
 * fallback.  Network: F inputs, two ReLU layers of H units, one output;
 * parameters laid out as W1[F*H] b1[H] W2[H*H] b2[H] W3[H] b3[1] (row-major,
 * W1[f*H + h]).  Training minimises the group-centred squared error
 * sum_g sum_{i in g} ((y_i - t_i) - mean_{j in g}(y_j - t_j))^2 / rows with
 * full-batch Adam (beta 0.9/0.999, eps 1e-8, decoupled weight decay), one
 * step captured in a hipGraph and replayed.  Rows of a group are contiguous:
 * group g is rows [group_off[g], group_off[g+1]).  All calls synchronise. */
typedef struct gm_mlp gm_mlp;
int gm_mlp_create(int32_t in_dim, int32_t hidden, uint64_t seed, gm_mlp** out);
void gm_mlp_destroy(gm_mlp* m);
int64_t gm_mlp_param_count(const gm_mlp* m);
int gm_mlp_get_params(const gm_mlp* m, float* out);
/* Gradients of the last training step's loss, same layout as the parameters. */
int gm_mlp_get_grads(const gm_mlp* m, float* out);
/* Loads parameters and resets the optimiser state. */
int gm_mlp_set_params(gm_mlp* m, const float* in);
int gm_mlp_predict(gm_mlp* m, const float* x, int32_t rows, float* y);
/* x[rows*F], t[rows]; final_loss (may be null) = loss before the last step. */
int gm_mlp_train(gm_mlp* m, const float* x, const float* t, const int32_t* group_off, int32_t groups, int32_t steps,
                 float lr, float weight_decay, float* final_loss);
const char* gm_mlp_last_error(void);

/* Number of visible HIP devices (0 without a GPU); never fails. */
int gm_device_count(void);

/* Message of the last failed call on this thread ("" if none). */
const char* gm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GRAPHINDOR_AMD_H */
