// Device side of the enumeration engine: plan and argument layouts, set
// operations, the schedule-driven DFS, the enumeration kernel template and
// the local-bitmap clique kernel template.  Included by gm_engine.hip (host
// half and C ABI) and by enum_instance.hip, which instantiates one
// enumerate_kernel<MODE, NMAX> per object so the instantiations compile in
// parallel.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "graphindor_amd.h"

namespace gmk {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;
constexpr int kMaxSets = GM_MAX_PREFIX + GM_MAX_PATTERN;  // prefixes + per-depth filtered loop sets

// Plan as the device reads it (wave-uniform scalar loads).
struct DevPlan {
    int32_t n, P, basic, nrestrict, k, vinduced, mode, iep_vn, iep_an;
    int32_t adj[GM_MAX_PATTERN * GM_MAX_PATTERN];
    int32_t father[GM_MAX_PREFIX], last[GM_MAX_PATTERN], next[GM_MAX_PREFIX], brk[GM_MAX_PREFIX];
    int32_t loopset[GM_MAX_PATTERN], rlast[GM_MAX_PATTERN], rnext[GM_MAX_PREFIX], rindex[GM_MAX_PREFIX];
    int32_t size_only[GM_MAX_PREFIX];
    int32_t clique_tail;  // clique mode: depth n-2 counts lane-parallel (see clique_tail_count)
    int32_t clique_chain; // clique mode on the plain k-clique chain: local-bitmap engine eligible
    int32_t tail_depth;   // pattern mode: depth counted lane-parallel by pattern_tail (-1: none)
    int32_t iep_id[GM_MAX_IEP_VERTEX], iep_vflag[GM_MAX_IEP_VERTEX], iep_vcoef[GM_MAX_IEP_VERTEX];
    int32_t iep_coef[GM_MAX_IEP], iep_flag[GM_MAX_IEP], iep_pos[GM_MAX_IEP];
};

struct SetRef {
    const uint32_t* ptr;
    const unsigned long long* bm;  // the set as a bitmap over vertex ids (a hub's neighbour list), or null
    int32_t len;
    int32_t pad;
};

// Elements of a clique-mode tail loop set staged in LDS per wave.
constexpr int kStage = 2048;

// Pattern-rule last level (pattern_tail): prefixes and IEP rows it handles.
constexpr int kTailMaxPrefix = 4;
constexpr int kTailMaxIep = 8;

// Everything the last level reads that does not depend on the candidate.
struct TailCtx {
    int32_t pid[kTailMaxPrefix];
    const uint32_t* fptr[kTailMaxPrefix];
    const unsigned long long* fbm[kTailMaxPrefix];  // the father's hub bitmap, or null
    int32_t flen[kTailMaxPrefix];
    bool alias[kTailMaxPrefix];
    int32_t base[kTailMaxIep];
    const uint32_t* sptr[kTailMaxIep];
    int32_t slen[kTailMaxIep];
    int32_t at[kTailMaxIep];  // position of the IEP row's prefix in pid[], or -1
    int qat;                  // position of the last loop set in pid[] (no IEP), or -1
    SetRef qset;              // the last loop set when it was built above D
};

struct WaveLds {
    SetRef sets[kMaxSets];
    int32_t ans[GM_MAX_IEP_VERTEX];
    TailCtx tail;  // pattern_tail's candidate-independent state (LDS, not registers)
};

template <int MODE>
struct WaveShared {
    WaveLds w;
    uint32_t stage[kStage];
};

struct Counters {
    unsigned long long bm_cursor;  // local-bitmap engine's root cursor
    unsigned long long cursor;     // enumerate kernel's task cursor
    unsigned long long bytes;
    unsigned long long tasks;
    unsigned int overflow;
    unsigned int timed_out;        // a wave passed the launch's time limit
    unsigned long long big_cursor; // block-per-root bitmap kernel's vertex cursor
    unsigned long long sm_cursor;  // small-root bitmap kernel's vertex cursor
    // The slice queue's counters on a cache line of their own, away from the
    // root cursor every wave claims from.
    alignas(128) unsigned long long qhead;  // slice queue: next slot to pop
    unsigned long long qtail;      // slice queue: slots reserved by pushers
    unsigned long long done_items; // (root, plan) items fully processed, donations included
    unsigned long long slices_done; // queued slices fully processed, their own donations included
    unsigned long long donations;  // donate_slices calls that queued slices
    unsigned long long deep_donations; // of those, donations of depth-3 slices
    unsigned long long retries;    // queue-head CAS attempts lost to another popper
    unsigned long long slice_ticks; // wall-clock ticks waves spent on donated slices (sum over waves)
    unsigned long long idle_ticks;  // wall-clock ticks waves spent waiting for slices (sum over waves)
};

// A slice of one root edge's work, donated by the wave that found a loop
// set long: candidates [s0, s1) of the depth-2 loop set of root edge e under
// plan pl of graph g and, below each of them, candidates [t0, t1) of the
// depth-3 loop set.  A depth-2 slice has t0 = 0, t1 = kAllCandidates; a
// depth-3 slice has s1 = s0 + 1 (one depth-2 candidate) and t0 > 0.
// `ready` is set last by the pusher and cleared by the popper, so the queue
// is all-clear between launches.
constexpr int32_t kAllCandidates = 0x7fffffff;
struct SliceItem {
    int64_t e;
    int32_t pl, g, s0, s1, t0, t1;
    int32_t ready;
};

struct KernelArgs {
    const int64_t* row;
    const uint32_t* col;
    int32_t nv;
    int64_t nnz;
    const DevPlan* const* plans;  // np plans counted in one launch
    int32_t np;
    const int64_t* gedge;         // ng+1 root-edge offsets: graph g owns edges [gedge[g], gedge[g+1])
    int32_t ng;
    unsigned long long* sums;     // [ng * np] raw sums
    int32_t tail_off;             // 1: never use the lane-parallel pattern tail
    int32_t tail_work;            // pattern tail: candidates walking more elements than this go wave-wide
    int32_t task_max;             // most work items one claim takes
    int64_t root_begin, root_end; // root edges counted: [gedge[0], gedge[ng])
    int64_t root_stride;          // counted root edges: root_begin + i * root_stride < root_end (root_begin is
                                  // the range's first multiple of root_stride)
    int64_t nroots;               // number of counted root edges
    uint32_t* scratch;
    int64_t wave_stride;  // uint32 per wave
    int32_t cap;          // capacity of a prefix slot
    int32_t cap_vi;       // capacity of a vertex-induced slot
    Counters* ctr;
    int32_t stats;
    int64_t root_min_outdeg;  // clique mode: skip roots v0 whose |N(v0)| <= this (taken by the bitmap engine)
    SliceItem* queue;         // slice queue (null: no donation)
    int64_t queue_cap;
    int32_t slice_len;        // depth-2 candidates per donated slice
    int32_t eager_donation;   // 1: donate long remainders at once, not only when roots run out
    int32_t deep_donation;    // 1: depth-3 loop sets are donated too, and slices count themselves done
    int64_t time_limit_ticks; // wall-clock ticks a wave may run (0: no limit)
    const unsigned long long* hub_bits;  // neighbour bitmaps of the hub vertices (null: none)
    const int32_t* hub_slot;             // per vertex: index of its bitmap, or -1
    int64_t hub_words;                   // 64-bit words per bitmap (ceil(nv / 64))
    int32_t stage_cap;                   // elements of tail fathers staged in LDS (<= kStage)
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int32_t unii(int32_t x) { return (int32_t)__builtin_amdgcn_readfirstlane((uint32_t)x); }
__device__ __forceinline__ int64_t uni64(int64_t x) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)x);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ uint32_t rdlane(uint32_t x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ int64_t rdlane64(int64_t x, int l) {
    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)(uint64_t)x, l);
    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((uint64_t)x >> 32), l);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int below(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Branchless search for `key` in sorted b[0:nb): the window halves with a
// trip count that depends only on nb.
__device__ __forceinline__ bool contains(const uint32_t* b, int32_t nb, uint32_t key) {
    if (nb <= 0) return false;
    const uint32_t* base = b;
    int32_t n = nb;
    while (n > 1) {
        int32_t half = n >> 1;
        base = (base[half] <= key) ? base + half : base;
        n -= half;
    }
    return *base == key;
}

// First index i with b[i] >= key (unsigned), single lane.
__device__ __forceinline__ int32_t lower_bound_u(const uint32_t* b, int32_t nb, uint32_t key) {
    int32_t lo = 0, len = nb;
    while (len > 0) {
        int32_t half = len >> 1;
        if (b[lo + half] < key) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

// Hub vertices (degree above a threshold) carry their neighbour set as a
// bitmap over vertex ids: membership in a hub's list is one load instead of
// a binary search of log2(degree) dependent loads.
__device__ __forceinline__ const unsigned long long* hub_bitmap(const KernelArgs& a, uint32_t v) {
    if (a.hub_bits == nullptr) return nullptr;
    const int32_t slot = a.hub_slot[v];
    return slot < 0 ? nullptr : a.hub_bits + (int64_t)slot * a.hub_words;
}

__device__ __forceinline__ bool bit_test(const unsigned long long* bm, uint32_t x) {
    return ((bm[x >> 6] >> (x & 63u)) & 1ull) != 0ull;
}

// x in the sorted list b[0:nb), by its bitmap when it has one.
__device__ __forceinline__ bool member_of(const uint32_t* b, int32_t nb, const unsigned long long* bm, uint32_t x) {
    return bm ? bit_test(bm, x) : contains(b, nb, x);
}

struct Wave {
    KernelArgs a;
    const DevPlan* pl;
    int tail_depth;
    WaveLds* lds;
    uint32_t* stage;
    uint32_t* scratch;
    uint32_t emb[GM_MAX_PATTERN + 1];
    unsigned long long sum;        // wave-uniform
    unsigned long long bytes;      // wave-uniform
    unsigned long long lane_sum;   // per lane (lane-parallel paths)
    unsigned long long lane_bytes; // per lane
    unsigned int overflow;
    int32_t slice_lo, slice_hi;    // depth-2 candidates this wave walks for the current root
    int32_t s3_lo, s3_hi;          // depth-3 candidates it walks below each of them
    int32_t deep;                  // 1: the current item is a depth-3 slice (its depth-2 candidate's
                                   // prefixes are rebuilt and charged by the donor)
    int32_t stats0;                // the launch's stats flag (a.stats is cleared during rebuilds)
    int32_t donate;                // 1: long depth-2 loop sets are split and donated
    int32_t donate3;               // 1: long depth-3 loop sets are split and donated
    int32_t cur_i2;                // index of the current depth-2 candidate in its loop set
    int32_t cur_pl, cur_gr;        // plan and graph of the current root (for donated slices)
    int64_t cur_e;                 // current root edge
    int64_t nitems;                // (root, plan) work items of the launch
    uint64_t t_start;              // wall clock at the wave's start
    int32_t aborted;               // 1: the time limit passed; unwind
    int64_t rebuild_ticks;         // wall-clock ticks the current root's depth-0/1 builds took
    int64_t rebuild2_ticks;        // wall-clock ticks the current depth-2 candidate's builds took
    int32_t donations;             // donations this wave made
    int32_t deep_donations;        // of those, of depth-3 slices
};

__device__ __forceinline__ SetRef get_set(const Wave& w, int pid) {
    SetRef s = w.lds->sets[pid];
    SetRef r;
    r.ptr = (const uint32_t*)uni64((int64_t)s.ptr);
    r.bm = (const unsigned long long*)uni64((int64_t)s.bm);
    r.len = unii(s.len);
    r.pad = 0;
    return r;
}

// Donation only pays for a remainder of at least this many wall-clock
// ticks (10 ns each at the 100 MHz constant clock: 500 us).  Below that,
// the queue traffic of many small donations costs more than the balance
// gains (house and the 4-motifs on R-MAT: up to 50x slower at 20 us).
constexpr int64_t kDonateMinTicks = 50000;

// Wall-clock ticks a popper pays to claim a slice before its first candidate
// (queue-head CAS under contention, the ready flag, the item's loads): a
// slice is sized to carry at least 8x this plus the prefix rebuild.  Without
// it, patterns whose prefixes cost nothing to rebuild (house, 4-motifs) were
// cut into 8-candidate slices, 10^5 per launch, whose claims serialised on
// the queue head.
constexpr int64_t kSliceClaimTicks = 300;

// Time limit: true (and the wave marked aborted) once the wave has run
// longer than the launch allows.  Uniform: every lane reads the same clock.
__device__ __forceinline__ bool out_of_time(Wave& w) {
    if (w.aborted) return true;
    if (w.a.time_limit_ticks <= 0) return false;
    const uint64_t now = uni64((int64_t)wall_clock64());
    if ((int64_t)(now - w.t_start) > w.a.time_limit_ticks) w.aborted = 1;
    return w.aborted != 0;
}

__device__ __forceinline__ void put_set(Wave& w, int pid, const uint32_t* ptr, int32_t len,
                                        const unsigned long long* bm = nullptr) {
    if (lane_id() == 0) {
        w.lds->sets[pid].ptr = ptr;
        w.lds->sets[pid].bm = bm;
        w.lds->sets[pid].len = len;
    }
    wave_sync();
}

template <int D>
__device__ __forceinline__ uint32_t emb_at(const Wave& w, int idx) {
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) r = (idx == j) ? w.emb[j] : r;
    return r;
}

template <int D>
__device__ __forceinline__ bool in_emb(const Wave& w, uint32_t v) {
    bool hit = false;
#pragma unroll
    for (int j = 0; j < D; ++j) hit |= (w.emb[j] == v);
    return hit;
}

// Wave-cooperative |a ∩ b|, compacted into out (out == nullptr: size only).
// Lane i takes element i of the shorter list and looks it up in the longer
// one: one bit test when the longer is a hub list with a bitmap (bma / bmb),
// else a branchless binary search.
__device__ __forceinline__ int32_t intersect(Wave& w, const uint32_t* a, int32_t na, const unsigned long long* bma,
                                             const uint32_t* b, int32_t nb, const unsigned long long* bmb,
                                             uint32_t* out, int32_t cap) {
    if (na > nb) {
        const uint32_t* t = a;
        a = b;
        b = t;
        int32_t tn = na;
        na = nb;
        nb = tn;
        const unsigned long long* tb = bma;
        bma = bmb;
        bmb = tb;
    }
    int32_t cnt = 0;
    if (na == 0) return 0;
    const int ln = lane_id();
    for (int32_t base = 0; base < na; base += kWave) {
        int32_t i = base + ln;
        bool hit = false;
        uint32_t x = 0;
        if (i < na) {
            x = a[i];
            hit = member_of(b, nb, bmb, x);
        }
        uint64_t m = ballot(hit);
        if (out != nullptr && hit) {
            int32_t pos = cnt + below(m);
            if (pos < cap) out[pos] = x;
        }
        cnt += (int32_t)__popcll(m);
    }
    if (out != nullptr) {
        if (cnt > cap) w.overflow = 1;
        wave_sync();
    }
    return cnt;
}

// |s[0:lim) \ emb[0:D)|
template <int D>
__device__ __forceinline__ int32_t subtract_emb(const Wave& w, const uint32_t* s, int32_t lim) {
    const int ln = lane_id();
    uint32_t key = emb_at<D>(w, ln);
    bool hit = (ln < D) && contains(s, lim, key);
    return lim - (int32_t)__popcll(ballot(hit));
}

// Prefix pid for the vertex just fixed, whose neighbour list is in[0:nin)
// (bm_in: its bitmap if it is a hub).
__device__ __forceinline__ void build(Wave& w, int pid, const uint32_t* in, int32_t nin,
                                      const unsigned long long* bm_in) {
    const DevPlan& p = *w.pl;
    int f = p.father[pid];
    if (f < 0) {
        put_set(w, pid, in, nin, bm_in);
        return;
    }
    SetRef fs = get_set(w, f);
    if (p.size_only[pid]) {
        int32_t c = intersect(w, fs.ptr, fs.len, fs.bm, in, nin, bm_in, nullptr, 0);
        put_set(w, pid, nullptr, c);
        if (w.a.stats) w.bytes += 4ull * (unsigned long long)(fs.len + nin);
        return;
    }
    uint32_t* out = w.scratch + (int64_t)pid * w.a.cap;
    int32_t c = intersect(w, fs.ptr, fs.len, fs.bm, in, nin, bm_in, out, w.a.cap);
    put_set(w, pid, out, c < w.a.cap ? c : w.a.cap);  // overflow is reported, never read past
    if (w.a.stats) w.bytes += 4ull * (unsigned long long)(fs.len + nin + c);
}

template <int D>
__device__ __forceinline__ uint32_t restriction_bound(const Wave& w) {
    const DevPlan& p = *w.pl;
    uint32_t b = 0xffffffffu;
    for (int i = p.rlast[D]; i != -1; i = p.rnext[i]) {
        uint32_t e = emb_at<D>(w, p.rindex[i]);
        b = e < b ? e : b;
    }
    return b;
}

// Vertex-induced filter (gpu/function/pattern_matching.cuh:145-206): keep the
// loop-set elements adjacent to no embedded vertex the pattern marks as a
// non-neighbour of vertex D.
template <int D>
__device__ __forceinline__ SetRef remove_anti_edge(Wave& w, SetRef in) {
    const DevPlan& p = *w.pl;
    uint32_t* out = w.scratch + (int64_t)p.P * w.a.cap + (int64_t)D * w.a.cap_vi;
    const int ln = lane_id();
    int32_t cnt = 0;
    for (int32_t base = 0; base < in.len; base += kWave) {
        int32_t i = base + ln;
        bool keep = false;
        uint32_t x = 0;
        if (i < in.len) {
            x = in.ptr[i];
            keep = true;
#pragma unroll
            for (int u = 0; u < D; ++u) {
                if (p.adj[u * p.n + D] != 0) continue;
                uint32_t v = w.emb[u];
                int64_t l = w.a.row[v], r = w.a.row[v + 1];
                if (contains(w.a.col + l, (int32_t)(r - l), x)) keep = false;
            }
        }
        uint64_t m = ballot(keep);
        if (keep) {
            int32_t pos = cnt + below(m);
            if (pos < w.a.cap_vi) out[pos] = x;
        }
        cnt += (int32_t)__popcll(m);
    }
    if (cnt > w.a.cap_vi) w.overflow = 1;
    wave_sync();
    SetRef r;
    r.ptr = out;
    r.bm = nullptr;
    r.len = cnt < w.a.cap_vi ? cnt : w.a.cap_vi;
    r.pad = 0;
    return r;
}

// IEP terminal (gpu/function/pattern_matching.cuh:110-139).
template <int D>
__device__ __forceinline__ void iep_terminal(Wave& w) {
    const DevPlan& p = *w.pl;
    for (int i = 0; i < p.iep_vn; ++i) {
        SetRef s = get_set(w, p.iep_id[i]);
        int32_t v;
        if (p.iep_vflag[i]) {
            v = s.len - p.iep_vcoef[i];
        } else {
            v = subtract_emb<D>(w, s.ptr, s.len);
            if (w.a.stats) w.bytes += 4ull * D;
        }
        if (lane_id() == 0) w.lds->ans[i] = v;
    }
    wave_sync();
    int last_pos = -1;
    long long val = 0;
    for (int pos = 0; pos < p.iep_an; ++pos) {
        long long x = unii(w.lds->ans[p.iep_pos[pos]]);
        if (pos == last_pos + 1) val = x;
        else if (val != 0) val = (long long)((unsigned long long)val * (unsigned long long)x);
        if (p.iep_flag[pos]) {
            last_pos = pos;
            w.sum += (unsigned long long)val * (unsigned long long)(long long)p.iep_coef[pos];
        }
    }
}

// Branchless membership test in an LDS-resident sorted array.
__device__ __forceinline__ bool lds_contains(const uint32_t* b, int32_t nb, uint32_t key) {
    if (nb <= 0) return false;
    int32_t base = 0, n = nb;
    while (n > 1) {
        int32_t half = n >> 1;
        base = (b[base + half] <= key) ? base + half : base;
        n -= half;
    }
    return b[base] == key;
}

// Clique rule at depth n-2: the count is sum over v in L of |L ∩ N(v)|, the
// size of the last loop set each candidate would build
// (gpu/gpu_clique.cu:224-250).  L is staged in LDS and every lane takes its
// own candidates, iterating the shorter of L and N(v) and searching the
// other.  Per candidate the byte model charges what the wave path does:
// 12 B (element + row pointers) + 4 (|L| + |N(v)|) for a size-only
// intersection.
__device__ __forceinline__ void clique_tail_count(Wave& w, SetRef L) {
    const int ln = lane_id();
    for (int32_t i = ln; i < L.len; i += kWave) w.stage[i] = L.ptr[i];
    wave_sync();
    unsigned long long acc = 0, bytes = 0;
    for (int32_t i = ln; i < L.len; i += kWave) {
        uint32_t v = w.stage[i];
        int64_t l = w.a.row[v], r = w.a.row[v + 1];
        int32_t d = (int32_t)(r - l);
        const uint32_t* nb = w.a.col + l;
        int32_t c = 0;
        if (d <= L.len) {
            for (int32_t j = 0; j < d; ++j) c += lds_contains(w.stage, L.len, nb[j]) ? 1 : 0;
        } else {
            for (int32_t j = 0; j < L.len; ++j) c += contains(nb, d, w.stage[j]) ? 1 : 0;
        }
        acc += (unsigned long long)c;
        bytes += 12ull + 4ull * (unsigned long long)(L.len + d);
    }
    w.lane_sum += acc;
    if (w.a.stats) w.lane_bytes += bytes;
    wave_sync();
}


// ---------------------------------------------------------------------------
// Last candidate level of the pattern rule, lane-parallel where it pays.
//
// At depth D = n-k-1 (IEP tail of k >= 2 vertices) or D = n-2 (no IEP) the
// sets the schedule builds for a candidate v are only read for their sizes
// and for membership of the embedded vertices: every prefix q of last[D]
// is S_q = F_q ∩ N(v) with F_q (its father) fixed above D, or N(v) itself.
// So a candidate's contribution is |S_q| for each q, the break-size prune,
// and then either the IEP terminal of depth D+1
// (gpu/function/pattern_matching.cuh:110-139) or the restricted last level
// (pattern_matching.cuh:240-257), with nothing materialised.
//
// Light candidates -- whose intersections walk at most kLaneTailWork
// elements in all -- are evaluated one per lane, each lane iterating the
// shorter list and searching the longer.  Heavy ones (hub neighbourhoods)
// would hold the whole wave up while one lane walks thousands of elements,
// so they are evaluated one at a time by the whole wave, 64 elements per
// step.  Both evaluate the same arithmetic; the fathers are staged in LDS
// when they fit.  Byte model: as the wave path, except that the builds of
// depth D are charged as size-only.
// ---------------------------------------------------------------------------
constexpr int32_t kLaneTailWork = 64;  // default light/heavy bound (gm_tuning.lane_tail > 0 sets it)

// |F ∩ N| over the elements below `limit`, iterating the shorter list (per
// lane) and looking its elements up in the other (by bitmap when it has one).
__device__ __forceinline__ int32_t lane_isect_count(const uint32_t* F, int32_t nF, const unsigned long long* fbm,
                                                    const uint32_t* Nv, int32_t nN, const unsigned long long* nbm,
                                                    uint32_t limit) {
    int32_t c = 0;
    if (nF <= nN) {
        for (int32_t t = 0; t < nF; ++t) {
            uint32_t x = F[t];
            if (x >= limit) break;
            c += member_of(Nv, nN, nbm, x) ? 1 : 0;
        }
    } else {
        for (int32_t t = 0; t < nN; ++t) {
            uint32_t x = Nv[t];
            if (x >= limit) break;
            c += member_of(F, nF, fbm, x) ? 1 : 0;
        }
    }
    return c;
}

// The same count by the whole wave (uniform arguments and result): lane i
// takes element i of each 64-element step of the shorter list.
__device__ __forceinline__ int32_t wave_isect_count(const uint32_t* F, int32_t nF, const unsigned long long* fbm,
                                                    const uint32_t* Nv, int32_t nN, const unsigned long long* nbm,
                                                    uint32_t limit) {
    const uint32_t* s = F;
    const uint32_t* l = Nv;
    const unsigned long long* lbm = nbm;
    int32_t ns = nF, nl = nN;
    if (nF > nN) {
        s = Nv;
        l = F;
        lbm = fbm;
        ns = nN;
        nl = nF;
    }
    const int32_t lim = (limit == 0xffffffffu) ? ns : lower_bound_u(s, ns, limit);
    int32_t c = 0;
    for (int32_t b = 0; b < lim; b += kWave) {
        const int32_t i = b + lane_id();
        const bool hit = i < lim && member_of(l, nl, lbm, s[i]);
        c += (int32_t)__popcll(ballot(hit));
    }
    return c;
}


// Contribution of candidate v (not embedded, below the bound) to the sum,
// and its byte charge beyond the 12 bytes of reading v and its row.  WAVE:
// every lane holds the same v and the counts are wave-wide; otherwise each
// lane evaluates its own v.
template <int D, bool WAVE>
__device__ __forceinline__ void tail_candidate(const Wave& w, const TailCtx& c, uint32_t v, const uint32_t* Nv,
                                               int32_t dv, const unsigned long long* vbm, unsigned long long& acc,
                                               unsigned long long& bytes) {
    const DevPlan& p = *w.pl;
    auto count = [&](const uint32_t* F, int32_t nF, const unsigned long long* fbm, const uint32_t* N2, int32_t n2,
                     uint32_t limit) -> int32_t {
        if constexpr (WAVE) return wave_isect_count(F, nF, fbm, N2, n2, vbm, limit);
        else return lane_isect_count(F, nF, fbm, N2, n2, vbm, limit);
    };
    int32_t sz[kTailMaxPrefix];
#pragma unroll
    for (int j = 0; j < kTailMaxPrefix; ++j) {
        sz[j] = 0;
        if (c.pid[j] == -1) continue;
        if (c.alias[j]) {
            sz[j] = dv;
        } else {
            sz[j] = count(c.fptr[j], c.flen[j], c.fbm[j], Nv, dv, 0xffffffffu);
            bytes += 4ull * (unsigned long long)(c.flen[j] + dv);
        }
        if (sz[j] == p.brk[c.pid[j]]) return;
    }
    // x in S_j  <=>  x in N(v) and (alias or x in F_j).
    auto member = [&](int j, uint32_t x) -> bool {
        bool in_n = member_of(Nv, dv, vbm, x);
        bool in_f = true;
#pragma unroll
        for (int jj = 0; jj < kTailMaxPrefix; ++jj)
            if (jj == j && !c.alias[jj]) in_f = member_of(c.fptr[jj], c.flen[jj], c.fbm[jj], x);
        return in_n && in_f;
    };
    if (p.k >= 2) {
        int32_t ans[kTailMaxIep];
#pragma unroll
        for (int e = 0; e < kTailMaxIep; ++e) {
            ans[e] = 0;
            if (e >= p.iep_vn) continue;
            const int j = c.at[e];
            if (j >= 0) {
                int32_t szj = 0;
#pragma unroll
                for (int jj = 0; jj < kTailMaxPrefix; ++jj)
                    if (jj == j) szj = sz[jj];
                if (p.iep_vflag[e]) {
                    ans[e] = szj - p.iep_vcoef[e];
                } else {
                    int32_t cnt = szj;
#pragma unroll
                    for (int x = 0; x < D; ++x) cnt -= member(j, w.emb[x]) ? 1 : 0;
                    ans[e] = cnt;
                    bytes += 4ull * (D + 1);
                }
            } else if (p.iep_vflag[e]) {
                ans[e] = c.base[e];
            } else {
                ans[e] = c.base[e] - (contains(c.sptr[e], c.slen[e], v) ? 1 : 0);
                bytes += 4ull * (D + 1);
            }
        }
        int last_pos = -1;
        long long val = 0;
        for (int pos = 0; pos < p.iep_an; ++pos) {
            const int ap = p.iep_pos[pos];
            long long x = 0;
#pragma unroll
            for (int e = 0; e < kTailMaxIep; ++e)
                if (e == ap) x = ans[e];
            if (pos == last_pos + 1) val = x;
            else if (val != 0) val = (long long)((unsigned long long)val * (unsigned long long)x);
            if (p.iep_flag[pos]) {
                last_pos = pos;
                acc += (unsigned long long)val * (unsigned long long)(long long)p.iep_coef[pos];
            }
        }
    } else {
        // Depth D+1 = n-1: |{x in S : x < bound'} \ emb'| with emb' = emb + v.
        uint32_t b1 = 0xffffffffu;
        for (int rr = p.rlast[D + 1]; rr != -1; rr = p.rnext[rr]) {
            const int idx = p.rindex[rr];
            uint32_t ev = (idx == D) ? v : emb_at<D>(w, idx);
            b1 = ev < b1 ? ev : b1;
        }
        int32_t res;
        if (c.qat >= 0) {
            bool al = false;
            const uint32_t* F = nullptr;
            const unsigned long long* Fbm = nullptr;
            int32_t nF = 0;
#pragma unroll
            for (int jj = 0; jj < kTailMaxPrefix; ++jj)
                if (jj == c.qat) {
                    al = c.alias[jj];
                    F = c.fptr[jj];
                    Fbm = c.fbm[jj];
                    nF = c.flen[jj];
                }
            res = al ? lower_bound_u(Nv, dv, b1) : count(F, nF, Fbm, Nv, dv, b1);
#pragma unroll
            for (int x = 0; x < D; ++x)
                if (w.emb[x] < b1 && member(c.qat, w.emb[x])) --res;
        } else {
            const int32_t lim = lower_bound_u(c.qset.ptr, c.qset.len, b1);
            res = lim;
#pragma unroll
            for (int x = 0; x < D; ++x)
                if (contains(c.qset.ptr, lim, w.emb[x])) --res;
            if (contains(c.qset.ptr, lim, v)) --res;
        }
        acc += (unsigned long long)(long long)res;
        bytes += 4ull * (D + 1);
    }
}

template <int D>
__device__ __forceinline__ void pattern_tail_loop(Wave& w, SetRef loop, uint32_t bound);
template <int D>
__device__ __forceinline__ int32_t maybe_donate(Wave& w, int32_t lo, int32_t base, int32_t hi, uint64_t t_loop0,
                                                int32_t step_mask);

template <int D>
__device__ __forceinline__ void pattern_tail(Wave& w, SetRef loop) {
    const DevPlan& p = *w.pl;
    const int ln = lane_id();
    const uint32_t bound = restriction_bound<D>(w);
    TailCtx c;  // built here, then kept in LDS for the candidate loop
    // Prefixes of depth D and their fathers (staged in LDS when they fit).
    int q = p.last[D];
    int32_t staged = 0;
#pragma unroll
    for (int j = 0; j < kTailMaxPrefix; ++j) {
        c.pid[j] = q;
        c.alias[j] = true;
        c.fptr[j] = nullptr;
        c.fbm[j] = nullptr;
        c.flen[j] = 0;
        if (q != -1) {
            int f = p.father[q];
            if (f >= 0) {
                SetRef fs = get_set(w, f);
                c.alias[j] = false;
                c.fptr[j] = fs.ptr;
                c.fbm[j] = fs.bm;
                c.flen[j] = fs.len;
                staged += fs.len;
            }
            q = p.next[q];
        }
    }
    if (staged <= w.a.stage_cap) {
        int32_t off = 0;
#pragma unroll
        for (int j = 0; j < kTailMaxPrefix; ++j) {
            if (c.pid[j] == -1 || c.alias[j]) continue;
            for (int32_t t = ln; t < c.flen[j]; t += kWave) w.stage[off + t] = c.fptr[j][t];
            c.fptr[j] = w.stage + off;
            off += c.flen[j];
        }
        wave_sync();
    }
    // IEP rows on sets built above D: |S| - coef, or |S \ emb[0:D)| (uniform).
#pragma unroll
    for (int e = 0; e < kTailMaxIep; ++e) {
        c.base[e] = 0;
        c.sptr[e] = nullptr;
        c.slen[e] = 0;
        c.at[e] = -1;
        if (p.k >= 2 && e < p.iep_vn) {
            int ip = p.iep_id[e];
#pragma unroll
            for (int j = 0; j < kTailMaxPrefix; ++j)
                if (c.pid[j] == ip) c.at[e] = j;
            if (c.at[e] < 0) {
                SetRef s = get_set(w, ip);
                c.sptr[e] = s.ptr;
                c.slen[e] = s.len;
                c.base[e] = p.iep_vflag[e] ? s.len - p.iep_vcoef[e] : subtract_emb<D>(w, s.ptr, s.len);
            }
        }
    }
    // Last loop set when there is no IEP tail.
    c.qat = -1;
    c.qset.ptr = nullptr;
    c.qset.bm = nullptr;
    c.qset.len = 0;
    c.qset.pad = 0;
    if (p.k == 0) {
        const int qlast = p.loopset[D + 1];
#pragma unroll
        for (int j = 0; j < kTailMaxPrefix; ++j)
            if (c.pid[j] == qlast) c.qat = j;
        if (c.qat < 0) c.qset = get_set(w, qlast);
    }
    if (ln == 0) w.lds->tail = c;
    wave_sync();
    return pattern_tail_loop<D>(w, loop, bound);
}

template <int D>
__device__ __forceinline__ void pattern_tail_loop(Wave& w, SetRef loop, uint32_t bound) {
    const int ln = lane_id();
    const TailCtx& c = w.lds->tail;
    unsigned long long acc = 0, bytes = 0;
    // At depths 2 and 3 a root's candidates may be split and donated exactly
    // as in the DFS loop: every candidate's contribution is its own.
    int32_t lo = 0, hi = loop.len;
    uint64_t t_loop0 = 0;
    if constexpr (D == 2 || D == 3) {
        const int32_t l0 = D == 2 ? w.slice_lo : w.s3_lo, h0 = D == 2 ? w.slice_hi : w.s3_hi;
        lo = l0 < loop.len ? l0 : loop.len;
        hi = h0 < loop.len ? h0 : loop.len;
        if (D == 2 ? w.donate : w.donate3) t_loop0 = (uint64_t)uni64((int64_t)wall_clock64());
    }
    for (int32_t b0 = lo; b0 < hi; b0 += kWave) {
        if (out_of_time(w)) break;
        if constexpr (D == 2 || D == 3) hi = maybe_donate<D>(w, lo, b0, hi, t_loop0, 15);
        const int32_t i = b0 + ln;
        const bool in_range = i < hi;
        const uint32_t v = in_range ? loop.ptr[i] : 0xffffffffu;
        const bool live = in_range && v < bound;
        const bool stop_after = ballot(in_range && v >= bound) != 0;
        bool heavy = false;
        if (live) {
            bytes += 4ull;
            if (!in_emb<D>(w, v)) {
                bytes += 8ull;
                const int64_t l = w.a.row[v], r = w.a.row[v + 1];
                const int32_t dv = (int32_t)(r - l);
                int32_t work = 0;
#pragma unroll
                for (int j = 0; j < kTailMaxPrefix; ++j)
                    if (c.pid[j] != -1 && !c.alias[j]) work += c.flen[j] < dv ? c.flen[j] : dv;
                heavy = work > w.a.tail_work;
                if (!heavy) tail_candidate<D, false>(w, c, v, w.a.col + l, dv, hub_bitmap(w.a, v), acc, bytes);
            }
        }
        // Heavy candidates, one at a time, wave-wide.
        uint64_t hm = ballot(heavy);
        while (hm != 0) {
            const int j = __ffsll((unsigned long long)hm) - 1;
            hm &= hm - 1;
            const uint32_t hv = rdlane(v, j);
            const int64_t l = w.a.row[hv], r = w.a.row[hv + 1];
            unsigned long long a2 = 0, b2 = 0;
            tail_candidate<D, true>(w, c, hv, w.a.col + l, (int32_t)(r - l), hub_bitmap(w.a, hv), a2, b2);
            if (ln == 0) {
                acc += a2;
                bytes += b2;
            }
        }
        if (stop_after) break;
    }
    w.lane_sum += acc;
    if (w.a.stats) w.lane_bytes += bytes;
    wave_sync();
}

// Pushes slices [lo + T, hi) of the current root's depth-D loop set (D = 2,
// or 3 below the current depth-2 candidate), T candidates each, and returns
// the end of the part this wave keeps.  Slots past the queue's capacity are
// kept too.
template <int D>
__device__ __forceinline__ int32_t donate_slices(Wave& w, int32_t lo, int32_t hi, int32_t T) {
    const int32_t keep_hi = lo + T;
    const int32_t nslices = (hi - keep_hi + T - 1) / T;
    unsigned long long base = 0;
    if (lane_id() == 0) base = atomicAdd(&w.a.ctr->qtail, (unsigned long long)nslices);
    base = (unsigned long long)uni64(__shfl((long long)base, 0));
    const int64_t cap = w.a.queue_cap;
    const int32_t queued = base >= (unsigned long long)cap ? 0
                           : (int32_t)(((int64_t)base + nslices <= cap) ? nslices : cap - (int64_t)base);
    if (queued > 0) {
        ++w.donations;
        if (D == 3) ++w.deep_donations;
    }
    // The last `queued` slices go to the queue; the ones before them that
    // did not fit stay with this wave, so its part remains contiguous.
    const int32_t first = nslices - queued;
    for (int32_t q = lane_id(); q < queued; q += kWave) {
        const int32_t k = first + q;
        SliceItem* it = w.a.queue + base + q;
        it->e = w.cur_e;
        it->pl = w.cur_pl;
        it->g = w.cur_gr;
        const int32_t c0 = keep_hi + k * T, c1 = keep_hi + (k + 1) * T < hi ? keep_hi + (k + 1) * T : hi;
        if (D == 2) {
            it->s0 = c0;
            it->s1 = c1;
            it->t0 = 0;
            it->t1 = kAllCandidates;
        } else {
            it->s0 = w.cur_i2;
            it->s1 = w.cur_i2 + 1;
            it->t0 = c0;
            it->t1 = c1;
        }
    }
    // One release fence (an L2 write-back) for the whole donation orders
    // every lane's item stores before its `ready` flags.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    for (int32_t q = lane_id(); q < queued; q += kWave)
        __hip_atomic_store(&w.a.queue[base + q].ready, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int32_t kept = keep_hi + first * T;
    return kept < hi ? kept : hi;
}

// Donation check of a depth-D loop (D = 2 or 3) at candidate `base` of
// [lo, hi): once no root is left to claim, a remainder that would keep this
// wave busy is cut into slices for idle waves; returns the new end of this
// wave's part.  Depth 3 donates only while waves wait on an empty queue:
// depth-2 slices come first, and a depth-3 slice costs its popper the
// depth-2 candidate's builds as well.  Slice length from this root's own
// costs: a slice's work (T candidates at the per-candidate time measured
// since t_loop0) is at least 8x the prefix rebuild and claim a popper pays
// for it.  The check runs at the 64-candidate steps whose index has no bit
// of step_mask set (the lane tail walks a step in about a microsecond, so it
// checks every 16th).  Wave-uniform.
template <int D>
__device__ __forceinline__ int32_t maybe_donate(Wave& w, int32_t lo, int32_t base, int32_t hi, uint64_t t_loop0,
                                                int32_t step_mask) {
    if (!(D == 2 ? w.donate : w.donate3) || hi - base <= 4 * w.a.slice_len) return hi;
    if ((((base - lo) / kWave) & step_mask) != 0) return hi;
    int32_t T = w.a.slice_len;
    bool go = w.a.eager_donation != 0;
    if (!go && base > lo) {
        unsigned long long c = 0, wait = 1;
        if (lane_id() == 0) {
            c = __hip_atomic_load(&w.a.ctr->cursor, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (D == 3) {
                const unsigned long long qh = __hip_atomic_load(&w.a.ctr->qhead, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const unsigned long long qt = __hip_atomic_load(&w.a.ctr->qtail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                wait = qh > qt ? 1 : 0;
            }
        }
        c = (unsigned long long)uni64(__shfl((long long)c, 0));
        wait = (unsigned long long)uni64(__shfl((long long)wait, 0));
        if ((int64_t)c >= w.nitems && wait) {
            const int64_t el = uni64((int64_t)wall_clock64()) - (int64_t)t_loop0;
            const int64_t per = el / (base - lo) + 1;  // ticks per candidate
            const int64_t rest = per * (int64_t)(hi - base);
            const int64_t rebuild = w.rebuild_ticks + (D == 3 ? w.rebuild2_ticks : 0);
            const int64_t want = (8 * (rebuild + kSliceClaimTicks)) / per + 1;
            T = (int32_t)(want < T ? T : (want > (hi - base) / 2 ? (hi - base) / 2 : want));
            go = rest > kDonateMinTicks && hi - base > 2 * T;
        }
    }
    return go ? donate_slices<D>(w, base, hi, T) : hi;
}

template <int D, int MODE, int NMAX>
__device__ __forceinline__ void dfs(Wave& w) {
    if constexpr (D >= NMAX) {
        return;
    } else {
        const DevPlan& p = *w.pl;
        const int n = p.n;
        if (MODE == GM_MODE_PATTERN && D == n - p.k) {
            iep_terminal<D>(w);
            return;
        }
        SetRef loop = get_set(w, p.loopset[D]);
        if (MODE == GM_MODE_PATTERN && p.vinduced) loop = remove_anti_edge<D>(w, loop);
        if (D == n - 1) {
            if (MODE == GM_MODE_CLIQUE) {
                w.sum += (unsigned long long)loop.len;
                return;
            }
            uint32_t bound = restriction_bound<D>(w);
            int32_t lim = unii(lower_bound_u(loop.ptr, loop.len, bound));
            w.sum += (unsigned long long)(long long)subtract_emb<D>(w, loop.ptr, lim);
            if (w.a.stats) w.bytes += 4ull * D;
            return;
        }
        if (MODE == GM_MODE_CLIQUE && D == n - 2 && p.clique_tail && loop.len <= kStage) {
            clique_tail_count(w, loop);
            return;
        }
        // The tail sits at depth n-2 at most (a plan has a last level below it),
        // so depths from GM_MAX_PATTERN - 1 on never instantiate it.
        if constexpr (D + 1 < GM_MAX_PATTERN) {
            if (MODE == GM_MODE_PATTERN && D == w.tail_depth) {
                pattern_tail<D>(w, loop);
                return;
            }
        }
        const uint32_t bound = (MODE == GM_MODE_PATTERN) ? restriction_bound<D>(w) : 0xffffffffu;
        const int ln = lane_id();
        int32_t lo = 0, hi = loop.len;
        uint64_t t_loop0 = 0;
        // Depth-3 slices exist for the pattern rule only; the clique rule
        // keeps its loops free of their bookkeeping.
        constexpr bool kSplit = D == 2 || (D == 3 && MODE == GM_MODE_PATTERN);
        constexpr bool kDeepHost = D == 2 && MODE == GM_MODE_PATTERN;
        if constexpr (kSplit) {
            // A root's depth-2 and depth-3 candidates may be split: this wave
            // walks [slice_lo, slice_hi) at depth 2 and [s3_lo, s3_hi) at
            // depth 3 (a donated slice, or the whole sets).
            const int32_t l0 = D == 2 ? w.slice_lo : w.s3_lo, h0 = D == 2 ? w.slice_hi : w.s3_hi;
            lo = l0 < loop.len ? l0 : loop.len;
            hi = h0 < loop.len ? h0 : loop.len;
            if (D == 2 ? w.donate : w.donate3) t_loop0 = (uint64_t)uni64((int64_t)wall_clock64());
        }
        for (int32_t base = lo; base < hi; base += kWave) {
            if constexpr (kSplit) hi = maybe_donate<D>(w, lo, base, hi, t_loop0, 0);
            int32_t i = base + ln;
            uint32_t cand = 0xffffffffu;
            int64_t rb = 0, re = 0;
            if (i < hi) {
                cand = loop.ptr[i];
                rb = w.a.row[cand];
                re = w.a.row[cand + 1];
            }
            int32_t cnum = hi - base < kWave ? hi - base : kWave;
            for (int j = 0; j < cnum; ++j) {
                uint32_t v = rdlane(cand, j);
                if (v >= bound) return;  // sorted: nothing further qualifies
                if (out_of_time(w)) return;
                uint64_t t_b2 = 0;
                if constexpr (kDeepHost) {
                    // A depth-3 slice rebuilds its one depth-2 candidate's
                    // prefixes, which the donor has charged.
                    w.a.stats = w.deep ? 0 : w.stats0;
                    if (w.donate3) t_b2 = (uint64_t)uni64((int64_t)wall_clock64());
                }
                if (w.a.stats) w.bytes += 4ull;
                if (MODE == GM_MODE_PATTERN && in_emb<D>(w, v)) continue;
                int64_t l = rdlane64(rb, j), r = rdlane64(re, j);
                if (w.a.stats) w.bytes += 8ull;
                const unsigned long long* vbm = hub_bitmap(w.a, v);
                bool prune = false;
                for (int pid = p.last[D]; pid != -1; pid = p.next[pid]) {
                    build(w, pid, w.a.col + l, (int32_t)(r - l), vbm);
                    if (get_set(w, pid).len == p.brk[pid]) {
                        prune = true;
                        break;
                    }
                }
                if constexpr (kDeepHost) {
                    w.a.stats = w.stats0;
                    if (w.donate3) w.rebuild2_ticks = uni64((int64_t)wall_clock64()) - (int64_t)t_b2;
                    w.cur_i2 = base + j;
                }
                if (prune) continue;
                w.emb[D] = v;
                dfs<D + 1, MODE, NMAX>(w);
            }
        }
    }
}

// Largest v with row[v] <= e (uniform trip count).
__device__ __forceinline__ int32_t edge_source(const int64_t* row, int32_t nv, int64_t e) {
    int32_t base = 0, n = nv;  // search in row[0:nv)
    while (n > 1) {
        int32_t half = n >> 1;
        base = (row[base + half] <= e) ? base + half : base;
        n -= half;
    }
    return base;
}

template <int MODE, int NMAX>
// Root edge (v0, v1): build the prefixes of depths 0 and 1 and walk the
// depth-2 candidates [lo, hi).  A donated slice (`donated`) rebuilds the
// root's prefixes without charging them again: the wave that found the root
// charged them.
__device__ __forceinline__ void process_root(Wave& w, uint32_t v0, uint32_t v1, int32_t lo, int32_t hi,
                                             bool donated, int32_t lo3 = 0, int32_t hi3 = kAllCandidates) {
    const DevPlan& p = *w.pl;
    const int32_t stats = w.a.stats;
    if (donated) w.a.stats = 0;
    const uint64_t t_build = w.a.queue ? (uint64_t)uni64((int64_t)wall_clock64()) : 0ull;
    int64_t l0 = w.a.row[v0], r0 = w.a.row[v0 + 1];
    const unsigned long long* bm0 = hub_bitmap(w.a, v0);
    for (int pid = p.last[0]; pid != -1; pid = p.next[pid]) build(w, pid, w.a.col + l0, (int32_t)(r0 - l0), bm0);
    int64_t l1 = w.a.row[v1], r1 = w.a.row[v1 + 1];
    const unsigned long long* bm1 = hub_bitmap(w.a, v1);
    for (int pid = p.last[1]; pid != -1; pid = p.next[pid]) {
        build(w, pid, w.a.col + l1, (int32_t)(r1 - l1), bm1);
        if (get_set(w, pid).len == 0 && pid < p.basic) {
            w.a.stats = stats;
            return;
        }
    }
    w.a.stats = stats;
    if (w.a.queue) w.rebuild_ticks = uni64((int64_t)wall_clock64()) - (int64_t)t_build;
    w.emb[0] = v0;
    w.emb[1] = v1;
    w.slice_lo = lo;
    w.slice_hi = hi;
    w.s3_lo = lo3;
    w.s3_hi = hi3;
    w.deep = (lo3 != 0 || hi3 != kAllCandidates) ? 1 : 0;
    w.donate = (!donated && w.a.queue != nullptr) ? 1 : 0;
    w.donate3 = (w.a.deep_donation && !w.deep && w.a.queue != nullptr) ? 1 : 0;
    dfs<2, MODE, NMAX>(w);
    w.a.stats = w.stats0;
}

// Claims the next pushed slice without waiting (lane 0 CAS on the queue head;
// -1 if none): the claim of a wave that still holds roots, under
// GM_COUNT_EAGER_DONATION.  Idle waves take tickets instead.
__device__ __forceinline__ int64_t pop_slice(const KernelArgs& a, unsigned long long& lost) {
    long long got = -1;
    if (lane_id() == 0) {
        unsigned long long h = __hip_atomic_load(&a.ctr->qhead, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (true) {
            const unsigned long long t = __hip_atomic_load(&a.ctr->qtail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long lim = t < (unsigned long long)a.queue_cap ? t : (unsigned long long)a.queue_cap;
            if (h >= lim) break;
            const unsigned long long prev = atomicCAS(&a.ctr->qhead, h, h + 1);
            if (prev == h) {
                got = (long long)h;
                break;
            }
            h = prev;
            ++lost;
            __builtin_amdgcn_s_sleep(2);  // contended: let the winner through
        }
    }
    return uni64(__shfl(got, 0));
}

// Sum of a per-lane value over the wave (all lanes receive it).
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
        lo = __shfl_xor(lo, off);
        hi = __shfl_xor(hi, off);
        x += ((unsigned long long)hi << 32) | lo;
    }
    return x;
}

// Graph of root edge e: largest g with gedge[g] <= e (uniform trip count).
__device__ __forceinline__ int32_t edge_graph(const int64_t* gedge, int32_t ng, int64_t e) {
    int32_t base = 0, n = ng;
    while (n > 1) {
        int32_t half = n >> 1;
        base = (gedge[base + half] <= e) ? base + half : base;
        n -= half;
    }
    return base;
}

// Adds the wave's partial sum for (graph, plan) to sums[] and clears it.
__device__ __forceinline__ void flush_sum(Wave& w, int32_t g, int32_t pi) {
    const unsigned long long total = w.sum + wave_sum(w.lane_sum);
    if (lane_id() == 0 && total != 0ull) atomicAdd(&w.a.sums[(int64_t)g * w.a.np + pi], total);
    w.sum = 0;
    w.lane_sum = 0;
}

// Persistent enumeration over tasks (root chunk c, plan pi), chunk-major so
// the plans of one chunk run close together and share its neighbour lists
// in cache.  A root's count goes to the (graph, plan) cell of its edge.
// NMAX: the largest pattern size among the launch's plans; the DFS is
// instantiated only that deep, which keeps registers (and so occupancy) to
// what the patterns need.  WPE: waves per SIMD the build must fit.  At 2 the
// pattern-rule DFS keeps all its state in registers (about 230 VGPRs); at 4
// it is held to 128 VGPRs and spills some 400 bytes per lane to scratch,
// which hides more of the latency of its scattered list reads: on R-MAT
// 2^22 house and 4-clique + pendant ran 1.46x and 1.49x faster, the 5-cycle
// (31 roots, bound by its few heaviest slices) 1.22x slower.
template <int MODE, int NMAX, int WPE>
__global__ void __launch_bounds__(kBlock, WPE) enumerate_kernel(KernelArgs a) {
    __shared__ WaveShared<MODE> lds_all[kWavesPerBlock];
    Wave w;
    w.a = a;
    const int wid = threadIdx.x / kWave;
    w.lds = &lds_all[wid].w;
    w.stage = lds_all[wid].stage;
    const int64_t gwave = (int64_t)blockIdx.x * kWavesPerBlock + wid;
    w.scratch = a.scratch + gwave * a.wave_stride;
    w.sum = 0;
    w.bytes = 0;
    w.lane_sum = 0;
    w.lane_bytes = 0;
    w.overflow = 0;
    w.aborted = 0;
    w.stats0 = a.stats;
    w.deep = 0;
    w.donate = 0;
    w.donate3 = 0;
    w.cur_i2 = 0;
    w.s3_lo = 0;
    w.s3_hi = kAllCandidates;
    w.rebuild_ticks = 0;
    w.rebuild2_ticks = 0;
    w.t_start = (uint64_t)uni64((int64_t)wall_clock64());
#pragma unroll
    for (int i = 0; i <= GM_MAX_PATTERN; ++i) w.emb[i] = 0;
    // A set that is never built reads as empty.
    for (int i = lane_id(); i < kMaxSets; i += kWave) {
        w.lds->sets[i].ptr = nullptr;
        w.lds->sets[i].bm = nullptr;
        w.lds->sets[i].len = 0;
    }
    wave_sync();
    const int ln = lane_id();
    // Work items are (logical root, plan) pairs, root-major, so the waves
    // that share a root edge's neighbour lists run its plans close together.
    // Guided self-scheduling: a claim takes remaining / (4 * waves) items,
    // at most task_max and at least 1, so large claims amortise the cursor
    // atomic early and single roots spread over the whole grid at the end
    // (a hub's root edges can each hold seconds of work).
    const int64_t nitems = a.nroots * (int64_t)a.np;
    w.nitems = nitems;
    const int64_t nwaves_total = (int64_t)gridDim.x * kWavesPerBlock;
    unsigned long long roots = 0;
    int32_t cur_g = 0, cur_p = 0;
    bool main_done = false;
    int idle_sleeps = 1;
    int64_t ticket = -1;  // queue slot this idle wave holds (wave-uniform)
    // Queue diagnostics (lane 0's CAS losses; wave-uniform tick sums).
    unsigned long long lost = 0, slice_ticks = 0, idle_ticks = 0;
    w.donations = 0;
    w.deep_donations = 0;
    while (true) {
        if (out_of_time(w)) break;
        // Donated slices first: they belong to roots already under way.
        // Lazy donation only pushes once the cursor is exhausted, so until
        // then the queue is empty and is not polled.
        if (a.queue && (main_done || a.eager_donation)) {
            const uint64_t t_pop = (uint64_t)uni64((int64_t)wall_clock64());
            int64_t slot = -1;
            if (main_done) {
                // An idle wave holds a ticket: the next queue slot, taken
                // with one fetch-add (a CAS claim costs one attempt per
                // contending wave, and thousands of waves idle at the end).
                // It waits until its slot is pushed, or leaves once nothing
                // can push any more and the ticket lies past the tail.  Only
                // roots push unless depth-3 donation is on, so then nothing
                // can once every root item is done; with depth-3 donation
                // depth-2 slices push too, so also every queued slot must
                // have been processed (slices_done == min(qtail, cap): a
                // slot in flight is queued but not done).  Counting items,
                // not waves, keeps the rule independent of how many of the
                // grid's waves are resident.
                if (ticket < 0) {
                    unsigned long long t = 0;
                    if (ln == 0) t = atomicAdd(&a.ctr->qhead, 1ull);
                    ticket = uni64(__shfl((long long)t, 0));
                }
                if (ticket >= a.queue_cap) break;  // the queue is used up: no slot will come
                unsigned long long tail = 0;
                if (ln == 0) tail = __hip_atomic_load(&a.ctr->qtail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                tail = (unsigned long long)uni64(__shfl((long long)tail, 0));
                if ((unsigned long long)ticket < tail) {
                    slot = ticket;
                    ticket = -1;
                } else {
                    unsigned long long fin = 0;
                    if (ln == 0)
                        fin = __hip_atomic_load(&a.ctr->done_items, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    fin = (unsigned long long)uni64(__shfl((long long)fin, 0));
                    if ((int64_t)fin >= nitems && !a.deep_donation) {
                        // Pairs with the release adds of done_items: every
                        // push reserved its slots before its root's add.
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                        if (ln == 0) tail = __hip_atomic_load(&a.ctr->qtail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        tail = (unsigned long long)uni64(__shfl((long long)tail, 0));
                        if ((unsigned long long)ticket >= tail) break;
                        continue;  // pushed meanwhile: the next pass takes it
                    }
                    if ((int64_t)fin >= nitems) {
                        // Every root is done; the tail is final once every
                        // queued slot is done too.  The relaxed reads gate the
                        // one acquire (an L2 invalidate on this XCD, which an
                        // idle wave must not repeat under the waves still
                        // walking slices): only when they say so is the
                        // acquire taken, pairing with the release adds of
                        // done_items and slices_done (every push reserved its
                        // slots before its root's or slice's add), and the
                        // tail read again after it.  slices_done is read
                        // before that tail, so a slice still in flight keeps
                        // the count below the tail read after it.
                        unsigned long long sd = 0;
                        if (ln == 0)
                            sd = __hip_atomic_load(&a.ctr->slices_done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        sd = (unsigned long long)uni64(__shfl((long long)sd, 0));
                        const unsigned long long seen = tail < (unsigned long long)a.queue_cap
                                                            ? tail : (unsigned long long)a.queue_cap;
                        if (sd >= seen) {
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                            if (ln == 0)
                                tail = __hip_atomic_load(&a.ctr->qtail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            tail = (unsigned long long)uni64(__shfl((long long)tail, 0));
                            const unsigned long long queued = tail < (unsigned long long)a.queue_cap
                                                                  ? tail : (unsigned long long)a.queue_cap;
                            if ((unsigned long long)ticket < tail) continue;  // pushed meanwhile: the next pass takes it
                            if (sd >= queued) break;  // the tail is final
                        }
                    }
                    // Back off exponentially; relaxed polls, nothing is read
                    // on their strength (an acquire would invalidate this
                    // XCD's L2 under the waves still walking adjacency lists).
                    for (int k = 0; k < idle_sleeps; ++k) __builtin_amdgcn_s_sleep(127);
                    idle_sleeps = idle_sleeps < 16 ? idle_sleeps * 2 : 16;
                    idle_ticks += uni64((int64_t)wall_clock64()) - (int64_t)t_pop;
                    continue;
                }
            } else {
                slot = pop_slice(a, lost);  // still holds roots: never wait
            }
            if (slot >= 0) {
                SliceItem* it = a.queue + slot;
                // The slot is reserved, so its pusher is writing it: spin
                // with relaxed loads, then acquire once.
                if (ln == 0)
                    while (__hip_atomic_load(&it->ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
                        __builtin_amdgcn_s_sleep(1);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                wave_sync();
                const int64_t e = uni64(__shfl((long long)__hip_atomic_load(&it->e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), 0));
                const int32_t pj = unii(__shfl(it->pl, 0));
                const int32_t gj = unii(__shfl(it->g, 0));
                const int32_t s0 = unii(__shfl(it->s0, 0));
                const int32_t s1 = unii(__shfl(it->s1, 0));
                const int32_t t0 = unii(__shfl(it->t0, 0));
                const int32_t t1 = unii(__shfl(it->t1, 0));
                if (ln == 0) __hip_atomic_store(&it->ready, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (gj != cur_g || pj != cur_p) {
                    flush_sum(w, cur_g, cur_p);
                    cur_g = gj;
                    cur_p = pj;
                }
                w.pl = a.plans[pj];
                w.tail_depth = a.tail_off ? -1 : w.pl->tail_depth;
                const uint32_t u1 = a.col[e];
                const uint32_t u0 = (uint32_t)edge_source(a.row, a.nv, e);
                // A depth-2 slice may donate depth-3 slices of its own root.
                w.cur_e = e;
                w.cur_pl = pj;
                w.cur_gr = gj;
                process_root<MODE, NMAX>(w, u0, u1, s0, s1, true, t0, t1);
                // Release: this slice's own pushes precede it.
                if (a.deep_donation && ln == 0 && !w.aborted)
                    __hip_atomic_fetch_add(&a.ctr->slices_done, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                slice_ticks += uni64((int64_t)wall_clock64()) - (int64_t)t_pop;
                idle_sleeps = 1;
                continue;
            }
        }
        if (main_done) break;
        idle_sleeps = 1;
        unsigned long long start = 0;
        int64_t take = 0;
        if (ln == 0) {
            const unsigned long long seen = __hip_atomic_load(&a.ctr->cursor, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int64_t rem = nitems - (int64_t)seen;
            take = rem / (4 * nwaves_total);
            take = take < 1 ? 1 : (take > a.task_max ? a.task_max : take);
            start = atomicAdd(&a.ctr->cursor, (unsigned long long)take);
        }
        start = (unsigned long long)uni64((int64_t)__shfl((long long)start, 0));
        take = uni64(__shfl((long long)take, 0));
        if ((int64_t)start >= nitems) {
            main_done = true;
            continue;
        }
        const int64_t i_end = (int64_t)start + take < nitems ? (int64_t)start + take : nitems;
        for (int64_t i0 = (int64_t)start; i0 < i_end; i0 += kWave) {
            const int64_t item = i0 + ln;
            bool ok = item < i_end;
            uint32_t v0 = 0, v1 = 0;
            int32_t g = 0, pl = 0;
            int64_t ecur = 0;
            if (ok) {
                const int64_t ri = a.np == 1 ? item : item / a.np;
                pl = a.np == 1 ? 0 : (int32_t)(item - ri * a.np);
                const int64_t e = a.root_begin + ri * a.root_stride;
                ecur = e;
                v1 = a.col[e];
                v0 = (uint32_t)edge_source(a.row, a.nv, e);
                if (a.ng > 1) g = edge_graph(a.gedge, a.ng, e);
                if (a.plans[pl]->rlast[1] != -1 && v0 <= v1) ok = false;
                if (MODE == GM_MODE_CLIQUE && a.root_min_outdeg > 0 && a.row[v0 + 1] - a.row[v0] <= a.root_min_outdeg)
                    ok = false;
            }
            uint64_t m = ballot(ok);
            while (m != 0) {
                int j = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                const uint32_t u0 = rdlane(v0, j), u1 = rdlane(v1, j);
                const int32_t gj = (int32_t)rdlane((uint32_t)g, j);
                const int32_t pj = (int32_t)rdlane((uint32_t)pl, j);
                if (gj != cur_g || pj != cur_p) {
                    flush_sum(w, cur_g, cur_p);
                    cur_g = gj;
                    cur_p = pj;
                }
                w.pl = a.plans[pj];
                w.tail_depth = a.tail_off ? -1 : w.pl->tail_depth;
                ++roots;
                if (a.stats) w.bytes += 24ull;
                w.cur_e = rdlane64(ecur, j);
                w.cur_pl = pj;
                w.cur_gr = gj;
                process_root<MODE, NMAX>(w, u0, u1, 0, 0x7fffffff, false);
            }
        }
        // Release: the pushes (tail reservations) of these roots precede it.
        if (a.queue && ln == 0 && !w.aborted)
            __hip_atomic_fetch_add(&a.ctr->done_items, (unsigned long long)(i_end - (int64_t)start), __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
    }
    flush_sum(w, cur_g, cur_p);
    const unsigned long long lane_bytes = a.stats ? wave_sum(w.lane_bytes) : 0ull;
    if (ln == 0) {
        atomicAdd(&a.ctr->tasks, roots);
        if (a.stats) atomicAdd(&a.ctr->bytes, w.bytes + lane_bytes);
        if (w.overflow) atomicOr(&a.ctr->overflow, 1u);
        if (w.aborted) atomicOr(&a.ctr->timed_out, 1u);
        if (a.queue) {
            if (w.donations) atomicAdd(&a.ctr->donations, (unsigned long long)w.donations);
            if (w.deep_donations) atomicAdd(&a.ctr->deep_donations, (unsigned long long)w.deep_donations);
            if (lost) atomicAdd(&a.ctr->retries, lost);
            if (slice_ticks) atomicAdd(&a.ctr->slice_ticks, slice_ticks);
            if (idle_ticks) atomicAdd(&a.ctr->idle_ticks, idle_ticks);
        }
    }
}



// ---------------------------------------------------------------------------
// Local-bitmap clique engine (GM_MODE_CLIQUE, plain k-clique chain).
//
// For a root v0 of out-degree d <= kBmMaxD every later candidate set is a
// subset of N(v0), so the enumeration below v0 runs on the d-vertex local
// DAG: row i is the bitmap (over positions in N(v0)) of N(N[i]) ∩ N(v0),
// built once per root.  Then a k-clique through v0 is a chain i -> j -> ...
// and every set operation of the schedule is a W-word AND (W = ceil(d/64)):
//   |S_{v0,i,j}|            = popcount(row_i & row_j)
//   |S_{v0,i,j} ∩ N(l)|     = popcount(row_i & row_j & row_l), l in row_i & row_j
// The (i, j) edges of the local DAG are spread evenly over the lanes through
// an exclusive scan of the row popcounts.  This is the local-subgraph bitmap
// idea of the reference's gpu/gpu_kclique.cu:359-400, applied to the
// schedule's clique chain; counts equal the list engine's.
//
// Byte model (HBM, per root): 16 (row pointers) + 4 d (N(v0)) +
// sum over u in N(v0) of (8 + 4 |N(u)|) for building the rows; the counting
// phase reads only the rows (on chip).
// ---------------------------------------------------------------------------
constexpr int kBmMaxD = 1024;      // largest root out-degree handled with bitmaps
constexpr int kBmMaxW = kBmMaxD / 64;
constexpr int kBmLdsWords = 1024;  // rows kept in LDS when d * W <= this

struct BmArgs {
    const int64_t* row;
    const uint32_t* col;
    int32_t nv;
    int32_t k;                 // clique size
    int32_t max_d;             // roots with out-degree above this are left to the list engine
    unsigned long long* rows_scratch;  // per wave kBmMaxD * kBmMaxW words
    unsigned long long* out;   // the plan's sum cell
    Counters* ctr;
    int32_t stats;
    int32_t big_min_d;         // roots with big_min_d < d and d * W <= big_cap_words go to the
    int32_t big_cap_words;     // block-per-root kernel (its rows in LDS); 0: none do
    int32_t small_max_d;       // roots with k-1 <= d <= small_max_d (<= 64) go to the small-root kernel
    int32_t tri;               // 1: every row holds larger ids only, -1: smaller ids only, 0: neither
    const int32_t* order;      // the vertices by out-degree, largest first (ties by id)
    int32_t n_big;             // order[0, n_big): out-degree above 64
    int32_t n_active;          // order[0, n_active): out-degree at least 2
};

// Roots the block-per-root kernel takes (the per-wave kernel skips them).
__device__ __forceinline__ bool bm_is_big(const BmArgs& a, int32_t d) {
    const int32_t W = (d + 63) >> 6;
    return a.big_cap_words > 0 && d > a.big_min_d && d > a.small_max_d && d <= a.max_d && d * W <= a.big_cap_words;
}

struct BmShared {
    uint32_t nbr[kBmMaxD];
    int32_t off[kBmMaxD + 1];
    unsigned long long rows[kBmLdsWords];
};

// Position of the r-th set bit (0-based) of x; x has more than r set bits.
__device__ __forceinline__ int select_bit64(unsigned long long x, int r) {
    int pos = 0;
#pragma unroll
    for (int width = 32; width >= 1; width >>= 1) {
        unsigned long long low = x & ((width == 64 ? ~0ull : ((1ull << width) - 1)));
        int c = __popcll(low);
        if (r >= c) {
            r -= c;
            x >>= width;
            pos += width;
        }
    }
    return pos;
}

template <int WMAX, int K>
__device__ unsigned long long bm_count_pair(const unsigned long long* rows, int W, int i, int j) {
    unsigned long long S[WMAX];
#pragma unroll
    for (int w = 0; w < WMAX; ++w) S[w] = (w < W) ? (rows[i * W + w] & rows[j * W + w]) : 0ull;
    unsigned long long acc = 0;
    if constexpr (K == 4) {
#pragma unroll
        for (int w = 0; w < WMAX; ++w) acc += __popcll(S[w]);
    } else {
#pragma unroll
        for (int w = 0; w < WMAX; ++w) {
            unsigned long long m = S[w];
            while (m) {
                int l = w * 64 + __builtin_ctzll(m);
                m &= m - 1;
                const unsigned long long* rl = rows + (int64_t)l * W;
                if constexpr (K == 5) {
#pragma unroll
                    for (int x = 0; x < WMAX; ++x)
                        if (x < W) acc += __popcll(S[x] & rl[x]);
                } else {
                    unsigned long long T[WMAX];
#pragma unroll
                    for (int x = 0; x < WMAX; ++x) T[x] = (x < W) ? (S[x] & rl[x]) : 0ull;
#pragma unroll
                    for (int x = 0; x < WMAX; ++x) {
                        unsigned long long mm = T[x];
                        while (mm) {
                            int q = x * 64 + __builtin_ctzll(mm);
                            mm &= mm - 1;
                            const unsigned long long* rq = rows + (int64_t)q * W;
                            if constexpr (K == 6) {
#pragma unroll
                                for (int y = 0; y < WMAX; ++y)
                                    if (y < W) acc += __popcll(T[y] & rq[y]);
                            } else {  // K == 7
#pragma unroll
                                for (int y = 0; y < WMAX; ++y) {
                                    if (y >= W) continue;
                                    unsigned long long mq = T[y] & rq[y];
                                    while (mq) {
                                        int z = y * 64 + __builtin_ctzll(mq);
                                        mq &= mq - 1;
                                        const unsigned long long* rz = rows + (int64_t)z * W;
#pragma unroll
                                        for (int t = 0; t < WMAX; ++t)
                                            if (t < W) acc += __popcll(T[t] & rq[t] & rz[t]);
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
    }
    return acc;
}

// bm_count_pair for K = 4, 5 on triangular rows.  When every row of the
// graph holds only larger ids (TRI = 1), the positions of N(v0) are in id
// order, so row i holds positions > i only: S = row_i & row_j lies above j,
// and S & row_l above l.  The word loops start at those bounds (TRI = -1,
// rows below their own position, is the mirror image).  On the degree-rank
// oriented R-MAT 2^20 this cuts the block kernel's 5-clique word operations
// from 1.24e11 to 3.1e10 for the same 1.7e10 (pair, l) items.
template <int WMAX, int K, int TRI>
__device__ __forceinline__ unsigned long long bm_count_pair_tri(const unsigned long long* rows, int W, int i, int j) {
    static_assert(K == 4 || K == 5, "triangular word ranges for K = 4, 5");
    const int lo = TRI > 0 ? ((j + 1) >> 6) : 0;  // first word that can hold a position above j
    const int hi = TRI < 0 ? ((j + 63) >> 6) : W; // words that can hold a position below j
    unsigned long long S[WMAX];
#pragma unroll
    for (int w = 0; w < WMAX; ++w) {
        S[w] = 0ull;
        if (w >= lo && w < hi) S[w] = rows[i * W + w] & rows[j * W + w];
    }
    unsigned long long acc = 0;
    if constexpr (K == 4) {
#pragma unroll
        for (int w = 0; w < WMAX; ++w) acc += __popcll(S[w]);
    } else {
#pragma unroll
        for (int x = 0; x < WMAX; ++x) {
            unsigned long long m = S[x];
            // Two items per step, so that their row loads are in flight together.
            while (m) {
                const int l1 = x * 64 + __builtin_ctzll(m);
                m &= m - 1;
                const unsigned long long* r1 = rows + (int64_t)l1 * W;
                const bool two = m != 0ull;
                const int l2 = two ? x * 64 + __builtin_ctzll(m) : l1;
                if (two) m &= m - 1;
                const unsigned long long* r2 = rows + (int64_t)l2 * W;
                if constexpr (TRI > 0) {
#pragma unroll
                    for (int y = x; y < WMAX; ++y)
                        if (y < W) {
                            const unsigned long long a1 = r1[y], a2 = r2[y];
                            acc += __popcll(S[y] & a1) + (two ? __popcll(S[y] & a2) : 0);
                        }
                } else {
#pragma unroll
                    for (int y = 0; y <= x; ++y) {
                        const unsigned long long a1 = r1[y], a2 = r2[y];
                        acc += __popcll(S[y] & a1) + (two ? __popcll(S[y] & a2) : 0);
                    }
                }
            }
        }
    }
    return acc;
}

template <int WMAX, int K>
__device__ unsigned long long bm_root(const BmArgs& a, BmShared& sh, unsigned long long* gl_rows, uint32_t v0,
                                      int64_t l0, int32_t d, unsigned long long& bytes) {
    const int ln = lane_id();
    const int W = (d + 63) >> 6;
    for (int i = ln; i < d; i += kWave) sh.nbr[i] = a.col[l0 + i];
    unsigned long long* rows = (d * W <= kBmLdsWords) ? sh.rows : gl_rows;
    for (int t = ln; t < d * W; t += kWave) rows[t] = 0ull;
    wave_sync();
    // Row i = positions in N(v0) of N(N[i]); lane-private rows, no conflicts.
    unsigned long long my_bytes = 0;
    for (int i = ln; i < d; i += kWave) {
        uint32_t u = sh.nbr[i];
        int64_t lu = a.row[u], ru = a.row[u + 1];
        int32_t du = (int32_t)(ru - lu);
        my_bytes += 8ull + 4ull * (unsigned long long)du;
        for (int32_t t = 0; t < du; ++t) {
            uint32_t x = a.col[lu + t];
            int32_t base = 0, n = d;
            while (n > 1) {
                int32_t half = n >> 1;
                base = (sh.nbr[base + half] <= x) ? base + half : base;
                n -= half;
            }
            if (sh.nbr[base] == x) rows[i * W + (base >> 6)] |= 1ull << (base & 63);
        }
    }
    bytes += my_bytes;
    wave_sync();
    unsigned long long acc = 0;
    if constexpr (K == 3) {
        for (int t = ln; t < d * W; t += kWave) acc += __popcll(rows[t]);
        return acc;
    } else {
        // Exclusive scan of row popcounts -> off[0..d].
        int32_t carry = 0;
        for (int base = 0; base < d; base += kWave) {
            int i = base + ln;
            int32_t c = 0;
            if (i < d)
                for (int w = 0; w < W; ++w) c += __popcll(rows[i * W + w]);
            int32_t incl = c;
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                int32_t y = __shfl_up(incl, off);
                if (ln >= off) incl += y;
            }
            if (i < d) sh.off[i] = carry + incl - c;
            carry += __shfl(incl, kWave - 1);
        }
        if (ln == 0) sh.off[d] = carry;
        wave_sync();
        const int32_t P = carry;
        for (int32_t pidx = ln; pidx < P; pidx += kWave) {
            // Row i with off[i] <= pidx < off[i+1].
            int32_t lo = 0, n = d;
            while (n > 1) {
                int32_t half = n >> 1;
                lo = (sh.off[lo + half] <= pidx) ? lo + half : lo;
                n -= half;
            }
            const int i = lo;
            int r = pidx - sh.off[i];
            int j = 0;
            for (int w = 0; w < W; ++w) {
                unsigned long long word = rows[i * W + w];
                int c = __popcll(word);
                if (r < c) {
                    j = w * 64 + select_bit64(word, r);
                    break;
                }
                r -= c;
            }
            acc += bm_count_pair<WMAX, K>(rows, W, i, j);
        }
        wave_sync();
        return acc;
    }
}

template <int WMAX, int K>
__global__ void __launch_bounds__(kWave) clique_bitmap_kernel(BmArgs a) {
    __shared__ BmShared sh;
    const int ln = lane_id();
    unsigned long long* gl_rows = a.rows_scratch + (int64_t)blockIdx.x * kBmMaxD * kBmMaxW;
    unsigned long long acc = 0, bytes = 0, wave_bytes = 0;
    unsigned long long roots = 0;
    // Roots of out-degree above 64, heaviest first, one per claim.
    while (true) {
        unsigned long long idx = 0;
        if (ln == 0) idx = atomicAdd(&a.ctr->bm_cursor, 1ull);
        idx = (unsigned long long)uni64((int64_t)__shfl((long long)idx, 0));
        if ((long long)idx >= a.n_big) break;
        {
            const int32_t v0 = unii(a.order[idx]);
            int64_t l0 = a.row[v0], r0 = a.row[v0 + 1];
            int32_t d = unii((int32_t)(r0 - l0));
            if (d < K - 1 || d > a.max_d || d <= a.small_max_d || bm_is_big(a, d)) continue;
            ++roots;
            if (a.stats) wave_bytes += 16ull + 4ull * (unsigned long long)d;
            if (d <= 64) acc += bm_root<1, K>(a, sh, gl_rows, (uint32_t)v0, l0, d, bytes);
            else if (d <= 256) acc += bm_root<4, K>(a, sh, gl_rows, (uint32_t)v0, l0, d, bytes);
            else acc += bm_root<WMAX, K>(a, sh, gl_rows, (uint32_t)v0, l0, d, bytes);
        }
    }
    const unsigned long long total = wave_sum(acc);
    const unsigned long long tb = a.stats ? wave_sum(bytes) : 0ull;
    if (ln == 0) {
        atomicAdd(a.out, total);
        atomicAdd(&a.ctr->tasks, roots);
        if (a.stats) atomicAdd(&a.ctr->bytes, tb + wave_bytes);
    }
}

// ---------------------------------------------------------------------------
// Block-per-root bitmap engine for the large roots.
//
// On skewed graphs the roots of out-degree a few hundred hold most of the
// k-clique work (R-MAT 2^20, k = 5: 0.7% of the roots, ~90% of the word
// operations), and one wave per root with its rows in global memory leaves
// the grid waiting on them.  Here a block of kBigThreads takes one root at a
// time: N(v0) and the d x W rows live in LDS (dynamic, sized at launch),
// each wave builds rows with its lanes over N(u) (LDS atomic OR), and the
// local edges (i, j) are handed out one per thread from an LDS counter, so a
// few heavy pairs cannot hold a wave.  Same arithmetic and byte charge as
// bm_root.
// ---------------------------------------------------------------------------
constexpr int kBigThreads = 512;

// Pair counting of the block kernel, sized to the root's row width.
template <int WMAX, int K, int TRI>
__device__ __forceinline__ unsigned long long big_pair(const unsigned long long* rows, int W, int i, int j) {
    if constexpr ((K == 4 || K == 5) && TRI != 0) return bm_count_pair_tri<WMAX, K, TRI>(rows, W, i, j);
    else return bm_count_pair<WMAX, K>(rows, W, i, j);
}

template <int K, int TRI>
__global__ void __launch_bounds__(kBigThreads) clique_bitmap_big_kernel(BmArgs a) {
    extern __shared__ unsigned long long big_rows[];
    __shared__ uint32_t nbr[kBmMaxD];
    __shared__ int32_t off[kBmMaxD + 1];
    __shared__ int32_t pair_ctr;
    __shared__ long long chunk_sh;
    __shared__ unsigned long long red[kBigThreads / kWave][2];
    const int tid = threadIdx.x, ln = tid & (kWave - 1), wv = tid / kWave;
    constexpr int nwv = kBigThreads / kWave;
    unsigned long long acc = 0, bytes = 0, roots = 0;
    // Roots of out-degree above 64, heaviest first, one per claim: the
    // last claims are the lightest roots, so the blocks finish together.
    while (true) {
        if (tid == 0) chunk_sh = (long long)atomicAdd(&a.ctr->big_cursor, 1ull);
        __syncthreads();
        const long long idx = chunk_sh;
        __syncthreads();
        if (idx >= a.n_big) break;
        {
            const int32_t v0 = a.order[idx];
            const int64_t l0 = a.row[v0], r0 = a.row[v0 + 1];
            const int32_t d = (int32_t)(r0 - l0);
            if (d < K - 1 || !bm_is_big(a, d)) continue;
            const int W = (d + 63) >> 6;
            if (tid == 0) {
                ++roots;
                if (a.stats) bytes += 16ull + 4ull * (unsigned long long)d;
            }
            for (int i = tid; i < d; i += kBigThreads) nbr[i] = a.col[l0 + i];
            for (int t = tid; t < d * W; t += kBigThreads) big_rows[t] = 0ull;
            __syncthreads();
            // Rows: wave wv builds rows wv, wv + nwv, ...; its lanes split N(u).
            for (int i = wv; i < d; i += nwv) {
                const uint32_t u = nbr[i];
                const int64_t lu = a.row[u], ru = a.row[u + 1];
                const int32_t du = (int32_t)(ru - lu);
                if (ln == 0 && a.stats) bytes += 8ull + 4ull * (unsigned long long)du;
                for (int32_t t = ln; t < du; t += kWave) {
                    const uint32_t x = a.col[lu + t];
                    int32_t base = 0, n = d;
                    while (n > 1) {
                        const int32_t half = n >> 1;
                        base = (nbr[base + half] <= x) ? base + half : base;
                        n -= half;
                    }
                    if (nbr[base] == x) atomicOr(&big_rows[i * W + (base >> 6)], 1ull << (base & 63));
                }
            }
            __syncthreads();
            if constexpr (K == 3) {
                for (int t = tid; t < d * W; t += kBigThreads) acc += __popcll(big_rows[t]);
                __syncthreads();
                continue;
            } else {
                // Row popcounts, then an exclusive scan by wave 0.
                for (int i = tid; i < d; i += kBigThreads) {
                    int32_t c = 0;
                    for (int w = 0; w < W; ++w) c += __popcll(big_rows[i * W + w]);
                    off[i] = c;
                }
                __syncthreads();
                if (wv == 0) {
                    int32_t carry = 0;
                    for (int base = 0; base < d; base += kWave) {
                        const int i = base + ln;
                        const int32_t c = i < d ? off[i] : 0;
                        int32_t incl = c;
#pragma unroll
                        for (int o = 1; o < kWave; o <<= 1) {
                            const int32_t y = __shfl_up(incl, o);
                            if (ln >= o) incl += y;
                        }
                        if (i < d) off[i] = carry + incl - c;
                        carry += __shfl(incl, kWave - 1);
                    }
                    if (ln == 0) {
                        off[d] = carry;
                        pair_ctr = 0;
                    }
                }
                __syncthreads();
                const int32_t P = off[d];
                while (true) {
                    // A wave claims 64 consecutive pairs with one LDS atomic.
                    int32_t pbase = 0;
                    if (ln == 0) pbase = atomicAdd(&pair_ctr, kWave);
                    pbase = __shfl(pbase, 0);
                    if (pbase >= P) break;
                    const int32_t pidx = pbase + ln;
                    if (pidx >= P) continue;  // the next claim ends the loop
                    int32_t lo = 0, n = d;
                    while (n > 1) {
                        const int32_t half = n >> 1;
                        lo = (off[lo + half] <= pidx) ? lo + half : lo;
                        n -= half;
                    }
                    const int i = lo;
                    int r = pidx - off[i];
                    int j = 0;
                    for (int w = 0; w < W; ++w) {
                        const unsigned long long word = big_rows[i * W + w];
                        const int c = __popcll(word);
                        if (r < c) {
                            j = w * 64 + select_bit64(word, r);
                            break;
                        }
                        r -= c;
                    }
                    // Word loops sized to the root's row width (W <= 16).
                    if (W <= 4) acc += big_pair<4, K, TRI>(big_rows, W, i, j);
                    else if (W <= 8) acc += big_pair<8, K, TRI>(big_rows, W, i, j);
                    else if (W <= 12) acc += big_pair<12, K, TRI>(big_rows, W, i, j);
                    else acc += big_pair<kBmMaxW, K, TRI>(big_rows, W, i, j);
                }
                __syncthreads();
            }
        }
    }
    const unsigned long long wa = wave_sum(acc), wb = a.stats ? wave_sum(bytes) : 0ull;
    if (ln == 0) {
        red[wv][0] = wa;
        red[wv][1] = wb;
    }
    __syncthreads();
    if (tid == 0) {
        unsigned long long ta = 0, tb = 0;
        for (int k = 0; k < nwv; ++k) {
            ta += red[k][0];
            tb += red[k][1];
        }
        atomicAdd(a.out, ta);
        atomicAdd(&a.ctr->tasks, roots);
        if (a.stats) atomicAdd(&a.ctr->bytes, tb);
    }
}

// ---------------------------------------------------------------------------
// Small roots: out-degree k-1 <= d <= 64, one word per row.
//
// One wave per root.  Rows are built from every N(u_p), u_p in N(v0), at
// once: the out-degrees are scanned over the lanes, and lane t takes element
// t of the concatenated lists (consecutive lanes read consecutive elements,
// so the loads coalesce), finds its list p and its position in N(v0) by
// binary searches in LDS and sets that bit of row p with an LDS atomic OR.
// Then lane p holds row p in a register and the counting loops run over
// wave-uniform pairs, each lane testing its own position:
//   k = 4: sum over i, j in row_i of |row_i & row_j|
//   k = 5: sum over i, j in row_i, l in S = row_i & row_j of |S & row_l|
//   (k = 6, 7: one and two more uniform levels),
// the same sums bm_root forms.  Byte model as bm_root.
// ---------------------------------------------------------------------------
constexpr int kSmallD = 64;
constexpr int kSmallChunk = 32;  // vertices per cursor claim

template <int K>
__device__ __forceinline__ unsigned long long small_root_count(unsigned long long r, int d) {
    const int ln = lane_id();
    unsigned long long acc = 0;
    if constexpr (K == 3) {
        acc = (unsigned long long)__popcll(r);
    } else {
        for (int i = 0; i < d; ++i) {
            const unsigned long long ri = (unsigned long long)rdlane64((int64_t)r, i);
            if (ri == 0ull) continue;
            if constexpr (K == 4) {
                if ((ri >> ln) & 1ull) acc += (unsigned long long)__popcll(ri & r);
            } else {
                unsigned long long m = ri;
                while (m) {
                    const int j = __builtin_ctzll(m);
                    m &= m - 1;
                    const unsigned long long S = ri & (unsigned long long)rdlane64((int64_t)r, j);
                    if constexpr (K == 5) {
                        if ((S >> ln) & 1ull) acc += (unsigned long long)__popcll(S & r);
                    } else {
                        unsigned long long m2 = S;
                        while (m2) {
                            const int l = __builtin_ctzll(m2);
                            m2 &= m2 - 1;
                            const unsigned long long S2 = S & (unsigned long long)rdlane64((int64_t)r, l);
                            if constexpr (K == 6) {
                                if ((S2 >> ln) & 1ull) acc += (unsigned long long)__popcll(S2 & r);
                            } else {
                                unsigned long long m3 = S2;
                                while (m3) {
                                    const int q = __builtin_ctzll(m3);
                                    m3 &= m3 - 1;
                                    const unsigned long long S3 = S2 & (unsigned long long)rdlane64((int64_t)r, q);
                                    if ((S3 >> ln) & 1ull) acc += (unsigned long long)__popcll(S3 & r);
                                }
                            }
                        }
                    }
                }
            }
        }
    }
    return acc;
}

template <int K>
__global__ void __launch_bounds__(kWave) clique_bitmap_small_kernel(BmArgs a) {
    __shared__ uint32_t nbr[kSmallD];
    __shared__ int32_t off[kSmallD];
    __shared__ int64_t beg[kSmallD];
    __shared__ unsigned long long rows[kSmallD];
    const int ln = lane_id();
    unsigned long long acc = 0, bytes = 0, roots = 0;
    // Chunks of the roots of out-degree 2..64 in decreasing out-degree:
    // each claim holds roots of one size, the heavy ones first.
    const int64_t nsmall = (int64_t)a.n_active - a.n_big;
    const int64_t nchunks = (nsmall + kSmallChunk - 1) / kSmallChunk;
    while (true) {
        unsigned long long chunk = 0;
        if (ln == 0) chunk = atomicAdd(&a.ctr->sm_cursor, 1ull);
        chunk = (unsigned long long)uni64((int64_t)__shfl((long long)chunk, 0));
        if ((int64_t)chunk >= nchunks) break;
        const int64_t ibeg = a.n_big + (int64_t)chunk * kSmallChunk;
        const int64_t iend = ibeg + kSmallChunk < a.n_active ? ibeg + kSmallChunk : a.n_active;
        for (int64_t oi = ibeg; oi < iend; ++oi) {
            const int32_t v0 = unii(a.order[oi]);
            const int64_t l0 = a.row[v0];
            const int32_t d = unii((int32_t)(a.row[v0 + 1] - l0));
            if (d < K - 1 || d > a.small_max_d) continue;
            ++roots;
            int32_t du = 0;
            if (ln < d) {
                const uint32_t u = a.col[l0 + ln];
                const int64_t lu = a.row[u];
                du = (int32_t)(a.row[u + 1] - lu);
                nbr[ln] = u;
                beg[ln] = lu;
                if (a.stats) bytes += 8ull + 4ull * (unsigned long long)du;
            }
            if (a.stats && ln == 0) bytes += 16ull + 4ull * (unsigned long long)d;
            int32_t incl = du;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int32_t y = __shfl_up(incl, o);
                if (ln >= o) incl += y;
            }
            const int32_t total = unii(__shfl(incl, kWave - 1));
            if (ln < d) off[ln] = incl - du;
            rows[ln] = 0ull;
            wave_sync();
            // Four elements per lane per step, their loads in flight together.
            constexpr int kIlp = 4;
            for (int32_t t0 = 0; t0 < total; t0 += kIlp * kWave) {
                uint32_t xs[kIlp];
                int32_t ps[kIlp];
#pragma unroll
                for (int k = 0; k < kIlp; ++k) {
                    const int32_t t = t0 + k * kWave + ln;
                    ps[k] = -1;
                    xs[k] = 0;
                    if (t < total) {
                        int32_t p = 0, n = d;  // the list holding element t: largest p with off[p] <= t
                        while (n > 1) {
                            const int32_t half = n >> 1;
                            p = (off[p + half] <= t) ? p + half : p;
                            n -= half;
                        }
                        ps[k] = p;
                        xs[k] = a.col[beg[p] + (t - off[p])];
                    }
                }
#pragma unroll
                for (int k = 0; k < kIlp; ++k) {
                    if (ps[k] < 0) continue;
                    const uint32_t x = xs[k];
                    int32_t q = 0, n = d;  // position of x in N(v0)
                    while (n > 1) {
                        const int32_t half = n >> 1;
                        q = (nbr[q + half] <= x) ? q + half : q;
                        n -= half;
                    }
                    if (nbr[q] == x) atomicOr(&rows[ps[k]], 1ull << q);
                }
            }
            wave_sync();
            const unsigned long long r = ln < d ? rows[ln] : 0ull;
            acc += small_root_count<K>(r, d);
            wave_sync();  // the next root rewrites the LDS arrays
        }
    }
    const unsigned long long total = wave_sum(acc);
    const unsigned long long tb = a.stats ? wave_sum(bytes) : 0ull;
    if (ln == 0) {
        atomicAdd(a.out, total);
        atomicAdd(&a.ctr->tasks, roots);
        if (a.stats) atomicAdd(&a.ctr->bytes, tb);
    }
}

// One instantiation of enumerate_kernel<MODE, NMAX, WPE> per object file
// (enum_instance.hip); the host half dispatches on (mode, nmax, wpe).  The
// pattern rule is built for 2 and 4 waves per SIMD, the clique rule for 2.
#define GMK_ENUM_DECL(M, N, W)                                                           \
    hipError_t enum_occupancy_##M##_##N##_##W(int* blocks_per_cu);                       \
    void enum_launch_##M##_##N##_##W(int blocks, hipStream_t s, const KernelArgs& a);
GMK_ENUM_DECL(0, 4, 2)
GMK_ENUM_DECL(0, 5, 2)
GMK_ENUM_DECL(0, 6, 2)
GMK_ENUM_DECL(0, 7, 2)
GMK_ENUM_DECL(0, 4, 4)
GMK_ENUM_DECL(0, 5, 4)
GMK_ENUM_DECL(0, 6, 4)
GMK_ENUM_DECL(0, 7, 4)
GMK_ENUM_DECL(1, 4, 2)
GMK_ENUM_DECL(1, 5, 2)
GMK_ENUM_DECL(1, 6, 2)
GMK_ENUM_DECL(1, 7, 2)
#undef GMK_ENUM_DECL

}  // namespace gmk
