// gm_engine.hip -- schedule-driven subgraph enumeration on gfx950 (wave64).
//
// One wavefront is one DFS worker.  The grid is persistent: each wave claims
// a chunk of 64 root edges with one atomic, filters it lane-parallel (the
// depth-1 restriction v1 < v0) and walks the survivors one by one.  Per
// depth, the candidate loop set is read 64 elements at a time (one per
// lane) together with their row pointers, and the wave steps through them
// with v_readlane, so the DFS control flow is wave-uniform.
//
// Set operations (gpu/component/gpu_vertex_set.cuh:67-248 in the reference):
//   intersection   lane i takes element i of the shorter list and does a
//                  branchless binary search in the longer one (uniform trip
//                  count); hits are compacted in order with ballot + mbcnt.
//   size only      the same without the writes.
//   subtraction    |S \ emb|: lane j searches emb[j] in S, one ballot.
// Counting rules follow gpu/src/gpu_pattern_matching.cuh:14-104 and
// gpu/function/pattern_matching.cuh:110-288 (GM_MODE_PATTERN) and
// gpu/gpu_clique.cu:215-364 (GM_MODE_CLIQUE).  Sums are uint64 and wrap like
// the reference's unsigned long long.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "graphindor_amd.h"

namespace {

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
    int32_t iep_id[GM_MAX_IEP], iep_vflag[GM_MAX_IEP], iep_vcoef[GM_MAX_IEP];
    int32_t iep_coef[GM_MAX_IEP], iep_flag[GM_MAX_IEP], iep_pos[GM_MAX_IEP];
};

struct SetRef {
    const uint32_t* ptr;
    int32_t len;
    int32_t pad;
};

struct WaveLds {
    SetRef sets[kMaxSets];
    int32_t ans[GM_MAX_IEP];
};

struct Counters {
    unsigned long long sum;
    unsigned long long cursor;
    unsigned long long bytes;
    unsigned long long tasks;
    unsigned int overflow;
    unsigned int pad;
};

struct KernelArgs {
    const int64_t* row;
    const uint32_t* col;
    int32_t nv;
    int64_t nnz;
    const DevPlan* plan;
    uint32_t* scratch;
    int64_t wave_stride;  // uint32 per wave
    int32_t cap;          // capacity of a prefix slot
    int32_t cap_vi;       // capacity of a vertex-induced slot
    int64_t nchunks;
    Counters* ctr;
    int32_t stats;
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

struct Wave {
    KernelArgs a;
    const DevPlan* pl;
    WaveLds* lds;
    uint32_t* scratch;
    uint32_t emb[GM_MAX_PATTERN + 1];
    unsigned long long sum;
    unsigned long long bytes;
    unsigned int overflow;
};

__device__ __forceinline__ SetRef get_set(const Wave& w, int pid) {
    SetRef s = w.lds->sets[pid];
    SetRef r;
    r.ptr = (const uint32_t*)uni64((int64_t)s.ptr);
    r.len = unii(s.len);
    r.pad = 0;
    return r;
}

__device__ __forceinline__ void put_set(Wave& w, int pid, const uint32_t* ptr, int32_t len) {
    if (lane_id() == 0) {
        w.lds->sets[pid].ptr = ptr;
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
__device__ int32_t intersect(Wave& w, const uint32_t* a, int32_t na, const uint32_t* b, int32_t nb, uint32_t* out,
                             int32_t cap) {
    if (na > nb) {
        const uint32_t* t = a;
        a = b;
        b = t;
        int32_t tn = na;
        na = nb;
        nb = tn;
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
            hit = contains(b, nb, x);
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

__device__ __forceinline__ void build(Wave& w, int pid, const uint32_t* in, int32_t nin) {
    const DevPlan& p = *w.pl;
    int f = p.father[pid];
    if (f < 0) {
        put_set(w, pid, in, nin);
        return;
    }
    SetRef fs = get_set(w, f);
    if (p.size_only[pid]) {
        int32_t c = intersect(w, fs.ptr, fs.len, in, nin, nullptr, 0);
        put_set(w, pid, nullptr, c);
        if (w.a.stats) w.bytes += 4ull * (unsigned long long)(fs.len + nin);
        return;
    }
    uint32_t* out = w.scratch + (int64_t)pid * w.a.cap;
    int32_t c = intersect(w, fs.ptr, fs.len, in, nin, out, w.a.cap);
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
__device__ SetRef remove_anti_edge(Wave& w, SetRef in) {
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
    r.len = cnt < w.a.cap_vi ? cnt : w.a.cap_vi;
    r.pad = 0;
    return r;
}

// IEP terminal (gpu/function/pattern_matching.cuh:110-139).
template <int D>
__device__ void iep_terminal(Wave& w) {
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

template <int D, int MODE>
__device__ void dfs(Wave& w) {
    if constexpr (D >= GM_MAX_PATTERN) {
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
        const uint32_t bound = (MODE == GM_MODE_PATTERN) ? restriction_bound<D>(w) : 0xffffffffu;
        const int ln = lane_id();
        for (int32_t base = 0; base < loop.len; base += kWave) {
            int32_t i = base + ln;
            uint32_t cand = 0xffffffffu;
            int64_t rb = 0, re = 0;
            if (i < loop.len) {
                cand = loop.ptr[i];
                rb = w.a.row[cand];
                re = w.a.row[cand + 1];
            }
            int32_t cnum = loop.len - base < kWave ? loop.len - base : kWave;
            for (int j = 0; j < cnum; ++j) {
                uint32_t v = rdlane(cand, j);
                if (v >= bound) return;  // sorted: nothing further qualifies
                if (w.a.stats) w.bytes += 4ull;
                if (MODE == GM_MODE_PATTERN && in_emb<D>(w, v)) continue;
                int64_t l = rdlane64(rb, j), r = rdlane64(re, j);
                if (w.a.stats) w.bytes += 8ull;
                bool prune = false;
                for (int pid = p.last[D]; pid != -1; pid = p.next[pid]) {
                    build(w, pid, w.a.col + l, (int32_t)(r - l));
                    if (get_set(w, pid).len == p.brk[pid]) {
                        prune = true;
                        break;
                    }
                }
                if (prune) continue;
                w.emb[D] = v;
                dfs<D + 1, MODE>(w);
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

template <int MODE>
__device__ void process_root(Wave& w, uint32_t v0, uint32_t v1) {
    const DevPlan& p = *w.pl;
    int64_t l0 = w.a.row[v0], r0 = w.a.row[v0 + 1];
    for (int pid = p.last[0]; pid != -1; pid = p.next[pid]) build(w, pid, w.a.col + l0, (int32_t)(r0 - l0));
    int64_t l1 = w.a.row[v1], r1 = w.a.row[v1 + 1];
    for (int pid = p.last[1]; pid != -1; pid = p.next[pid]) {
        build(w, pid, w.a.col + l1, (int32_t)(r1 - l1));
        if (get_set(w, pid).len == 0 && pid < p.basic) return;
    }
    w.emb[0] = v0;
    w.emb[1] = v1;
    dfs<2, MODE>(w);
}

template <int MODE>
__global__ void __launch_bounds__(kBlock) enumerate_kernel(KernelArgs a) {
    __shared__ WaveLds lds_all[kWavesPerBlock];
    Wave w;
    w.a = a;
    w.pl = a.plan;
    const int wid = threadIdx.x / kWave;
    w.lds = &lds_all[wid];
    const int64_t gwave = (int64_t)blockIdx.x * kWavesPerBlock + wid;
    w.scratch = a.scratch + gwave * a.wave_stride;
    w.sum = 0;
    w.bytes = 0;
    w.overflow = 0;
#pragma unroll
    for (int i = 0; i <= GM_MAX_PATTERN; ++i) w.emb[i] = 0;
    // A set that is never built reads as empty.
    for (int i = lane_id(); i < kMaxSets; i += kWave) {
        w.lds->sets[i].ptr = nullptr;
        w.lds->sets[i].len = 0;
    }
    wave_sync();
    const DevPlan& p = *w.pl;
    const bool restricted1 = p.rlast[1] != -1;
    const int ln = lane_id();
    unsigned long long tasks = 0;
    while (true) {
        unsigned long long chunk = 0;
        if (ln == 0) chunk = atomicAdd(&a.ctr->cursor, 1ull);
        chunk = (unsigned long long)uni64((int64_t)__shfl((long long)chunk, 0));
        if ((long long)chunk >= a.nchunks) break;
        int64_t e = (int64_t)chunk * kWave + ln;
        bool ok = e < a.nnz;
        uint32_t v0 = 0, v1 = 0;
        if (ok) {
            v1 = a.col[e];
            v0 = (uint32_t)edge_source(a.row, a.nv, e);
            if (restricted1 && v0 <= v1) ok = false;
        }
        uint64_t m = ballot(ok);
        while (m != 0) {
            int j = __ffsll((unsigned long long)m) - 1;
            m &= m - 1;
            uint32_t u0 = rdlane(v0, j), u1 = rdlane(v1, j);
            ++tasks;
            if (a.stats) w.bytes += 24ull;
            process_root<MODE>(w, u0, u1);
        }
    }
    if (ln == 0) {
        atomicAdd(&a.ctr->sum, w.sum);
        atomicAdd(&a.ctr->tasks, tasks);
        if (a.stats) atomicAdd(&a.ctr->bytes, w.bytes);
        if (w.overflow) atomicOr(&a.ctr->overflow, 1u);
    }
}

// ---------------------------------------------------------------------------
// Host side of the C ABI.
// ---------------------------------------------------------------------------

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(e_ == hipErrorOutOfMemory ? GM_ERR_OOM : GM_ERR_HIP,                       \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                        \
    } while (0)

// Grow-only per-device scratch shared by calls (one host thread at a time
// per device, as documented for the handles).
struct DeviceCache {
    std::mutex mu;
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    Counters* ctr = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

DeviceCache& cache_for(int dev) {
    static std::mutex mu;
    static std::vector<DeviceCache*> caches;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)caches.size() <= dev) caches.resize(dev + 1, nullptr);
    if (!caches[dev]) caches[dev] = new DeviceCache();
    return *caches[dev];
}

}  // namespace

struct gm_graph {
    int device;
    int32_t n;
    int64_t nnz;
    int64_t* d_row;
    uint32_t* d_col;
    int32_t max_deg;
    int32_t second_deg;
};

struct gm_plan {
    DevPlan host;
    DevPlan* d_plan;
    int device;
};

extern "C" {

const char* gm_last_error(void) { return g_err.c_str(); }

int gm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int gm_graph_create(const int64_t* row_ptr, const int32_t* col, int32_t n, int64_t nnz, int32_t max_isect,
                    gm_graph** out) {
    (void)max_isect;
    if (!out || !row_ptr || (nnz > 0 && !col) || n < 0 || nnz < 0) return fail(GM_ERR_INVALID, "gm_graph_create: bad arguments");
    *out = nullptr;
    if (n > 0x7ffffffe) return fail(GM_ERR_INVALID, "gm_graph_create: too many vertices");
    if (row_ptr[0] != 0 || row_ptr[n] != nnz) return fail(GM_ERR_INVALID, "gm_graph_create: row_ptr must run from 0 to nnz");
    // Validate: monotone rows, ids in range, strictly increasing within a row
    // (sorted, no duplicates) and no self-loops.  Kernel indexing relies on it.
    int64_t d1 = 0, d2 = 0;
    for (int32_t v = 0; v < n; ++v) {
        int64_t l = row_ptr[v], r = row_ptr[v + 1];
        if (r < l) return fail(GM_ERR_INVALID, "gm_graph_create: row_ptr not monotone at row " + std::to_string(v));
        for (int64_t e = l; e < r; ++e) {
            int32_t u = col[e];
            if (u < 0 || u >= n) return fail(GM_ERR_INVALID, "gm_graph_create: neighbour id out of range in row " + std::to_string(v));
            if (u == v) return fail(GM_ERR_INVALID, "gm_graph_create: self-loop at vertex " + std::to_string(v));
            if (e > l && col[e - 1] >= u) return fail(GM_ERR_INVALID, "gm_graph_create: row " + std::to_string(v) + " not strictly increasing");
        }
        int64_t d = r - l;
        if (d > d1) { d2 = d1; d1 = d; }
        else if (d > d2) d2 = d;
    }
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    gm_graph* g = new gm_graph();
    g->device = dev;
    g->n = n;
    g->nnz = nnz;
    g->max_deg = (int32_t)d1;
    g->second_deg = (int32_t)d2;
    g->d_row = nullptr;
    g->d_col = nullptr;
    hipError_t e1 = hipMalloc(&g->d_row, sizeof(int64_t) * ((size_t)n + 1));
    hipError_t e2 = hipMalloc(&g->d_col, sizeof(uint32_t) * (size_t)(nnz > 0 ? nnz : 1));
    if (e1 != hipSuccess || e2 != hipSuccess) {
        if (g->d_row) (void)hipFree(g->d_row);
        if (g->d_col) (void)hipFree(g->d_col);
        delete g;
        return fail(GM_ERR_OOM, "gm_graph_create: device allocation failed");
    }
    hipError_t c1 = hipMemcpy(g->d_row, row_ptr, sizeof(int64_t) * ((size_t)n + 1), hipMemcpyHostToDevice);
    hipError_t c2 = nnz > 0 ? hipMemcpy(g->d_col, col, sizeof(uint32_t) * (size_t)nnz, hipMemcpyHostToDevice) : hipSuccess;
    if (c1 != hipSuccess || c2 != hipSuccess) {
        (void)hipFree(g->d_row);
        (void)hipFree(g->d_col);
        delete g;
        return fail(GM_ERR_HIP, "gm_graph_create: upload failed");
    }
    *out = g;
    return GM_OK;
}

void gm_graph_destroy(gm_graph* g) {
    if (!g) return;
    (void)hipFree(g->d_row);
    (void)hipFree(g->d_col);
    delete g;
}

int gm_plan_create(const gm_plan_desc* d, gm_plan** out) {
    if (!d || !out) return fail(GM_ERR_INVALID, "gm_plan_create: null argument");
    *out = nullptr;
    const int n = d->size;
    if (n < 3 || n > GM_MAX_PATTERN) return fail(GM_ERR_UNSUPPORTED, "gm_plan_create: pattern size must be 3.." + std::to_string(GM_MAX_PATTERN));
    const int P = d->total_prefix_num;
    const int Pmax = n * (n - 1) / 2 + 1;
    if (P < 1 || P > Pmax || d->basic_prefix_num < 0 || d->basic_prefix_num > P)
        return fail(GM_ERR_INVALID, "gm_plan_create: bad prefix counts");
    const int k = d->in_exclusion_optimize_num;
    if (k < 0 || k > n - 2 || k == 1) return fail(GM_ERR_INVALID, "gm_plan_create: bad IEP depth");
    if (d->iep_vertex_num < 0 || d->iep_vertex_num > GM_MAX_IEP || d->iep_array_size < 0 || d->iep_array_size > GM_MAX_IEP)
        return fail(GM_ERR_INVALID, "gm_plan_create: IEP tables too large");
    if (k > 0 && (d->iep_vertex_num == 0 || d->iep_array_size == 0)) return fail(GM_ERR_INVALID, "gm_plan_create: IEP depth without tables");
    if (d->mode != GM_MODE_PATTERN && d->mode != GM_MODE_CLIQUE) return fail(GM_ERR_INVALID, "gm_plan_create: unknown mode");
    // gpu_clique ignores the vertex-induced flag its two-argument schedule
    // carries (gpu/gpu_clique.cu:215-258); it takes no IEP tail.
    if (d->mode == GM_MODE_CLIQUE && k != 0) return fail(GM_ERR_INVALID, "gm_plan_create: clique mode takes no IEP");
    DevPlan h;
    std::memset(&h, 0, sizeof(h));
    h.n = n;
    h.P = P;
    h.basic = d->basic_prefix_num;
    h.nrestrict = d->total_restrict_num;
    h.k = k;
    h.vinduced = (d->is_vertex_induced && d->mode == GM_MODE_PATTERN) ? 1 : 0;
    h.mode = d->mode;
    h.iep_vn = k > 0 ? d->iep_vertex_num : 0;
    h.iep_an = k > 0 ? d->iep_array_size : 0;
    for (int i = 0; i < n * n; ++i) h.adj[i] = d->adj_mat[i] ? 1 : 0;
    auto okp = [&](int x) { return x >= -1 && x < P; };
    for (int i = 0; i < P; ++i) {
        h.father[i] = d->father_prefix_id[i];
        h.next[i] = d->next[i];
        h.brk[i] = d->break_size[i];
        if (!okp(h.father[i]) || !okp(h.next[i]) || h.father[i] >= i) return fail(GM_ERR_INVALID, "gm_plan_create: bad prefix link");
    }
    for (int i = 0; i < n; ++i) {
        h.last[i] = d->last[i];
        h.loopset[i] = d->loop_set_prefix_id[i];
        h.rlast[i] = d->restrict_last[i];
        if (!okp(h.last[i]) || !okp(h.loopset[i])) return fail(GM_ERR_INVALID, "gm_plan_create: bad prefix id");
        if (i >= 1 && h.loopset[i] < 0) return fail(GM_ERR_INVALID, "gm_plan_create: missing loop set");
        if (h.rlast[i] < -1 || h.rlast[i] >= Pmax) return fail(GM_ERR_INVALID, "gm_plan_create: bad restriction list");
    }
    // Restriction lists: indices in range, each entry refers to an earlier vertex.
    for (int i = 0; i < Pmax; ++i) {
        h.rnext[i] = d->restrict_next[i];
        h.rindex[i] = d->restrict_index[i];
    }
    for (int v = 0; v < n; ++v) {
        int steps = 0;
        for (int r = h.rlast[v]; r != -1; r = h.rnext[r]) {
            if (r < 0 || r >= Pmax || ++steps > Pmax) return fail(GM_ERR_INVALID, "gm_plan_create: bad restriction chain");
            if (h.rindex[r] < 0 || h.rindex[r] >= v) return fail(GM_ERR_INVALID, "gm_plan_create: restriction must refer to an earlier vertex");
        }
        steps = 0;
        for (int q = h.last[v]; q != -1; q = h.next[q])
            if (++steps > P) return fail(GM_ERR_INVALID, "gm_plan_create: prefix list cycle");
    }
    // Each prefix is built when the last vertex of its list is fixed; a father
    // must be built at an earlier depth, a loop set before its depth, an IEP
    // set before the IEP depth.
    int owner[GM_MAX_PREFIX];
    for (int q = 0; q < P; ++q) owner[q] = -1;
    for (int v = 0; v < n; ++v)
        for (int q = h.last[v]; q != -1; q = h.next[q]) {
            if (owner[q] != -1) return fail(GM_ERR_INVALID, "gm_plan_create: prefix in two lists");
            owner[q] = v;
        }
    for (int q = 0; q < P; ++q) {
        if (owner[q] == -1) return fail(GM_ERR_INVALID, "gm_plan_create: prefix in no list");
        if (h.father[q] >= 0 && owner[h.father[q]] >= owner[q]) return fail(GM_ERR_INVALID, "gm_plan_create: father built after child");
    }
    for (int v = 1; v < n; ++v)
        if (owner[h.loopset[v]] >= v) return fail(GM_ERR_INVALID, "gm_plan_create: loop set built too late");
    for (int i = 0; i < h.iep_vn; ++i) {
        h.iep_id[i] = d->iep_vertex_id[i];
        h.iep_vflag[i] = d->iep_vertex_flag[i] ? 1 : 0;
        h.iep_vcoef[i] = d->iep_vertex_coef[i];
        if (h.iep_id[i] < 0 || h.iep_id[i] >= P) return fail(GM_ERR_INVALID, "gm_plan_create: bad IEP prefix");
        if (owner[h.iep_id[i]] >= n - k) return fail(GM_ERR_INVALID, "gm_plan_create: IEP set built too late");
    }
    for (int i = 0; i < h.iep_an; ++i) {
        h.iep_coef[i] = d->iep_coef[i];
        h.iep_flag[i] = d->iep_flag[i] ? 1 : 0;
        h.iep_pos[i] = d->iep_ans_pos[i];
        if (h.iep_pos[i] < 0 || h.iep_pos[i] >= h.iep_vn) return fail(GM_ERR_INVALID, "gm_plan_create: bad IEP position");
    }
    // Size-only prefixes, derived from the plan structure rather than taken
    // from the plan's flags: no child prefix, not a loop set before the IEP
    // depth, and every IEP row reading it uses |S| - coef.
    if (k > 0) {
        for (int i = 0; i < h.iep_vn; ++i) h.size_only[h.iep_id[i]] = 1;
        for (int i = 0; i < h.iep_vn; ++i)
            if (!h.iep_vflag[i]) h.size_only[h.iep_id[i]] = 0;
        for (int q = 0; q < P; ++q)
            if (h.father[q] >= 0) h.size_only[h.father[q]] = 0;
        for (int dd = 1; dd < n - k; ++dd) h.size_only[h.loopset[dd]] = 0;
    }
    gm_plan* p = new gm_plan();
    p->host = h;
    p->d_plan = nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        delete p;
        return fail(GM_ERR_HIP, "gm_plan_create: no device");
    }
    p->device = dev;
    if (hipMalloc(&p->d_plan, sizeof(DevPlan)) != hipSuccess) {
        delete p;
        return fail(GM_ERR_OOM, "gm_plan_create: device allocation failed");
    }
    if (hipMemcpy(p->d_plan, &p->host, sizeof(DevPlan), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p->d_plan);
        delete p;
        return fail(GM_ERR_HIP, "gm_plan_create: upload failed");
    }
    *out = p;
    return GM_OK;
}

void gm_plan_destroy(gm_plan* p) {
    if (!p) return;
    (void)hipFree(p->d_plan);
    delete p;
}

int gm_count_multi(const gm_graph* g, const gm_plan* const* plans, int32_t np, const gm_tuning* t, int32_t flags,
                   void* stream, uint64_t* raw_sums, gm_stats* stats);

int gm_count(const gm_graph* g, const gm_plan* p, const gm_tuning* t, int32_t flags, void* stream, uint64_t* raw_sum,
             gm_stats* stats) {
    return gm_count_multi(g, &p, 1, t, flags, stream, raw_sum, stats);
}

int gm_count_multi(const gm_graph* g, const gm_plan* const* plans, int32_t np, const gm_tuning* t, int32_t flags,
                   void* stream, uint64_t* raw_sums, gm_stats* stats) {
    if (!g || !plans || np <= 0 || !raw_sums) return fail(GM_ERR_INVALID, "gm_count: bad arguments");
    for (int i = 0; i < np; ++i)
        if (!plans[i]) return fail(GM_ERR_INVALID, "gm_count: null plan");
    HIP_TRY(hipSetDevice(g->device));
    hipStream_t s = (hipStream_t)stream;
    DeviceCache& dc = cache_for(g->device);
    std::lock_guard<std::mutex> lock(dc.mu);
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    int waves_per_cu = (t && t->waves_per_cu > 0) ? t->waves_per_cu : 8;
    int blocks = std::max(1, cus * waves_per_cu / kWavesPerBlock);
    const int64_t nwaves = (int64_t)blocks * kWavesPerBlock;
    const int32_t cap = std::max(1, g->second_deg);
    const int32_t cap_vi = std::max(1, g->max_deg);
    if (!dc.ctr) HIP_TRY(hipMalloc(&dc.ctr, sizeof(Counters)));
    if (!dc.ev0) {
        HIP_TRY(hipEventCreate(&dc.ev0));
        HIP_TRY(hipEventCreate(&dc.ev1));
    }
    double total_ms = 0;
    unsigned long long total_bytes = 0, total_tasks = 0;
    for (int pi = 0; pi < np; ++pi) {
        const gm_plan* p = plans[pi];
        const int n = p->host.n;
        const int64_t stride = (int64_t)p->host.P * cap + (int64_t)(p->host.vinduced ? n : 0) * cap_vi;
        const size_t need = (size_t)std::max<int64_t>(stride, 1) * (size_t)nwaves * sizeof(uint32_t);
        if (need > dc.scratch_bytes) {
            if (dc.scratch) HIP_TRY(hipFree(dc.scratch));
            dc.scratch = nullptr;
            dc.scratch_bytes = 0;
            HIP_TRY(hipMalloc(&dc.scratch, need));
            dc.scratch_bytes = need;
        }
        HIP_TRY(hipMemsetAsync(dc.ctr, 0, sizeof(Counters), s));
        KernelArgs a;
        a.row = g->d_row;
        a.col = g->d_col;
        a.nv = g->n;
        a.nnz = g->nnz;
        a.plan = p->d_plan;
        a.scratch = (uint32_t*)dc.scratch;
        a.wave_stride = stride;
        a.cap = cap;
        a.cap_vi = cap_vi;
        a.nchunks = (g->nnz + kWave - 1) / kWave;
        a.ctr = dc.ctr;
        a.stats = (flags & GM_COUNT_ALG_BYTES) ? 1 : 0;
        HIP_TRY(hipEventRecord(dc.ev0, s));
        if (p->host.mode == GM_MODE_CLIQUE)
            hipLaunchKernelGGL(enumerate_kernel<GM_MODE_CLIQUE>, dim3(blocks), dim3(kBlock), 0, s, a);
        else
            hipLaunchKernelGGL(enumerate_kernel<GM_MODE_PATTERN>, dim3(blocks), dim3(kBlock), 0, s, a);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(dc.ev1, s));
        Counters c;
        HIP_TRY(hipMemcpyAsync(&c, dc.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, dc.ev0, dc.ev1));
        if (c.overflow) return fail(GM_ERR_INVALID, "gm_count: an intersection exceeded its scratch slot (graph/plan mismatch)");
        raw_sums[pi] = c.sum;
        total_ms += ms;
        total_bytes += c.bytes;
        total_tasks += c.tasks;
    }
    if (stats) {
        stats->kernel_ms = total_ms;
        stats->alg_bytes = total_bytes;
        stats->root_tasks = total_tasks;
        stats->grid_waves = (int32_t)nwaves;
        stats->flags_used = flags;
    }
    return GM_OK;
}

}  // extern "C"
