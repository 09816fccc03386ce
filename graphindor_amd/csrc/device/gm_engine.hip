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
#include <chrono>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "graphindor_amd.h"

#include "engine_kernels.h"

namespace {

using namespace gmk;

// Adds a clique launch's counters into the call's totals.
__global__ void accumulate_counters_kernel(Counters* total, const Counters* part) {
    total->bytes += part->bytes;
    total->tasks += part->tasks;
    total->overflow |= part->overflow;
    total->timed_out |= part->timed_out;
}

// Hub bitmaps: block b writes the bitmap of hub hubs[b] (its neighbour set
// over vertex ids) and records its slot.
__global__ void hub_build_kernel(const int64_t* row, const uint32_t* col, const int32_t* hubs, int64_t words,
                                 unsigned long long* bits, int32_t* slot) {
    const int32_t h = hubs[blockIdx.x];
    unsigned long long* bm = bits + (int64_t)blockIdx.x * words;
    for (int64_t i = threadIdx.x; i < words; i += blockDim.x) bm[i] = 0ull;
    __syncthreads();
    for (int64_t e = row[h] + threadIdx.x; e < row[h + 1]; e += blockDim.x) {
        const uint32_t x = col[e];
        atomicOr(&bm[x >> 6], 1ull << (x & 63u));
    }
    if (threadIdx.x == 0) slot[h] = (int32_t)blockIdx.x;
}

// Zeroes n 64-bit words (a kernel, so that it may be captured in a graph).
__global__ void zero_u64_kernel(unsigned long long* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
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
// Kernel instantiation for a launch whose largest pattern has n vertices.
int nmax_for(int n) { return n <= 4 ? 4 : n; }

// Depth-3 donation (GM_COUNT_DEEP_DONATION) is refused until its stall on
// hub-heavy graphs is understood: a 5-cycle sample of R-MAT 2^22 that takes
// 8 s without it ran past 180 s with it.  The kernels keep the code.
const char* const kDeepDonationOff =
    "GM_COUNT_DEEP_DONATION is switched off: depth-3 donation stalled on hub-heavy graphs";

// Waves per SIMD of the build a launch uses.  The clique rule has the
// 2-wave build only.  The pattern rule runs the 4-wave build (128 VGPRs,
// some state in scratch) when the tuning asks for more than 8 resident waves
// per CU, the 2-wave build (all state in registers) when it asks for 8 or
// fewer; by default the 4-wave build for launches of at least kWideMinRoots
// root edges.  Small launches keep the 2-wave build: a launch of the 4-wave
// build costs about 1 ms more (3,072 per-graph 4-motif launches over the
// 512-graph batch: 9.4 s of kernel time against 6.5 s), and a launch of few
// roots is bound by its heaviest slices, which the faster waves finish sooner
// (the 31-root 5-cycle sample: 2.8 s against 3.4 s).
constexpr int64_t kWideMinRoots = 1 << 14;
int enum_wpe(int mode, int req_waves, int64_t roots) {
    if (mode != GM_MODE_PATTERN) return 2;
    if (req_waves > 0) return req_waves > 8 ? 4 : 2;
    return roots >= kWideMinRoots ? 4 : 2;
}

// Root edges e of [lo, hi) with e % stride == 0.
int64_t sampled_roots(int64_t lo, int64_t hi, int64_t stride) {
    const int64_t first = (lo + stride - 1) / stride * stride;
    return first < hi ? (hi - first + stride - 1) / stride : 0;
}

template <int MODE>
hipError_t enum_occupancy(int nmax, int wpe, int* blocks_per_cu) {
    if (MODE == 0 && wpe == 4) {
        switch (nmax) {
            case 4: return enum_occupancy_0_4_4(blocks_per_cu);
            case 5: return enum_occupancy_0_5_4(blocks_per_cu);
            case 6: return enum_occupancy_0_6_4(blocks_per_cu);
            default: return enum_occupancy_0_7_4(blocks_per_cu);
        }
    }
    switch (nmax) {
        case 4: return MODE == 0 ? enum_occupancy_0_4_2(blocks_per_cu) : enum_occupancy_1_4_2(blocks_per_cu);
        case 5: return MODE == 0 ? enum_occupancy_0_5_2(blocks_per_cu) : enum_occupancy_1_5_2(blocks_per_cu);
        case 6: return MODE == 0 ? enum_occupancy_0_6_2(blocks_per_cu) : enum_occupancy_1_6_2(blocks_per_cu);
        default: return MODE == 0 ? enum_occupancy_0_7_2(blocks_per_cu) : enum_occupancy_1_7_2(blocks_per_cu);
    }
}

template <int MODE>
void enum_launch(int nmax, int wpe, int blocks, hipStream_t s, const KernelArgs& a) {
    if (MODE == 0 && wpe == 4) {
        switch (nmax) {
            case 4: enum_launch_0_4_4(blocks, s, a); break;
            case 5: enum_launch_0_5_4(blocks, s, a); break;
            case 6: enum_launch_0_6_4(blocks, s, a); break;
            default: enum_launch_0_7_4(blocks, s, a); break;
        }
        return;
    }
    switch (nmax) {
        case 4: MODE == 0 ? enum_launch_0_4_2(blocks, s, a) : enum_launch_1_4_2(blocks, s, a); break;
        case 5: MODE == 0 ? enum_launch_0_5_2(blocks, s, a) : enum_launch_1_5_2(blocks, s, a); break;
        case 6: MODE == 0 ? enum_launch_0_6_2(blocks, s, a) : enum_launch_1_6_2(blocks, s, a); break;
        default: MODE == 0 ? enum_launch_0_7_2(blocks, s, a) : enum_launch_1_7_2(blocks, s, a); break;
    }
}

// Persistent grid for a kernel: `waves_per_cu` (0 = all the occupancy allows).
int grid_blocks(int cus, int blocks_per_cu, int waves_per_cu) {
    blocks_per_cu = std::max(1, blocks_per_cu);
    const int w = waves_per_cu > 0 ? std::min(waves_per_cu, blocks_per_cu * kWavesPerBlock) : blocks_per_cu * kWavesPerBlock;
    return std::max(1, cus * w / kWavesPerBlock);
}

// Counters per call: [0] totals, [1] one clique list launch, [2] one bitmap
// count, [3] bitmap totals, [4] one block-per-root launch, [5] its totals.
constexpr int kNumCounters = 6;

struct DeviceCache {
    std::mutex mu;
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    void* bm_rows = nullptr;
    size_t bm_bytes = 0;
    void* plan_ptrs = nullptr;
    size_t plan_ptrs_bytes = 0;
    void* gedge = nullptr;
    size_t gedge_bytes = 0;
    void* sums = nullptr;
    size_t sums_bytes = 0;
    Counters* ctr = nullptr;  // [0] call totals, [1] one clique launch, [2] one bitmap launch, [3] bitmap totals
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> evpool;  // start/stop pairs around individual launches
    SliceItem* queue = nullptr;      // donated depth-2 slices (ready flags all-clear between launches)
};

// Slots of the slice queue, each used once per launch: 160 MiB of items per
// device.  At 2^20 slots the house sample of R-MAT 2^22 (1/4096 of the
// roots) pushed 1,023,669 slices, and with 1.5x the waves it filled the
// queue, after which idle waves leave and the rest ran 4x slower.
constexpr int64_t kQueueCap = 1 << 22;

// The slice queue's counters of a list-engine launch (wall clock: 100 MHz).
void fill_queue_stats(const Counters& c, gm_stats* st) {
    st->slices = c.qtail < (unsigned long long)kQueueCap ? c.qtail : (unsigned long long)kQueueCap;
    st->donations = c.donations;
    st->deep_donations = c.deep_donations;
    st->queue_retries = c.retries;
    st->slice_wave_ms = (double)c.slice_ticks * 1e-5;
    st->idle_wave_ms = (double)c.idle_ticks * 1e-5;
}

// Slice length from the tuning: > 0 donates in slices of at least that
// length, 0 selects the default, -1 does not donate.  On R-MAT 2^14 and 2^16
// (root-sampled, lane tail off) donation took house from 36 to 19 ms and
// 123 to 48 ms, k4+pendant from 7.9 to 4.4 ms and 20.5 to 5.7 ms, the
// 5-cycle from 1.42 to 0.65 s and 10.0 to 3.3 s; with the lane tail on,
// house and k4+pendant donate nothing (their depth 2 is the tail) and the
// 5-cycle gains 2.5x.
constexpr int32_t kDefaultSliceLen = 8;
int32_t slice_len_of(const gm_tuning* t) {
    if (t && t->slice_len < 0) return 0;
    return (t && t->slice_len > 0) ? t->slice_len : kDefaultSliceLen;
}

// Father elements the pattern tail stages in LDS: gm_tuning.lds_set_cap
// (0 = the whole per-wave arena, -1 = never).
int32_t stage_cap_of(const gm_tuning* t) {
    if (t && t->lds_set_cap < 0) return -1;
    if (t && t->lds_set_cap > 0) return std::min<int32_t>(t->lds_set_cap, kStage);
    return kStage;
}

hipError_t alloc_queue(SliceItem** q) {
    if (*q) return hipSuccess;
    hipError_t e = hipMalloc((void**)q, sizeof(SliceItem) * kQueueCap);
    if (e != hipSuccess) return e;
    return hipMemset(*q, 0, sizeof(SliceItem) * kQueueCap);
}

DeviceCache& cache_for(int dev) {
    static std::mutex mu;
    static std::vector<DeviceCache*> caches;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)caches.size() <= dev) caches.resize(dev + 1, nullptr);
    if (!caches[dev]) caches[dev] = new DeviceCache();
    return *caches[dev];
}

// Depth whose candidates pattern_tail counts lane-parallel, or -1: the last
// candidate depth (n-k-1 with an IEP tail, n-2 without), at least 2, with at
// most kTailMaxPrefix prefixes built there and kTailMaxIep IEP rows; not for
// the vertex-induced rule.  The oracle applies the same rule to its byte model.
int pattern_tail_depth(const DevPlan& h) {
    if (h.mode != GM_MODE_PATTERN || h.vinduced) return -1;
    const int D = h.k >= 2 ? h.n - h.k - 1 : (h.k == 0 ? h.n - 2 : -1);
    if (D < 2) return -1;
    int np = 0;
    for (int q = h.last[D]; q != -1; q = h.next[q]) ++np;
    if (np > kTailMaxPrefix) return -1;
    if (h.k >= 2 && h.iep_vn > kTailMaxIep) return -1;
    return D;
}

// One 64-lane block per resident slot of the bitmap kernel (its LDS bounds
// residency at 10 blocks per CU).
int bm_blocks(int cus) { return cus * 10; }

// Block-per-root kernel: rows of d x W words in dynamic LDS, up to what a
// block may allocate beside the kernel's static LDS.
constexpr int kBigMinD = 64;

hipError_t big_rows_capacity(int device, int32_t* cap_words, size_t* dyn_bytes) {
    int maxsh = 0;
    hipError_t e = hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    if (e != hipSuccess) return e;
    const size_t static_bytes = sizeof(uint32_t) * kBmMaxD + sizeof(int32_t) * (kBmMaxD + 1) +
                                64 + sizeof(unsigned long long) * 2 * (kBigThreads / kWave) + 64;
    const size_t avail = (size_t)maxsh > static_bytes ? (size_t)maxsh - static_bytes : 0;
    const size_t words = std::min<size_t>(avail / sizeof(unsigned long long), (size_t)kBmMaxD * kBmMaxW);
    *cap_words = (int32_t)words;
    *dyn_bytes = words * sizeof(unsigned long long);
    return hipSuccess;
}

// Resident blocks of the block-per-root kernel with `dyn` bytes of row LDS:
// one launch fills every slot (the kernel claims roots from a cursor).
// The block kernel for clique size k and row triangularity tri (used for k = 4, 5).
template <int K, int TRI>
void* big_kernel_fn() { return reinterpret_cast<void*>(&clique_bitmap_big_kernel<K, TRI>); }

void* big_kernel_for(int k, int tri) {
    switch (k) {
        case 3: return big_kernel_fn<3, 0>();
        case 4: return tri > 0 ? big_kernel_fn<4, 1>() : tri < 0 ? big_kernel_fn<4, -1>() : big_kernel_fn<4, 0>();
        case 5: return tri > 0 ? big_kernel_fn<5, 1>() : tri < 0 ? big_kernel_fn<5, -1>() : big_kernel_fn<5, 0>();
        case 6: return big_kernel_fn<6, 0>();
        case 7: return big_kernel_fn<7, 0>();
        default: return nullptr;
    }
}

int big_blocks(int k, int tri, int cus, size_t dyn) {
    int occ = 0;
    void* fn = big_kernel_for(k, tri);
    hipError_t e = fn ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, kBigThreads, dyn) : hipErrorInvalidValue;
    if (e != hipSuccess || occ < 1) occ = 1;
    return cus * std::min(occ, 4);
}

hipError_t launch_bitmap_big(int k, int tri, int blocks, size_t dyn, const BmArgs& b, hipStream_t s) {
    void* fn = big_kernel_for(k, tri);
    if (!fn) return hipErrorInvalidValue;
    BmArgs arg = b;
    void* params[] = {&arg};
    return hipLaunchKernel(fn, dim3(blocks), dim3(kBigThreads), params, dyn, s);
}

// Small-root kernel: one 64-lane block per resident slot.
template <int K>
void* small_kernel_fn() { return reinterpret_cast<void*>(&clique_bitmap_small_kernel<K>); }

void* small_kernel_for(int k) {
    switch (k) {
        case 3: return small_kernel_fn<3>();
        case 4: return small_kernel_fn<4>();
        case 5: return small_kernel_fn<5>();
        case 6: return small_kernel_fn<6>();
        case 7: return small_kernel_fn<7>();
        default: return nullptr;
    }
}

hipError_t launch_bitmap_small(int k, int cus, const BmArgs& b, hipStream_t s) {
    void* fn = small_kernel_for(k);
    if (!fn) return hipErrorInvalidValue;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, kWave, 0) != hipSuccess || occ < 1) occ = 8;
    BmArgs arg = b;
    void* params[] = {&arg};
    return hipLaunchKernel(fn, dim3(cus * std::min(occ, 32)), dim3(kWave), params, 0, s);
}

hipError_t launch_bitmap(int k, int blocks, const BmArgs& b, hipStream_t s) {
    switch (k) {
        case 3: hipLaunchKernelGGL((clique_bitmap_kernel<kBmMaxW, 3>), dim3(blocks), dim3(kWave), 0, s, b); break;
        case 4: hipLaunchKernelGGL((clique_bitmap_kernel<kBmMaxW, 4>), dim3(blocks), dim3(kWave), 0, s, b); break;
        case 5: hipLaunchKernelGGL((clique_bitmap_kernel<kBmMaxW, 5>), dim3(blocks), dim3(kWave), 0, s, b); break;
        case 6: hipLaunchKernelGGL((clique_bitmap_kernel<kBmMaxW, 6>), dim3(blocks), dim3(kWave), 0, s, b); break;
        case 7: hipLaunchKernelGGL((clique_bitmap_kernel<kBmMaxW, 7>), dim3(blocks), dim3(kWave), 0, s, b); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

// Static workspace of gm_count_static: counters, sums and the list engine's
// prefix slots of a count that allocates nothing.  Each device holds its own
// copy from module load to process end.  Host code never uses this name as
// an address: static_workspace() looks up the device's own address once.
static __device__ __align__(256) unsigned char scratch[GM_STATIC_WORKSPACE_BYTES];

namespace {

constexpr int kMaxStaticDevices = 64;
std::mutex g_ws_lookup_mu;
unsigned char* g_ws_ptr[kMaxStaticDevices] = {};
std::mutex g_ws_count_mu[kMaxStaticDevices];

std::mutex& static_ws_mutex(int dev) { return g_ws_count_mu[dev]; }

// The device address of `scratch` on `dev` (the current device), looked up
// the first time a call runs there.
int static_workspace(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxStaticDevices) return fail(GM_ERR_UNSUPPORTED, "gm_count_static: device index too large");
    std::lock_guard<std::mutex> lock(g_ws_lookup_mu);
    if (!g_ws_ptr[dev]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
        void* ptr = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(GM_ERR_HIP, std::string("gm_count_static: hipGetSymbolAddress: ") + hipGetErrorString(e));
        g_ws_ptr[dev] = (unsigned char*)ptr;
    }
    *out = g_ws_ptr[dev];
    return GM_OK;
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
    int32_t tri;  // 1: every row holds larger ids only; -1: smaller ids only; 0: neither (or no edges)
    int32_t* d_order;  // vertices by out-degree, largest first (ties by id)
    int32_t n_big;     // out-degree above 64 (the bitmap engine's block and per-wave roots)
    int32_t n_active;  // out-degree at least 2
    std::vector<int32_t> h_order;  // d_order on the host
    std::vector<int32_t> h_deg;    // out-degree of h_order[i]
    // Hub bitmaps (pattern rule), built at first use for a threshold.
    int32_t hub_deg_built = 0;     // threshold they were built for (0: none built)
    int64_t n_hubs = 0;
    int64_t hub_words = 0;
    unsigned long long* d_hub_bits = nullptr;
    int32_t* d_hub_slot = nullptr;
};

namespace {

// Hub bitmaps are off unless asked for: a bitmap over vertex ids turns a
// binary search in a hub's list into one load, but its lines are scattered
// over n/8 bytes, and on R-MAT 2^22 the house pattern ran 2.7x slower with
// them while the 4-clique + pendant ran 1.38x faster.  A tuner that measures
// both picks the threshold per pattern.
constexpr int32_t kDefaultHubDegree = 0;
constexpr size_t kHubBytesCap = (size_t)4 << 30;

// Builds (or reuses) the hub bitmaps of graph g for threshold T on stream s.
int ensure_hub_bitmaps(gm_graph* g, int32_t T, hipStream_t s) {
    if (g->hub_deg_built == T) return GM_OK;
    if (g->d_hub_bits) (void)hipFree(g->d_hub_bits);
    g->d_hub_bits = nullptr;
    g->n_hubs = 0;
    g->hub_deg_built = 0;
    g->hub_words = ((int64_t)g->n + 63) / 64;
    int64_t H = 0;
    while (H < (int64_t)g->h_deg.size() && g->h_deg[(size_t)H] >= T) ++H;
    const size_t per = (size_t)g->hub_words * sizeof(unsigned long long);
    if (per > 0 && (size_t)H * per > kHubBytesCap) H = (int64_t)(kHubBytesCap / per);
    if (H > 0) {
        if (!g->d_hub_slot) HIP_TRY(hipMalloc(&g->d_hub_slot, sizeof(int32_t) * (size_t)g->n));
        HIP_TRY(hipMalloc(&g->d_hub_bits, per * (size_t)H));
        int32_t* d_hubs = nullptr;
        HIP_TRY(hipMalloc(&d_hubs, sizeof(int32_t) * (size_t)H));
        hipError_t e = hipMemsetAsync(g->d_hub_slot, 0xff, sizeof(int32_t) * (size_t)g->n, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_hubs, g->h_order.data(), sizeof(int32_t) * (size_t)H, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(hub_build_kernel, dim3((unsigned)H), dim3(256), 0, s, g->d_row, g->d_col, d_hubs,
                               g->hub_words, g->d_hub_bits, g->d_hub_slot);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(d_hubs);
        if (e != hipSuccess) {
            (void)hipFree(g->d_hub_bits);
            g->d_hub_bits = nullptr;
            return fail(GM_ERR_HIP, std::string("hub bitmaps: ") + hipGetErrorString(e));
        }
    }
    g->n_hubs = H;
    g->hub_deg_built = T;
    return GM_OK;
}

}  // namespace

struct gm_batch;

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
    bool upper = nnz > 0, lower = nnz > 0;
    for (int32_t v = 0; v < n; ++v) {
        int64_t l = row_ptr[v], r = row_ptr[v + 1];
        if (r < l) return fail(GM_ERR_INVALID, "gm_graph_create: row_ptr not monotone at row " + std::to_string(v));
        for (int64_t e = l; e < r; ++e) {
            int32_t u = col[e];
            if (u < 0 || u >= n) return fail(GM_ERR_INVALID, "gm_graph_create: neighbour id out of range in row " + std::to_string(v));
            if (u == v) return fail(GM_ERR_INVALID, "gm_graph_create: self-loop at vertex " + std::to_string(v));
            if (e > l && col[e - 1] >= u) return fail(GM_ERR_INVALID, "gm_graph_create: row " + std::to_string(v) + " not strictly increasing");
        }
        // Rows are sorted: their first and last ids decide the orientation.
        if (r > l) {
            upper = upper && col[l] > v;
            lower = lower && col[r - 1] < v;
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
    g->tri = upper ? 1 : (lower ? -1 : 0);
    // The bitmap kernels claim their roots heaviest first: a counting sort of
    // the vertices by decreasing out-degree, stable in id.
    std::vector<int32_t> order((size_t)n);
    {
        std::vector<int64_t> start((size_t)d1 + 2, 0);
        for (int32_t v = 0; v < n; ++v) ++start[(size_t)(d1 - (row_ptr[v + 1] - row_ptr[v])) + 1];
        for (size_t i = 1; i < start.size(); ++i) start[i] += start[i - 1];
        for (int32_t v = 0; v < n; ++v) order[(size_t)start[(size_t)(d1 - (row_ptr[v + 1] - row_ptr[v]))]++] = v;
    }
    g->h_deg.resize((size_t)n);
    for (int32_t i = 0; i < n; ++i) g->h_deg[(size_t)i] = (int32_t)(row_ptr[order[(size_t)i] + 1] - row_ptr[order[(size_t)i]]);
    g->n_big = 0;
    g->n_active = 0;
    for (int32_t v = 0; v < n; ++v) {
        const int64_t d = row_ptr[v + 1] - row_ptr[v];
        g->n_big += d > kSmallD ? 1 : 0;
        g->n_active += d >= 2 ? 1 : 0;
    }
    g->d_order = nullptr;
    g->d_row = nullptr;
    g->d_col = nullptr;
    hipError_t e1 = hipMalloc(&g->d_row, sizeof(int64_t) * ((size_t)n + 1));
    hipError_t e2 = hipMalloc(&g->d_col, sizeof(uint32_t) * (size_t)(nnz > 0 ? nnz : 1));
    hipError_t e3 = hipMalloc(&g->d_order, sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        if (g->d_row) (void)hipFree(g->d_row);
        if (g->d_col) (void)hipFree(g->d_col);
        if (g->d_order) (void)hipFree(g->d_order);
        delete g;
        return fail(GM_ERR_OOM, "gm_graph_create: device allocation failed");
    }
    hipError_t c1 = hipMemcpy(g->d_row, row_ptr, sizeof(int64_t) * ((size_t)n + 1), hipMemcpyHostToDevice);
    hipError_t c2 = nnz > 0 ? hipMemcpy(g->d_col, col, sizeof(uint32_t) * (size_t)nnz, hipMemcpyHostToDevice) : hipSuccess;
    hipError_t c3 = n > 0 ? hipMemcpy(g->d_order, order.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice)
                          : hipSuccess;
    if (c1 != hipSuccess || c2 != hipSuccess || c3 != hipSuccess) {
        (void)hipFree(g->d_row);
        (void)hipFree(g->d_col);
        (void)hipFree(g->d_order);
        delete g;
        return fail(GM_ERR_HIP, "gm_graph_create: upload failed");
    }
    g->h_order = std::move(order);
    *out = g;
    return GM_OK;
}

void gm_graph_destroy(gm_graph* g) {
    if (!g) return;
    (void)hipFree(g->d_row);
    (void)hipFree(g->d_col);
    (void)hipFree(g->d_order);
    if (g->d_hub_bits) (void)hipFree(g->d_hub_bits);
    if (g->d_hub_slot) (void)hipFree(g->d_hub_slot);
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
    if (d->iep_vertex_num < 0 || d->iep_vertex_num > GM_MAX_IEP_VERTEX || d->iep_array_size < 0 || d->iep_array_size > GM_MAX_IEP)
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
    h.tail_depth = pattern_tail_depth(h);
    if (h.mode == GM_MODE_CLIQUE && n >= 3) {
        // Only the size of the last loop set is read (gpu/gpu_clique.cu:229-232).
        int q = h.loopset[n - 1];
        bool has_child = false;
        for (int c = 0; c < P; ++c) has_child |= (h.father[c] == q);
        bool earlier = false;
        for (int dd = 1; dd < n - 1; ++dd) earlier |= (h.loopset[dd] == q);
        if (h.father[q] >= 0 && !has_child && !earlier) h.size_only[q] = 1;
        // Depth n-2 can be counted lane-parallel when it builds exactly that
        // set, from its own loop set.
        if (n >= 4 && h.size_only[q] && h.last[n - 2] == q && h.next[q] == -1 && h.father[q] == h.loopset[n - 2])
            h.clique_tail = 1;
        // The plain chain: complete pattern, one prefix per depth, each the
        // father of the next, prefix of depth d the loop set of d+1.
        bool chain = h.nrestrict == 0;
        for (int i = 0; i < n && chain; ++i)
            for (int j = 0; j < n; ++j)
                if (i != j && !h.adj[i * n + j]) chain = false;
        int prev = -1;
        for (int dd = 0; dd + 1 < n && chain; ++dd) {
            int qd = h.last[dd];
            if (qd < 0 || h.next[qd] != -1 || h.father[qd] != prev || h.loopset[dd + 1] != qd) chain = false;
            prev = qd;
        }
        h.clique_chain = chain ? 1 : 0;
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

// The counting call behind gm_count, gm_count_multi and gm_count_batch
// (fixed_ws == nullptr: device buffers come from the per-device cache, grown
// on demand) and gm_count_static (fixed_ws: the device address of the static
// workspace, GM_STATIC_WORKSPACE_BYTES, from which every buffer is carved;
// nothing is allocated, slice donation is off, no HIP event is used and the
// list engine's grid is cut to the waves whose prefix slots fit).
static int count_impl(const gm_graph* g, const int64_t* graph_edge_offsets, int32_t ng, const gm_plan* const* plans,
                      int32_t np, const gm_tuning* t, int32_t flags, void* stream, uint64_t* raw_sums, gm_stats* stats,
                      unsigned char* fixed_ws) {
    if (!g || !plans || np <= 0 || !raw_sums || ng < 1) return fail(GM_ERR_INVALID, "gm_count: bad arguments");
    for (int i = 0; i < np; ++i)
        if (!plans[i]) return fail(GM_ERR_INVALID, "gm_count: null plan");
    if (flags & GM_COUNT_DEEP_DONATION)
        return fail(GM_ERR_UNSUPPORTED, kDeepDonationOff);
    if (graph_edge_offsets) {
        if (graph_edge_offsets[0] < 0 || graph_edge_offsets[ng] > g->nnz)
            return fail(GM_ERR_INVALID, "gm_count_batch: edge offsets outside [0, nnz]");
        for (int i = 0; i < ng; ++i)
            if (graph_edge_offsets[i + 1] < graph_edge_offsets[i])
                return fail(GM_ERR_INVALID, "gm_count_batch: edge offsets not monotone");
    } else if (ng > 1) {
        return fail(GM_ERR_INVALID, "gm_count_batch: edge offsets required for more than one graph");
    }
    HIP_TRY(hipSetDevice(g->device));
    hipStream_t s = (hipStream_t)stream;
    DeviceCache fixed_dc;  // fixed_ws: views of the static workspace, owned by nobody
    DeviceCache& dc = fixed_ws ? fixed_dc : cache_for(g->device);
    std::unique_lock<std::mutex> lock(fixed_ws ? static_ws_mutex(g->device) : dc.mu);
    const auto t_host0 = std::chrono::steady_clock::now();
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    // Default grids: every block the occupancy of the launched kernel allows
    // (persistent waves; the kernels exit when the task cursor runs out).
    int nmax_pat = 0, nmax_clq = 0;
    for (int i = 0; i < np; ++i) {
        int& m = plans[i]->host.mode == GM_MODE_PATTERN ? nmax_pat : nmax_clq;
        m = std::max(m, nmax_for(plans[i]->host.n));
    }
    const int req_waves = (t && t->waves_per_cu > 0) ? t->waves_per_cu : 0;
    // The static workspace runs the 2-wave build: the 4-wave build's spill
    // space is private scratch the runtime would have to provide.
    const int64_t roots_est =
        sampled_roots(graph_edge_offsets ? graph_edge_offsets[0] : 0, graph_edge_offsets ? graph_edge_offsets[ng] : g->nnz,
                      (t && t->root_stride > 1) ? t->root_stride : 1);
    const int wpe_pat = fixed_ws ? 2 : enum_wpe(GM_MODE_PATTERN, req_waves, roots_est);
    int blocks_pat = 0, blocks_clq = 0;
    if (nmax_pat) {
        int occ = 0;
        HIP_TRY(enum_occupancy<GM_MODE_PATTERN>(nmax_pat, wpe_pat, &occ));
        blocks_pat = grid_blocks(cus, occ, req_waves);
    }
    if (nmax_clq) {
        int occ = 0;
        HIP_TRY(enum_occupancy<GM_MODE_CLIQUE>(nmax_clq, 2, &occ));
        blocks_clq = grid_blocks(cus, occ, req_waves);
    }
    int64_t nwaves = (int64_t)std::max(blocks_pat, blocks_clq) * kWavesPerBlock;
    const int32_t cap = std::max(1, g->second_deg);
    const int32_t cap_vi = std::max(1, g->max_deg);
    const int task_max = (t && t->chunk_edges > 0) ? t->chunk_edges : kWave;
    int bm_max = (t && t->bitmap_max_degree != 0) ? std::min(t->bitmap_max_degree, kBmMaxD) : kBmMaxD;
    if (fixed_ws && bm_max > 0) {
        // Without the per-wave bitmap kernel's global rows, the bitmap
        // engine takes only the roots the small-root and block kernels
        // cover; the list engine counts the rest.
        int32_t cover = std::min(kSmallD, bm_max);
        const int big_min = (t && t->bitmap_block_degree > 0) ? t->bitmap_block_degree : kBigMinD;
        if (!(t && t->bitmap_block_degree < 0) && big_min <= kSmallD) {
            int32_t cap_words = 0;
            size_t dyn = 0;
            HIP_TRY(big_rows_capacity(g->device, &cap_words, &dyn));
            int32_t d = bm_max;
            while (d > cover && (int64_t)d * ((d + 63) / 64) > cap_words) --d;
            cover = std::max(cover, d);
        }
        bm_max = cover;
    }
    const int stats_on = (flags & GM_COUNT_ALG_BYTES) ? 1 : 0;
    const int64_t root_stride = (t && t->root_stride > 1) ? t->root_stride : 1;

    // Plans in launch order: pattern plans first (one fused launch), then
    // clique plans.
    std::vector<int> order;
    for (int i = 0; i < np; ++i)
        if (plans[i]->host.mode == GM_MODE_PATTERN) order.push_back(i);
    const int np_pat = (int)order.size();
    if (root_stride > 1 && np_pat != np) return fail(GM_ERR_INVALID, "gm_count: root_stride applies to GM_MODE_PATTERN plans only");
    for (int i = 0; i < np; ++i)
        if (plans[i]->host.mode == GM_MODE_CLIQUE) order.push_back(i);
    const int np_clq = np - np_pat;

    int64_t stride = 1;
    for (int i = 0; i < np; ++i) {
        const DevPlan& h = plans[i]->host;
        stride = std::max<int64_t>(stride, (int64_t)h.P * cap + (int64_t)(h.vinduced ? h.n : 0) * cap_vi);
    }
    auto grow = [&](void** buf, size_t* have, size_t need) -> hipError_t {
        if (need <= *have) return hipSuccess;
        if (*buf) {
            hipError_t e = hipFree(*buf);
            if (e != hipSuccess) return e;
        }
        *buf = nullptr;
        *have = 0;
        hipError_t e = hipMalloc(buf, need);
        if (e == hipSuccess) *have = need;
        return e;
    };
    // Is the list engine launched at all?  Pattern plans always; a clique
    // plan unless the bitmap engine takes every root.
    bool list_needed = np_pat > 0;
    for (int c = 0; c < np_clq; ++c) {
        const gm_plan* p = plans[order[np_pat + c]];
        const bool bm = ng == 1 && !graph_edge_offsets && p->host.clique_chain && bm_max > 0;
        if (!(bm && g->max_deg <= bm_max)) list_needed = true;
    }
    if (fixed_ws) {
        // Carve the buffers from the static workspace; the prefix slots get
        // what is left, and the list grid shrinks to the waves they hold.
        size_t off = 0;
        auto carve = [&](size_t bytes) {
            unsigned char* p = fixed_ws + off;
            off += (bytes + 255) & ~(size_t)255;
            return (void*)p;
        };
        dc.ctr = (Counters*)carve(sizeof(Counters) * kNumCounters);
        dc.plan_ptrs = carve(sizeof(DevPlan*) * (size_t)np);
        dc.gedge = carve(sizeof(int64_t) * (size_t)(ng + 1));
        dc.sums = carve(sizeof(unsigned long long) * (size_t)ng * (size_t)np);
        if (off > (size_t)GM_STATIC_WORKSPACE_BYTES)
            return fail(GM_ERR_UNSUPPORTED, "gm_count_static: the counters of this call exceed the static workspace");
        dc.scratch = fixed_ws + off;
        dc.scratch_bytes = (size_t)GM_STATIC_WORKSPACE_BYTES - off;
        if (list_needed) {
            const int64_t slot_bytes = stride * (int64_t)sizeof(uint32_t);
            const int64_t fit_blocks = (int64_t)dc.scratch_bytes / (slot_bytes * kWavesPerBlock);
            if (fit_blocks < 1)
                return fail(GM_ERR_UNSUPPORTED, "gm_count_static: one block of list-engine waves needs " +
                                                    std::to_string(slot_bytes * kWavesPerBlock) +
                                                    " bytes of prefix slots; the static workspace has " +
                                                    std::to_string(dc.scratch_bytes));
            if (blocks_pat > fit_blocks) blocks_pat = (int)fit_blocks;
            if (blocks_clq > fit_blocks) blocks_clq = (int)fit_blocks;
            nwaves = (int64_t)std::max(blocks_pat, blocks_clq) * kWavesPerBlock;
        }
    } else {
        HIP_TRY(grow(&dc.scratch, &dc.scratch_bytes, (size_t)stride * (size_t)nwaves * sizeof(uint32_t)));
        HIP_TRY(grow(&dc.plan_ptrs, &dc.plan_ptrs_bytes, sizeof(DevPlan*) * (size_t)np));
        HIP_TRY(grow(&dc.gedge, &dc.gedge_bytes, sizeof(int64_t) * (size_t)(ng + 1)));
        HIP_TRY(grow(&dc.sums, &dc.sums_bytes, sizeof(unsigned long long) * (size_t)ng * (size_t)np));
        if (!dc.ctr) HIP_TRY(hipMalloc(&dc.ctr, sizeof(Counters) * kNumCounters));
        if (!dc.ev0) {
            HIP_TRY(hipEventCreate(&dc.ev0));
            HIP_TRY(hipEventCreate(&dc.ev1));
        }
        // Event pairs: the fused pattern launch, then a list and a bitmap launch per clique plan.
        const size_t nev = 2 * (size_t)(1 + 3 * np_clq);
        while (dc.evpool.size() < nev) {
            hipEvent_t ev = nullptr;
            HIP_TRY(hipEventCreate(&ev));
            dc.evpool.push_back(ev);
        }
    }
    // HIP events around the launches (the static-workspace call times on the host).
    const bool timed = fixed_ws == nullptr;
    auto rec = [&](hipEvent_t ev) -> hipError_t { return timed ? hipEventRecord(ev, s) : hipSuccess; };
    std::vector<int> ev_list, ev_bitmap, ev_block;  // indices of start events used
    std::vector<const DevPlan*> hptrs(np);
    for (int k = 0; k < np; ++k) hptrs[k] = plans[order[k]]->d_plan;
    HIP_TRY(hipMemcpyAsync(dc.plan_ptrs, hptrs.data(), sizeof(DevPlan*) * np, hipMemcpyHostToDevice, s));
    std::vector<int64_t> hgedge(2, 0);
    if (graph_edge_offsets) hgedge.assign(graph_edge_offsets, graph_edge_offsets + ng + 1);
    else hgedge[1] = g->nnz;
    HIP_TRY(hipMemcpyAsync(dc.gedge, hgedge.data(), sizeof(int64_t) * (ng + 1), hipMemcpyHostToDevice, s));
    unsigned long long* d_sums = (unsigned long long*)dc.sums;
    const DevPlan* const* d_plans = (const DevPlan* const*)dc.plan_ptrs;
    const int64_t ncells = (int64_t)ng * np;
    const int64_t nctr_words = (int64_t)(sizeof(Counters) * kNumCounters / sizeof(unsigned long long));

    HIP_TRY(rec(dc.ev0));
    hipLaunchKernelGGL(zero_u64_kernel, dim3(64), dim3(256), 0, s, d_sums, ncells);
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, (unsigned long long*)dc.ctr, nctr_words);
    HIP_TRY(hipGetLastError());
    KernelArgs a;
    std::memset(&a, 0, sizeof(a));
    a.row = g->d_row;
    a.col = g->d_col;
    a.nv = g->n;
    a.nnz = g->nnz;
    a.gedge = (const int64_t*)dc.gedge;
    a.ng = ng;
    a.scratch = (uint32_t*)dc.scratch;
    a.wave_stride = stride;
    a.cap = cap;
    a.cap_vi = cap_vi;
    a.ctr = dc.ctr;
    a.stats = stats_on;
    a.tail_off = (t && t->lane_tail < 0) ? 1 : 0;
    a.tail_work = (t && t->lane_tail > 0) ? t->lane_tail : kLaneTailWork;
    a.task_max = task_max;
    a.stage_cap = stage_cap_of(t);
    // Hub bitmaps for the pattern rule (not in the static workspace, which
    // allocates nothing).
    int64_t hubs_used = 0;
    const int32_t hub_deg = (t && t->hub_degree != 0) ? t->hub_degree : kDefaultHubDegree;
    if (np_pat > 0 && !fixed_ws && hub_deg > 0 && g->max_deg >= hub_deg) {
        const int rc = ensure_hub_bitmaps(const_cast<gm_graph*>(g), hub_deg, s);
        if (rc != GM_OK) return rc;
        if (g->n_hubs > 0) {
            a.hub_bits = g->d_hub_bits;
            a.hub_slot = g->d_hub_slot;
            a.hub_words = g->hub_words;
            hubs_used = g->n_hubs;
        }
    }
    // Slice donation applies to the pattern rule; clique plans rely on the
    // degree orientation and the bitmap engine to bound a root's work.
    if (np_pat > 0 && !fixed_ws) {
        const int32_t sl = slice_len_of(t);
        if (sl > 0) {
            HIP_TRY(alloc_queue(&dc.queue));
            a.queue = dc.queue;
            a.queue_cap = kQueueCap;
            a.slice_len = sl;
            a.eager_donation = (flags & GM_COUNT_EAGER_DONATION) ? 1 : 0;
        a.deep_donation = (flags & GM_COUNT_DEEP_DONATION) ? 1 : 0;
        }
    }
    if (t && t->time_limit_ms > 0) {
        int khz = 0;
        HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g->device));
        a.time_limit_ticks = (int64_t)t->time_limit_ms * (int64_t)std::max(khz, 1);
    }
    // Sampling is absolute: the counted roots are the edges e of the range
    // with e % root_stride == 0, so splitting a range never changes the sample.
    a.root_begin = (hgedge[0] + root_stride - 1) / root_stride * root_stride;
    a.root_end = hgedge[ng];
    a.root_stride = root_stride;
    a.nroots = a.root_begin < a.root_end ? (a.root_end - a.root_begin + root_stride - 1) / root_stride : 0;
    if (np_pat > 0) {
        a.plans = d_plans;
        a.np = np_pat;
        a.sums = d_sums;
        a.root_min_outdeg = 0;
        HIP_TRY(rec(dc.evpool.empty() ? nullptr : dc.evpool[0]));
        enum_launch<GM_MODE_PATTERN>(nmax_pat, wpe_pat, blocks_pat, s, a);
        HIP_TRY(hipGetLastError());
        HIP_TRY(rec(dc.evpool.empty() ? nullptr : dc.evpool[1]));
        ev_list.push_back(0);
    }
    // Clique plans: one at a time; on a single graph the local-bitmap engine
    // takes the roots of out-degree <= bm_max first.
    Counters* ctr2 = dc.ctr + 1;
    Counters* ctr_bm = dc.ctr + 2;
    Counters* ctr_bm_total = dc.ctr + 3;
    Counters* ctr_blk = dc.ctr + 4;
    Counters* ctr_blk_total = dc.ctr + 5;
    // Donation is a pattern-rule mechanism; the clique launches run without
    // the queue (the degree orientation and the bitmap engines bound a
    // clique root's work).
    a.queue = nullptr;
    a.queue_cap = 0;
    a.hub_bits = nullptr;
    a.hub_slot = nullptr;
    for (int c = 0; c < np_clq; ++c) {
        const int evl = 2 + 6 * c, evb = evl + 2, evk = evl + 4;
        const gm_plan* p = plans[order[np_pat + c]];
        unsigned long long* cell = d_sums + (int64_t)ng * np_pat + c;  // stride 1 per graph below
        const bool use_bitmap = ng == 1 && !graph_edge_offsets && p->host.clique_chain && bm_max > 0;
        hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, (unsigned long long*)ctr2,
                           (int64_t)(sizeof(Counters) / sizeof(unsigned long long)));
        a.ctr = ctr2;
        a.root_min_outdeg = 0;
        if (use_bitmap) {
            BmArgs b;
            b.row = g->d_row;
            b.col = g->d_col;
            b.nv = g->n;
            b.k = p->host.n;
            b.max_d = bm_max;
            b.rows_scratch = nullptr;
            b.order = g->d_order;
            b.n_big = g->n_big;
            b.n_active = g->n_active;
            b.out = cell;
            b.ctr = ctr_bm;
            b.stats = stats_on;
            size_t big_dyn = 0;
            b.big_min_d = (t && t->bitmap_block_degree > 0) ? t->bitmap_block_degree : kBigMinD;
            b.big_cap_words = 0;
            if (!(t && t->bitmap_block_degree < 0)) {
                HIP_TRY(big_rows_capacity(g->device, &b.big_cap_words, &big_dyn));
                // Rows only as large as this graph's longest bitmap root
                // needs, so that more blocks are resident per CU (a root
                // whose rows exceed the capacity goes to the per-wave kernel).
                const int32_t dmax = std::min<int32_t>(g->max_deg, bm_max);
                const int32_t need = dmax * ((dmax + 63) / 64);
                if (need > 0 && need < b.big_cap_words) {
                    b.big_cap_words = need;
                    big_dyn = (size_t)need * sizeof(unsigned long long);
                }
            }
            // Roots of out-degree up to 64 (one word per row) go to the
            // small-root kernel; the per-wave kernel takes the roots of
            // out-degree above 64 that the block kernel does not, and is
            // launched only when such roots can exist.
            b.small_max_d = std::min(kSmallD, bm_max);
            b.tri = g->tri;
            const int32_t dmax_bm = std::min<int32_t>(g->max_deg, bm_max);
            bool need_wave = false;
            if (dmax_bm > b.small_max_d) {
                if (b.big_cap_words <= 0 || b.big_min_d > b.small_max_d) need_wave = true;
                else if ((int64_t)dmax_bm * ((dmax_bm + 63) / 64) > b.big_cap_words) need_wave = true;
            }
            hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, (unsigned long long*)ctr_bm,
                               (int64_t)(sizeof(Counters) / sizeof(unsigned long long)));
            HIP_TRY(rec(timed ? dc.evpool[evb] : nullptr));
            HIP_TRY(launch_bitmap_small(p->host.n, cus, b, s));
            if (need_wave) {
                if (fixed_ws) return fail(GM_ERR_HIP, "gm_count_static: internal: per-wave bitmap rows needed");
                HIP_TRY(grow(&dc.bm_rows, &dc.bm_bytes,
                             (size_t)bm_blocks(cus) * kBmMaxD * kBmMaxW * sizeof(unsigned long long)));
                b.rows_scratch = (unsigned long long*)dc.bm_rows;
                HIP_TRY(launch_bitmap(p->host.n, bm_blocks(cus), b, s));
            }
            if (b.big_cap_words > 0) {
                // The large roots: their own counters and events, so their share is reported.
                hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, (unsigned long long*)ctr_blk,
                                   (int64_t)(sizeof(Counters) / sizeof(unsigned long long)));
                BmArgs bb = b;
                bb.ctr = ctr_blk;
                HIP_TRY(rec(timed ? dc.evpool[evk] : nullptr));
                HIP_TRY(launch_bitmap_big(p->host.n, b.tri, big_blocks(p->host.n, b.tri, cus, big_dyn), big_dyn, bb, s));
                HIP_TRY(rec(timed ? dc.evpool[evk + 1] : nullptr));
                ev_block.push_back(evk);
                hipLaunchKernelGGL(accumulate_counters_kernel, dim3(1), dim3(1), 0, s, ctr_bm, ctr_blk);
                hipLaunchKernelGGL(accumulate_counters_kernel, dim3(1), dim3(1), 0, s, ctr_blk_total, ctr_blk);
            }
            HIP_TRY(rec(timed ? dc.evpool[evb + 1] : nullptr));
            ev_bitmap.push_back(evb);
            hipLaunchKernelGGL(accumulate_counters_kernel, dim3(1), dim3(1), 0, s, dc.ctr, ctr_bm);
            hipLaunchKernelGGL(accumulate_counters_kernel, dim3(1), dim3(1), 0, s, ctr_bm_total, ctr_bm);
            a.root_min_outdeg = bm_max;
        }
        // Per-graph cells of this plan are ng apart: give the kernel the
        // column as its own 1-plan table.
        a.plans = d_plans + np_pat + c;
        a.np = 1;
        a.sums = cell;
        if (ng > 1) {
            // cells for graph g at cell[g * np_clq]: use a private region instead
            a.sums = d_sums + (int64_t)ng * np_pat + (int64_t)c * ng;
        }
        // With the bitmap engine on and no row longer than its limit (a
        // degree-oriented graph: out-degree 666 on R-MAT 2^20), every root
        // of the list launch fails the root_min_outdeg filter -- no sum, no
        // bytes, no task -- so the launch, a 9.5 ms scan there, is skipped.
        if (!(use_bitmap && g->max_deg <= bm_max)) {
            HIP_TRY(rec(timed ? dc.evpool[evl] : nullptr));
            enum_launch<GM_MODE_CLIQUE>(nmax_clq, 2, blocks_clq, s, a);
            HIP_TRY(hipGetLastError());
            HIP_TRY(rec(timed ? dc.evpool[evl + 1] : nullptr));
            ev_list.push_back(evl);
        }
        hipLaunchKernelGGL(accumulate_counters_kernel, dim3(1), dim3(1), 0, s, dc.ctr, ctr2);
        a.ctr = dc.ctr;
    }
    HIP_TRY(rec(dc.ev1));
    std::vector<unsigned long long> hs(ncells);
    Counters c, cb, ck;
    HIP_TRY(hipMemcpyAsync(hs.data(), d_sums, sizeof(unsigned long long) * ncells, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&c, dc.ctr, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&cb, ctr_bm_total, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&ck, ctr_blk_total, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0;
    if (timed) HIP_TRY(hipEventElapsedTime(&ms, dc.ev0, dc.ev1));
    else ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host0).count();
    double list_ms = 0, bitmap_ms = 0, block_ms = 0;
    if (timed) {
        for (int i : ev_list) {
            float x = 0;
            HIP_TRY(hipEventElapsedTime(&x, dc.evpool[i], dc.evpool[i + 1]));
            list_ms += x;
        }
        for (int i : ev_bitmap) {
            float x = 0;
            HIP_TRY(hipEventElapsedTime(&x, dc.evpool[i], dc.evpool[i + 1]));
            bitmap_ms += x;
        }
        for (int i : ev_block) {
            float x = 0;
            HIP_TRY(hipEventElapsedTime(&x, dc.evpool[i], dc.evpool[i + 1]));
            block_ms += x;
        }
    }
    if (c.overflow) return fail(GM_ERR_INVALID, "gm_count: an intersection exceeded its scratch slot (graph/plan mismatch)");
    if (c.timed_out) {
        // Slices left in the queue keep their ready flags: clear them so the
        // next launch starts from an all-clear queue.
        if (np_pat > 0 && dc.queue) HIP_TRY(hipMemset(dc.queue, 0, sizeof(SliceItem) * kQueueCap));
        return fail(GM_ERR_TIMEOUT, "gm_count: the time limit passed before the count finished");
    }
    // Scatter back to raw_sums[graph * np + plan] in the caller's plan order.
    for (int gi = 0; gi < ng; ++gi) {
        for (int k = 0; k < np_pat; ++k) raw_sums[(int64_t)gi * np + order[k]] = hs[(int64_t)gi * np_pat + k];
        for (int cc = 0; cc < np_clq; ++cc)
            raw_sums[(int64_t)gi * np + order[np_pat + cc]] =
                ng > 1 ? hs[(int64_t)ng * np_pat + (int64_t)cc * ng + gi] : hs[(int64_t)ng * np_pat + cc];
    }
    if (stats) {
        stats->kernel_ms = ms;
        stats->alg_bytes = c.bytes;
        stats->root_tasks = c.tasks;
        stats->grid_waves = (int32_t)nwaves;
        stats->flags_used = flags;
        stats->list_ms = list_ms;
        stats->bitmap_ms = bitmap_ms;
        stats->bitmap_alg_bytes = cb.bytes;
        stats->bitmap_block_ms = block_ms;
        stats->bitmap_block_alg_bytes = ck.bytes;
        stats->hub_vertices = hubs_used;
        fill_queue_stats(c, stats);
    }
    return GM_OK;
}

int gm_count_batch(const gm_graph* g, const int64_t* graph_edge_offsets, int32_t ng, const gm_plan* const* plans,
                   int32_t np, const gm_tuning* t, int32_t flags, void* stream, uint64_t* raw_sums, gm_stats* stats) {
    return count_impl(g, graph_edge_offsets, ng, plans, np, t, flags, stream, raw_sums, stats, nullptr);
}

int gm_count_static(const gm_graph* g, const gm_plan* p, const gm_tuning* t, int32_t flags, void* stream,
                    uint64_t* raw_sum, gm_stats* stats) {
    if (!g || !p || !raw_sum) return fail(GM_ERR_INVALID, "gm_count_static: bad arguments");
    unsigned char* ws = nullptr;
    const int rc = static_workspace(g->device, &ws);
    if (rc != GM_OK) return rc;
    return count_impl(g, nullptr, 1, &p, 1, t, flags, stream, raw_sum, stats, ws);
}

int gm_count_multi(const gm_graph* g, const gm_plan* const* plans, int32_t np, const gm_tuning* t, int32_t flags,
                   void* stream, uint64_t* raw_sums, gm_stats* stats) {
    return gm_count_batch(g, nullptr, 1, plans, np, t, flags, stream, raw_sums, stats);
}

struct gm_batch {
    int device = 0;
    int32_t ng = 0, np = 0;
    int64_t nwaves = 0;
    int stats = 0;
    int wpe = 2;  // waves per SIMD of the pattern-rule build the graph launches
    void* d_plans = nullptr;
    void* d_gedge = nullptr;
    void* d_sums = nullptr;
    void* d_ctr = nullptr;
    void* d_scratch = nullptr;
    SliceItem* d_queue = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

void gm_batch_destroy(gm_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    for (void* q : {b->d_plans, b->d_gedge, b->d_sums, b->d_ctr, b->d_scratch, (void*)b->d_queue})
        if (q) (void)hipFree(q);
    delete b;
}

int gm_batch_create(const gm_graph* g, const int64_t* graph_edge_offsets, int32_t ng, const gm_plan* const* plans,
                    int32_t np, const gm_tuning* t, int32_t flags, gm_batch** out) {
    if (!out) return fail(GM_ERR_INVALID, "gm_batch_create: null output");
    *out = nullptr;
    if (!g || !plans || np <= 0 || ng < 1) return fail(GM_ERR_INVALID, "gm_batch_create: bad arguments");
    if (ng > 1 && !graph_edge_offsets) return fail(GM_ERR_INVALID, "gm_batch_create: edge offsets required");
    if (flags & GM_COUNT_DEEP_DONATION) return fail(GM_ERR_UNSUPPORTED, kDeepDonationOff);
    for (int i = 0; i < np; ++i) {
        if (!plans[i]) return fail(GM_ERR_INVALID, "gm_batch_create: null plan");
        if (plans[i]->host.mode != GM_MODE_PATTERN)
            return fail(GM_ERR_UNSUPPORTED, "gm_batch_create: only pattern-rule plans are batched");
    }
    std::vector<int64_t> hgedge(2, 0);
    if (graph_edge_offsets) {
        hgedge.assign(graph_edge_offsets, graph_edge_offsets + ng + 1);
        if (hgedge[0] < 0 || hgedge[ng] > g->nnz) return fail(GM_ERR_INVALID, "gm_batch_create: edge offsets outside [0, nnz]");
        for (int i = 0; i < ng; ++i)
            if (hgedge[i + 1] < hgedge[i]) return fail(GM_ERR_INVALID, "gm_batch_create: edge offsets not monotone");
    } else {
        hgedge[1] = g->nnz;
    }
    HIP_TRY(hipSetDevice(g->device));
    gm_batch* b = new gm_batch();
    b->device = g->device;
    b->ng = ng;
    b->np = np;
    b->stats = (flags & GM_COUNT_ALG_BYTES) ? 1 : 0;
    auto bail = [&](int code, const std::string& msg) {
        gm_batch_destroy(b);
        return fail(code, msg);
    };
    int nmax = 0;
    for (int i = 0; i < np; ++i) nmax = std::max(nmax, nmax_for(plans[i]->host.n));
    int cus = 0, occ = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device) != hipSuccess ||
        enum_occupancy<GM_MODE_PATTERN>(nmax,
                                        b->wpe = enum_wpe(GM_MODE_PATTERN, (t && t->waves_per_cu > 0) ? t->waves_per_cu : 0,
                                                          sampled_roots(hgedge[0], hgedge[ng],
                                                                        (t && t->root_stride > 1) ? t->root_stride : 1)),
                                        &occ) != hipSuccess)
        return bail(GM_ERR_HIP, "gm_batch_create: device query failed");
    const int blocks = grid_blocks(cus, occ, (t && t->waves_per_cu > 0) ? t->waves_per_cu : 0);
    b->nwaves = (int64_t)blocks * kWavesPerBlock;
    const int32_t cap = std::max(1, g->second_deg), cap_vi = std::max(1, g->max_deg);
    int64_t stride = 1;
    for (int i = 0; i < np; ++i) {
        const DevPlan& h = plans[i]->host;
        stride = std::max<int64_t>(stride, (int64_t)h.P * cap + (int64_t)(h.vinduced ? h.n : 0) * cap_vi);
    }
    const int64_t ncells = (int64_t)ng * np;
    if (hipMalloc(&b->d_plans, sizeof(DevPlan*) * np) != hipSuccess ||
        hipMalloc(&b->d_gedge, sizeof(int64_t) * (ng + 1)) != hipSuccess ||
        hipMalloc(&b->d_sums, sizeof(unsigned long long) * ncells) != hipSuccess ||
        hipMalloc(&b->d_ctr, sizeof(Counters)) != hipSuccess ||
        hipMalloc(&b->d_scratch, (size_t)stride * (size_t)b->nwaves * sizeof(uint32_t)) != hipSuccess)
        return bail(GM_ERR_OOM, "gm_batch_create: device allocation failed");
    std::vector<const DevPlan*> hptrs(np);
    for (int i = 0; i < np; ++i) hptrs[i] = plans[i]->d_plan;
    if (hipMemcpy(b->d_plans, hptrs.data(), sizeof(DevPlan*) * np, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_gedge, hgedge.data(), sizeof(int64_t) * (ng + 1), hipMemcpyHostToDevice) != hipSuccess)
        return bail(GM_ERR_HIP, "gm_batch_create: upload failed");
    KernelArgs a;
    std::memset(&a, 0, sizeof(a));
    a.row = g->d_row;
    a.col = g->d_col;
    a.nv = g->n;
    a.nnz = g->nnz;
    a.plans = (const DevPlan* const*)b->d_plans;
    a.np = np;
    a.gedge = (const int64_t*)b->d_gedge;
    a.ng = ng;
    a.sums = (unsigned long long*)b->d_sums;
    a.tail_off = (t && t->lane_tail < 0) ? 1 : 0;
    a.tail_work = (t && t->lane_tail > 0) ? t->lane_tail : kLaneTailWork;
    a.task_max = (t && t->chunk_edges > 0) ? t->chunk_edges : kWave;
    a.stage_cap = stage_cap_of(t);
    {
        // The hub bitmaps are built before the capture and belong to the graph.
        const int32_t hub_deg = (t && t->hub_degree != 0) ? t->hub_degree : kDefaultHubDegree;
        if (hub_deg > 0 && g->max_deg >= hub_deg) {
            const int rc = ensure_hub_bitmaps(const_cast<gm_graph*>(g), hub_deg, nullptr);
            if (rc != GM_OK) {
                gm_batch_destroy(b);
                return rc;
            }
            if (g->n_hubs > 0) {
                a.hub_bits = g->d_hub_bits;
                a.hub_slot = g->d_hub_slot;
                a.hub_words = g->hub_words;
            }
        }
    }
    if (slice_len_of(t) > 0) {
        if (alloc_queue(&b->d_queue) != hipSuccess) return bail(GM_ERR_OOM, "gm_batch_create: queue allocation failed");
        a.queue = b->d_queue;
        a.queue_cap = kQueueCap;
        a.slice_len = slice_len_of(t);
        a.eager_donation = (flags & GM_COUNT_EAGER_DONATION) ? 1 : 0;
        a.deep_donation = (flags & GM_COUNT_DEEP_DONATION) ? 1 : 0;
    }
    if (t && t->time_limit_ms > 0) {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g->device) != hipSuccess)
            return bail(GM_ERR_HIP, "gm_batch_create: device query failed");
        a.time_limit_ticks = (int64_t)t->time_limit_ms * (int64_t)std::max(khz, 1);
    }
    const int64_t root_stride = (t && t->root_stride > 1) ? t->root_stride : 1;
    a.root_begin = (hgedge[0] + root_stride - 1) / root_stride * root_stride;
    a.root_end = hgedge[ng];
    a.scratch = (uint32_t*)b->d_scratch;
    a.wave_stride = stride;
    a.cap = cap;
    a.cap_vi = cap_vi;
    a.ctr = (Counters*)b->d_ctr;
    a.stats = b->stats;
    a.root_stride = root_stride;
    a.nroots = a.root_begin < a.root_end ? (a.root_end - a.root_begin + root_stride - 1) / root_stride : 0;
    // Capture: zero the cells and counters with kernels (no memset nodes),
    // then the fused launch.
    hipStream_t cap_stream = nullptr;
    if (hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking) != hipSuccess)
        return bail(GM_ERR_HIP, "gm_batch_create: stream creation failed");
    hipError_t e = hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(zero_u64_kernel, dim3(64), dim3(256), 0, cap_stream, (unsigned long long*)b->d_sums, ncells);
        hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, cap_stream, (unsigned long long*)b->d_ctr,
                           (int64_t)(sizeof(Counters) / sizeof(unsigned long long)));
        enum_launch<GM_MODE_PATTERN>(nmax, b->wpe, blocks, cap_stream, a);
        hipError_t e2 = hipStreamEndCapture(cap_stream, &b->graph);
        e = (e2 != hipSuccess) ? e2 : hipGetLastError();
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cap_stream);
    if (e != hipSuccess) return bail(GM_ERR_HIP, std::string("gm_batch_create: graph capture failed: ") + hipGetErrorString(e));
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
        return bail(GM_ERR_HIP, "gm_batch_create: event creation failed");
    *out = b;
    return GM_OK;
}

int gm_batch_run(gm_batch* b, void* stream, uint64_t* raw_sums, gm_stats* stats) {
    if (!b || !raw_sums) return fail(GM_ERR_INVALID, "gm_batch_run: bad arguments");
    HIP_TRY(hipSetDevice(b->device));
    hipStream_t s = (hipStream_t)stream;
    const int64_t ncells = (int64_t)b->ng * b->np;
    HIP_TRY(hipEventRecord(b->ev0, s));
    HIP_TRY(hipGraphLaunch(b->exec, s));
    HIP_TRY(hipEventRecord(b->ev1, s));
    Counters c;
    HIP_TRY(hipMemcpyAsync(raw_sums, b->d_sums, sizeof(unsigned long long) * ncells, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&c, b->d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (c.overflow) return fail(GM_ERR_INVALID, "gm_batch_run: an intersection exceeded its scratch slot");
    if (c.timed_out) {
        // The batch's own queue keeps the ready flags of slices nobody took:
        // clear them so the next replay starts from an all-clear queue.
        if (b->d_queue) HIP_TRY(hipMemset(b->d_queue, 0, sizeof(SliceItem) * kQueueCap));
        return fail(GM_ERR_TIMEOUT, "gm_batch_run: the time limit passed before the count finished");
    }
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (stats) {
        stats->kernel_ms = ms;
        stats->alg_bytes = c.bytes;
        stats->root_tasks = c.tasks;
        stats->grid_waves = (int32_t)b->nwaves;
        stats->flags_used = b->stats;
        stats->list_ms = ms;
        stats->bitmap_ms = 0;
        stats->bitmap_alg_bytes = 0;
        stats->bitmap_block_ms = 0;
        stats->bitmap_block_alg_bytes = 0;
        stats->hub_vertices = 0;
        fill_queue_stats(c, stats);
    }
    return GM_OK;
}

int gm_count(const gm_graph* g, const gm_plan* p, const gm_tuning* t, int32_t flags, void* stream, uint64_t* raw_sum,
             gm_stats* stats) {
    return gm_count_batch(g, nullptr, 1, &p, 1, t, flags, stream, raw_sum, stats);
}

}  // extern "C"
