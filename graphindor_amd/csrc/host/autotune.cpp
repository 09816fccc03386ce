// Batched-count autotuner (see autotune.h).
#include "autotune.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>

std::string degree_class(const int64_t* row_ptr, int32_t n) {
    int64_t maxd = 0;
    for (int32_t v = 0; v < n; ++v) maxd = std::max<int64_t>(maxd, row_ptr[v + 1] - row_ptr[v]);
    double mean = n > 0 ? (double)row_ptr[n] / n : 0.0;
    int lm = mean > 1 ? (int)std::lround(std::log2(mean)) : 0;
    int lx = maxd > 1 ? (int)std::lround(std::log2((double)maxd)) : 0;
    return "d" + std::to_string(lm) + "m" + std::to_string(lx);
}

namespace {

bool cache_lookup(const char* path, const std::string& key, TuneChoice& out) {
    if (!path) return false;
    std::ifstream in(path);
    std::string line;
    bool found = false;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string k;
        TuneChoice c;
        if (ls >> k >> c.waves_per_cu >> c.chunk_edges >> c.lane_tail >> c.sample_ms && k == key) {
            out = c;
            found = true;  // the last line for a key wins
        }
    }
    out.from_cache = found ? 1 : 0;
    return found;
}

void cache_store(const char* path, const std::string& key, const TuneChoice& c) {
    if (!path) return;
    std::ofstream out(path, std::ios::app);
    out << key << ' ' << c.waves_per_cu << ' ' << c.chunk_edges << ' ' << c.lane_tail << ' ' << c.sample_ms << '\n';
}

}  // namespace

TuneChoice autotune_batch(const gm_graph* g, const int64_t* edge_offsets, int32_t ng, const gm_plan* const* plans,
                          int32_t np, const std::string& key, const char* cache_path, int32_t root_stride) {
    TuneChoice best;
    if (cache_lookup(cache_path, key, best)) return best;
    // Sample: every graph, every (16 * root_stride)-th root edge.
    const int32_t sample_stride = std::max(1, root_stride) * 16;
    const int32_t k = ng;
    std::vector<uint64_t> sums((size_t)k * np);
    const int waves[] = {8, 12, 16};
    const int chunks[] = {16, 64, 256};
    const int tails[] = {0, -1};
    bool have = false;
    for (int w : waves)
        for (int c : chunks)
            for (int tl : tails) {
                gm_tuning t{};
                t.waves_per_cu = w;
                t.chunk_edges = c;
                t.lane_tail = tl;
                t.root_stride = sample_stride;
                gm_stats st{};
                // one warm-up, then the timed run
                if (gm_count_batch(g, edge_offsets, k, plans, np, &t, 0, nullptr, sums.data(), &st) != GM_OK ||
                    gm_count_batch(g, edge_offsets, k, plans, np, &t, 0, nullptr, sums.data(), &st) != GM_OK)
                    throw std::runtime_error(std::string("autotune: ") + gm_last_error());
                ++best.tried;
                if (!have || st.kernel_ms < best.sample_ms) {
                    have = true;
                    best.waves_per_cu = w;
                    best.chunk_edges = c;
                    best.lane_tail = tl;
                    best.sample_ms = st.kernel_ms;
                }
            }
    cache_store(cache_path, key, best);
    return best;
}
