// Autotuner for batched counts.
//
// A tuning is (resident waves per CU, root edges per work-steal task,
// lane-parallel last level on/off): how many wavefronts share the root
// tasks, how large a task is, and whether the last candidate level runs
// per lane.  For a (pattern set, degree class) the tuner times each
// candidate on a sample of the batch -- every graph, the root edges e with
// e % (16 * root_stride) == 0, root_stride being the sampling of the runs
// it tunes for (1 = full runs) -- keeps the fastest and remembers it in a
// cache file (one line per key: key waves chunk tail ms).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "graphindor_amd.h"

struct TuneChoice {
    int waves_per_cu = 0;  // 0: library default (full occupancy)
    int chunk_edges = 0;   // 0: library default
    int lane_tail = 0;     // 0: on, -1: off
    double sample_ms = 0;  // time of the chosen tuning on the sample
    int from_cache = 0;
    int tried = 0;
};

// Degree class of a CSR: "d<log2 mean degree>m<log2 max degree>".
std::string degree_class(const int64_t* row_ptr, int32_t n);

TuneChoice autotune_batch(const gm_graph* g, const int64_t* edge_offsets, int32_t ng, const gm_plan* const* plans,
                          int32_t np, const std::string& key, const char* cache_path, int32_t root_stride = 1);
