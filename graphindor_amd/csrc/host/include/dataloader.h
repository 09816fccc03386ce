// .g binary CSR files and complete-graph builders (reference
// include/dataloader.h:37-71, src/dataloader.cpp:21-124, 630-771).
//
// File layout (little endian): a 32-byte header
//   int32 v_cnt; 4 bytes padding; int64 e_cnt; uint32 max_intersection_size;
//   uint32 checksum; int64 tri_cnt
// whose eight 32-bit words (padding included) sum to 0 mod 2^32, then
// int64 vertex[v_cnt + 1] and int32 edge[e_cnt] (sorted neighbour lists).
#pragma once
#include "graph.h"

class DataLoader {
public:
    // Read a .g file; prints "load_graph: ..." messages like the reference and
    // returns false on a missing file, bad header, checksum or short read.
    bool fast_load(Graph*& g, const char* path);
    // K_n built the way the reference builds it for schedule validation.
    bool load_data(Graph*& g, int clique_size);
    bool load_complete(Graph*& g, int clique_size);
    // Write `g` as a .g file with the given max_intersection_size.
    static bool dump(const Graph& g, const char* path, uint32_t max_intersection_size);
};
