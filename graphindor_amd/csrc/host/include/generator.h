// Seeded synthetic graphs for the benchmark and parity workloads.
//
// Randomness: SplitMix64 (Steele, Lea & Flood 2014).  Edge sample i of a
// graph with seed s draws from the stream seeded with mix(s, i), so the
// result does not depend on the number of threads.
//
// R-MAT (Chakrabarti, Zhan & Faloutsos 2004): 2^scale vertices,
// edge_factor * 2^scale samples; per sample and per level a quadrant draw
// with probabilities (a, b, c, d), no noise.  Then: symmetrise, drop
// self-loops and duplicates, relabel vertices by a seeded Fisher-Yates
// permutation, sort neighbour lists.
//
// Erdos-Renyi G(n, p): each unordered pair independently with probability
// p, drawn by geometric skipping (Batagelj & Brandes 2005) per row.
//
// Both fill tri_cnt with the exact triangle count and report
// max_intersection_size = the second-largest degree, as the reference's
// loaders compute it (src/dataloader.cpp:296-301).
#pragma once
#include <cstdint>

#include "graph.h"

Graph* generate_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, bool permute,
                     int32_t* max_isect_out);
Graph* generate_erdos_renyi(int32_t n, double p, uint64_t seed, int32_t* max_isect_out);
// A batch of `count` graphs held as one CSR with globally numbered vertices.
// Graph i (seed seed_base + i) is, by a coin flip of its own stream, R-MAT
// ((a,b,c,d) = (0.57,0.19,0.19,0.05) over 2^ceil(log2 n) ids, edges with an
// endpoint >= n drawn again) or Erdos-Renyi G(n, d/(n-1)); n is log-uniform
// in [2^log2_min, 2^log2_max] and the average degree d uniform in
// [deg_min, deg_max].  vertex_off / edge_off receive count+1 offsets.
Graph* generate_batch(int count, uint64_t seed_base, double log2_min, double log2_max, double deg_min,
                      double deg_max, int64_t* vertex_off, int64_t* edge_off);
// Exact triangle count of a symmetric CSR (OpenMP).
long long count_triangles(const Graph& g);
// Second-largest degree (the reference's max_intersection_size).
int32_t second_largest_degree(const Graph& g);
