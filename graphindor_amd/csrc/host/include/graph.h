// Graph: CSR adjacency (reference include/graph.h:12-82).  Counting runs on
// the GPU through the C ABI of include/graphindor_amd.h; the device copy of
// the CSR is uploaded on first use and kept until the graph changes.
#pragma once
#include <cstdint>

#include "schedule_IEP.h"
#include "types.h"
#include "vertex_set.h"

struct gm_graph;

// Result of one device count (timing of the counting region).
struct GpuRunInfo {
    unsigned long long raw_sum = 0;  // before the division by the IEP redundancy
    double kernel_ms = 0;            // HIP-event time of the kernel(s)
    double region_s = 0;             // wall time of launch + sync + readback
    unsigned long long alg_bytes = 0;
};

class Graph {
public:
    v_index_t v_cnt;
    e_index_t e_cnt;
    long long tri_cnt;
    double max_running_time = 60 * 60 * 24;
    v_index_t *edge, *edge_from;
    e_index_t* vertex;

    Graph() : v_cnt(0), e_cnt(0), tri_cnt(0), edge(nullptr), edge_from(nullptr), vertex(nullptr) {}
    ~Graph();
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;

    // Embeddings of the scheduled pattern, divided by the IEP redundancy
    // (src/graph.cpp:305-355 semantics), counted on the GPU.  `clique` is
    // accepted for interface compatibility and, as in the reference, does
    // not change the counting rule.
    long long pattern_matching(const Schedule_IEP& schedule, bool clique = false);

    // Per-motif counts of every connected k-vertex pattern, printed as the
    // reference prints them (src/graph.cpp:948-985).
    void motif_counting(int pattern_size);

    // Exact triangle count (each triangle once).
    long long triangle_counting();

    void build_reverse_edges();

    // Device copy of the CSR (uploaded on first call); library-owned.
    gm_graph* device_graph();
    // Drop the device copy after the CSR arrays were modified.
    void invalidate_device();

private:
    gm_graph* dev_ = nullptr;
};

// Keep only the edges (v, u) with u < v: one direction per undirected edge
// (src/graph.cpp:710-735).
void reduce_edges_for_clique(Graph& g);

// Orient every undirected edge of a symmetric graph from the endpoint with
// the smaller (degree, id) to the other; `g` receives a new graph with half
// the CSR entries (src/graph.cpp:821-859).  Each k-clique is then one chain
// of out-neighbours, and out-degrees stay near sqrt(2|E|) on skewed graphs.
void degree_orientation_init(Graph* original_g, Graph*& g);
// degree_orientation_init with vertices renumbered by (degree, id) rank, so
// that every kept edge runs from a smaller id to a larger one (same counts).
void degree_rank_orientation(Graph* original_g, Graph*& g);
// The symmetric graph with vertices renumbered by (degree, id), ascending or
// descending (the text loader's oriented_type 2 / 1); counts are unchanged.
void degree_relabel(Graph* original_g, Graph*& g, bool ascending);

// Device drivers replacing the file-local ones of gpu/gpu_graph.cu:22-51,
// gpu/gpu_mc.cu:29-59 and gpu/gpu_clique.cu:366-570.  They return the count
// after the division by the IEP redundancy; `info` (optional) receives the
// raw sum and timings.  Failures throw std::runtime_error with the C-ABI
// message.
unsigned long long gpu_pattern_matching(Graph* g, const Schedule_IEP& schedule, GpuRunInfo* info = nullptr);
unsigned long long gpu_clique_matching(Graph* g, const Schedule_IEP& schedule, GpuRunInfo* info = nullptr);
