// .g reader/writer and complete-graph builders (reference
// src/dataloader.cpp:21-124, 630-771).
#include "dataloader.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

#pragma pack(push, 1)
struct GHeader {
    int32_t v_cnt;
    uint32_t pad;
    int64_t e_cnt;
    uint32_t max_intersection_size;
    uint32_t checksum;
    int64_t tri_cnt;
};
#pragma pack(pop)
static_assert(sizeof(GHeader) == 32, "the .g header is 32 bytes");

uint32_t word_sum(const GHeader& h) {
    uint32_t w[8];
    std::memcpy(w, &h, sizeof(w));
    uint32_t s = 0;
    for (uint32_t x : w) s += x;
    return s;
}

}  // namespace

bool DataLoader::fast_load(Graph*& g, const char* path) {
    FILE* fp = std::fopen(path, "rb");
    if (!fp) {
        printf("load_graph: cannot open file %s.\n", path);
        return false;
    }
    GHeader h;
    bool ok = std::fread(&h, sizeof(h), 1, fp) == 1;
    if (!ok) printf("load_graph: bad header.\n");
    if (ok && word_sum(h) != 0) {
        printf("load_graph: checksum != 0. stop.\n");
        ok = false;
    }
    if (ok && (h.v_cnt < 0 || h.e_cnt < 0)) {
        printf("load_graph: bad header.\n");
        ok = false;
    }
    Graph* ng = nullptr;
    if (ok) {
        ng = new Graph();
        ng->v_cnt = h.v_cnt;
        ng->e_cnt = h.e_cnt;
        ng->tri_cnt = h.tri_cnt;
        VertexSet::max_intersection_size = (int)h.max_intersection_size;
        ng->vertex = new e_index_t[(size_t)h.v_cnt + 1];
        ng->edge = new v_index_t[h.e_cnt > 0 ? (size_t)h.e_cnt : 1];
        if (std::fread(ng->vertex, sizeof(e_index_t), (size_t)h.v_cnt + 1, fp) != (size_t)h.v_cnt + 1) {
            printf("load_graph: failed to load vertexes.\n");
            ok = false;
        } else if (std::fread(ng->edge, sizeof(v_index_t), (size_t)h.e_cnt, fp) != (size_t)h.e_cnt) {
            printf("load_graph: failed to load edges.\n");
            ok = false;
        }
    }
    std::fclose(fp);
    if (!ok) {
        delete ng;
        return false;
    }
    printf("load_graph: %u vertexes, %lu edges\n", (unsigned)ng->v_cnt, (unsigned long)ng->e_cnt);
    g = ng;
    return true;
}

bool DataLoader::dump(const Graph& g, const char* path, uint32_t max_intersection_size) {
    GHeader h;
    std::memset(&h, 0, sizeof(h));
    h.v_cnt = g.v_cnt;
    h.e_cnt = g.e_cnt;
    h.max_intersection_size = max_intersection_size;
    h.tri_cnt = g.tri_cnt;
    h.checksum = 0u - word_sum(h);
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return false;
    bool ok = std::fwrite(&h, sizeof(h), 1, fp) == 1 &&
              std::fwrite(g.vertex, sizeof(e_index_t), (size_t)g.v_cnt + 1, fp) == (size_t)g.v_cnt + 1 &&
              std::fwrite(g.edge, sizeof(v_index_t), (size_t)g.e_cnt, fp) == (size_t)g.e_cnt;
    return std::fclose(fp) == 0 && ok;
}

// K_n with sorted neighbour lists; raises the process-wide intersection bound
// to n-1 like the reference (src/dataloader.cpp:661-662).
bool DataLoader::load_complete(Graph*& g, int n) {
    if (n < 2) return false;
    g = new Graph();
    g->v_cnt = n;
    g->e_cnt = (e_index_t)n * (n - 1);
    g->vertex = new e_index_t[n + 1];
    g->edge = new v_index_t[g->e_cnt];
    e_index_t pos = 0;
    for (int v = 0; v < n; ++v) {
        g->vertex[v] = pos;
        for (int u = 0; u < n; ++u)
            if (u != v) g->edge[pos++] = u;
    }
    g->vertex[n] = pos;
    VertexSet::max_intersection_size = std::max(VertexSet::max_intersection_size, n - 1);
    return true;
}

bool DataLoader::load_data(Graph*& g, int clique_size) { return load_complete(g, clique_size); }
