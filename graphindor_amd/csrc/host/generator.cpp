// Seeded R-MAT / Erdos-Renyi generators and exact triangle counting.
#include "generator.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <parallel/algorithm>
#include <vector>

namespace {

inline uint64_t splitmix_next(uint64_t& state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Independent stream for (seed, index).
inline uint64_t stream_seed(uint64_t seed, uint64_t index) {
    uint64_t s = seed ^ 0x6A09E667F3BCC909ULL;
    uint64_t a = splitmix_next(s);
    uint64_t t = a ^ (index * 0xD1B54A32D192ED03ULL);
    return splitmix_next(t);
}

// Uniform double in [0, 1) from the top 53 bits.
inline double unit(uint64_t& state) { return (splitmix_next(state) >> 11) * (1.0 / 9007199254740992.0); }

// Build a symmetric CSR from undirected pairs (u != v); duplicates removed.
Graph* build_csr(int32_t n, std::vector<std::pair<int32_t, int32_t> >& pairs) {
    std::vector<int64_t> key;
    key.reserve(pairs.size() * 2);
    for (const auto& p : pairs) {
        if (p.first == p.second) continue;
        key.push_back(((int64_t)p.first << 32) | (uint32_t)p.second);
        key.push_back(((int64_t)p.second << 32) | (uint32_t)p.first);
    }
    std::vector<std::pair<int32_t, int32_t> >().swap(pairs);
    // libstdc++'s parallel sort over the OpenMP threads; the keys are
    // integers, so the result is the same sequence as a serial sort's.
    __gnu_parallel::sort(key.begin(), key.end());
    key.erase(std::unique(key.begin(), key.end()), key.end());
    Graph* g = new Graph();
    g->v_cnt = n;
    g->e_cnt = (e_index_t)key.size();
    g->vertex = new e_index_t[(size_t)n + 1];
    g->edge = new v_index_t[key.empty() ? 1 : key.size()];
    std::fill(g->vertex, g->vertex + n + 1, 0);
    for (size_t i = 0; i < key.size(); ++i) {
        ++g->vertex[(key[i] >> 32) + 1];
        g->edge[i] = (int32_t)(key[i] & 0xffffffffLL);
    }
    for (int32_t v = 0; v < n; ++v) g->vertex[v + 1] += g->vertex[v];
    return g;
}

}  // namespace

Graph* generate_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, bool permute,
                     int32_t* max_isect_out) {
    const int32_t n = (int32_t)1 << scale;
    const int64_t m = (int64_t)edge_factor << scale;
    const double ab = a + b, abc = a + b + c;
    std::vector<std::pair<int32_t, int32_t> > pairs((size_t)m);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
        uint64_t st = stream_seed(seed, (uint64_t)i);
        int32_t u = 0, v = 0;
        for (int l = 0; l < scale; ++l) {
            double r = unit(st);
            int bu = 0, bv = 0;
            if (r < a) {
            } else if (r < ab) {
                bv = 1;
            } else if (r < abc) {
                bu = 1;
            } else {
                bu = bv = 1;
            }
            u = (u << 1) | bu;
            v = (v << 1) | bv;
        }
        pairs[i] = std::make_pair(u, v);
    }
    if (permute) {
        std::vector<int32_t> perm(n);
        std::iota(perm.begin(), perm.end(), 0);
        uint64_t st = stream_seed(seed, ~0ULL);
        for (int32_t i = n - 1; i > 0; --i) {
            int32_t j = (int32_t)(splitmix_next(st) % (uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < m; ++i) pairs[i] = std::make_pair(perm[pairs[i].first], perm[pairs[i].second]);
    }
    Graph* g = build_csr(n, pairs);
    g->tri_cnt = count_triangles(*g);
    if (max_isect_out) *max_isect_out = second_largest_degree(*g);
    return g;
}

Graph* generate_erdos_renyi(int32_t n, double p, uint64_t seed, int32_t* max_isect_out) {
    std::vector<std::vector<std::pair<int32_t, int32_t> > > rows(n);
    const double lq = std::log1p(-p);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t u = 0; u < n; ++u) {
        uint64_t st = stream_seed(seed, (uint64_t)u);
        int64_t v = u;  // candidates v in (u, n)
        while (true) {
            double r = unit(st);
            int64_t skip = (int64_t)std::floor(std::log1p(-r) / lq);
            v += 1 + skip;
            if (v >= n) break;
            rows[u].push_back(std::make_pair(u, (int32_t)v));
        }
    }
    std::vector<std::pair<int32_t, int32_t> > pairs;
    size_t total = 0;
    for (const auto& r : rows) total += r.size();
    pairs.reserve(total);
    for (auto& r : rows) {
        pairs.insert(pairs.end(), r.begin(), r.end());
        std::vector<std::pair<int32_t, int32_t> >().swap(r);
    }
    Graph* g = build_csr(n, pairs);
    g->tri_cnt = count_triangles(*g);
    if (max_isect_out) *max_isect_out = second_largest_degree(*g);
    return g;
}

// Each triangle once: orient u -> v when (deg u, u) < (deg v, v) and merge.
long long count_triangles(const Graph& g) {
    const int32_t n = g.v_cnt;
    std::vector<int64_t> off((size_t)n + 1, 0);
    auto less = [&](int32_t x, int32_t y) {
        int64_t dx = g.vertex[x + 1] - g.vertex[x], dy = g.vertex[y + 1] - g.vertex[y];
        return dx < dy || (dx == dy && x < y);
    };
    for (int32_t v = 0; v < n; ++v)
        for (e_index_t e = g.vertex[v]; e < g.vertex[v + 1]; ++e)
            if (less(v, g.edge[e])) ++off[v + 1];
    for (int32_t v = 0; v < n; ++v) off[v + 1] += off[v];
    std::vector<int32_t> out((size_t)off[n] + 1);
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t v = 0; v < n; ++v) {
        int64_t p = off[v];
        for (e_index_t e = g.vertex[v]; e < g.vertex[v + 1]; ++e)
            if (less(v, g.edge[e])) out[p++] = g.edge[e];
    }
    long long total = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total)
    for (int32_t v = 0; v < n; ++v) {
        for (int64_t e = off[v]; e < off[v + 1]; ++e) {
            int32_t u = out[e];
            int64_t i = off[v], j = off[u];
            while (i < off[v + 1] && j < off[u + 1]) {
                if (out[i] < out[j]) ++i;
                else if (out[i] > out[j]) ++j;
                else { ++total; ++i; ++j; }
            }
        }
    }
    return total;
}

int32_t second_largest_degree(const Graph& g) {
    int64_t d1 = 0, d2 = 0;
    for (int32_t v = 0; v < g.v_cnt; ++v) {
        int64_t d = g.vertex[v + 1] - g.vertex[v];
        if (d > d1) { d2 = d1; d1 = d; }
        else if (d > d2) d2 = d;
    }
    return (int32_t)d2;
}

namespace {

// One graph of the batch, local ids [0, n).
void batch_member(uint64_t seed, double log2_min, double log2_max, double deg_min, double deg_max,
                  std::vector<std::pair<int32_t, int32_t> >& pairs, int32_t& n_out) {
    uint64_t st = stream_seed(seed, 0x5eed5eedULL);
    const bool rmat = (splitmix_next(st) & 1) != 0;
    const double lg = log2_min + unit(st) * (log2_max - log2_min);
    const int32_t n = std::max<int32_t>(2, (int32_t)std::llround(std::pow(2.0, lg)));
    const double d = deg_min + unit(st) * (deg_max - deg_min);
    const int64_t m = (int64_t)std::llround(d * n / 2.0);
    n_out = n;
    pairs.clear();
    if (rmat) {
        int scale = 1;
        while ((1LL << scale) < n) ++scale;
        const double a = 0.57, ab = 0.76, abc = 0.95;
        for (int64_t i = 0; i < m; ++i) {
            int32_t u, v;
            uint64_t es = stream_seed(seed, (uint64_t)i + 1);
            do {
                u = 0;
                v = 0;
                for (int l = 0; l < scale; ++l) {
                    double r = unit(es);
                    int bu = 0, bv = 0;
                    if (r < a) {
                    } else if (r < ab) {
                        bv = 1;
                    } else if (r < abc) {
                        bu = 1;
                    } else {
                        bu = bv = 1;
                    }
                    u = (u << 1) | bu;
                    v = (v << 1) | bv;
                }
            } while (u >= n || v >= n);
            pairs.push_back(std::make_pair(u, v));
        }
    } else {
        const double p = std::min(1.0, d / (double)(n - 1));
        const double lq = std::log1p(-p);
        for (int32_t u = 0; u < n && p > 0; ++u) {
            uint64_t rs = stream_seed(seed, (uint64_t)u + 1);
            int64_t v = u;
            while (true) {
                double r = unit(rs);
                v += 1 + (p >= 1.0 ? 0 : (int64_t)std::floor(std::log1p(-r) / lq));
                if (v >= n) break;
                pairs.push_back(std::make_pair(u, (int32_t)v));
            }
        }
    }
}

}  // namespace

Graph* generate_batch(int count, uint64_t seed_base, double log2_min, double log2_max, double deg_min,
                      double deg_max, int64_t* vertex_off, int64_t* edge_off) {
    std::vector<Graph*> parts(count, nullptr);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < count; ++i) {
        std::vector<std::pair<int32_t, int32_t> > pairs;
        int32_t n = 0;
        batch_member(seed_base + (uint64_t)i, log2_min, log2_max, deg_min, deg_max, pairs, n);
        parts[i] = build_csr(n, pairs);
    }
    int64_t nv = 0, ne = 0;
    for (int i = 0; i < count; ++i) {
        vertex_off[i] = nv;
        edge_off[i] = ne;
        nv += parts[i]->v_cnt;
        ne += parts[i]->e_cnt;
    }
    vertex_off[count] = nv;
    edge_off[count] = ne;
    Graph* g = new Graph();
    g->v_cnt = (v_index_t)nv;
    g->e_cnt = ne;
    g->vertex = new e_index_t[(size_t)nv + 1];
    g->edge = new v_index_t[ne > 0 ? (size_t)ne : 1];
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < count; ++i) {
        const Graph* p = parts[i];
        for (v_index_t v = 0; v < p->v_cnt; ++v) g->vertex[vertex_off[i] + v] = edge_off[i] + p->vertex[v];
        for (e_index_t e = 0; e < p->e_cnt; ++e) g->edge[edge_off[i] + e] = (v_index_t)(vertex_off[i] + p->edge[e]);
    }
    g->vertex[nv] = ne;
    long long tri = 0;
    for (int i = 0; i < count; ++i) {
        tri += count_triangles(*parts[i]);
        delete parts[i];
    }
    g->tri_cnt = tri;
    return g;
}
