// Pattern (reference src/pattern.cpp:9-343).
#include "pattern.h"

#include <algorithm>
#include <cstring>

Pattern::Pattern(int _size, bool clique) : adj_mat(new int[_size * _size]()), size(_size) {
    if (clique)
        for (int i = 0; i < size; ++i)
            for (int j = 0; j < i; ++j) add_edge(i, j);
}

// Row-major '0'/'1' string; any '1' at (i,j) adds the undirected edge
// (src/pattern.cpp:22-31).
Pattern::Pattern(int _size, const char* buffer) : adj_mat(new int[_size * _size]()), size(_size) {
    for (int i = 0; i < size; ++i)
        for (int j = 0; j < size; ++j)
            if (buffer[INDEX(i, j, size)] == '1') add_edge(i, j);
}

Pattern::Pattern(const Pattern& p) : adj_mat(new int[p.size * p.size]), size(p.size) {
    std::memcpy(adj_mat, p.adj_mat, sizeof(int) * size * size);
}

Pattern& Pattern::operator=(const Pattern& p) {
    if (this != &p) {
        int* m = new int[p.size * p.size];
        std::memcpy(m, p.adj_mat, sizeof(int) * p.size * p.size);
        delete[] adj_mat;
        adj_mat = m;
        size = p.size;
    }
    return *this;
}

Pattern::~Pattern() { delete[] adj_mat; }

void Pattern::add_edge(int x, int y) { adj_mat[INDEX(x, y, size)] = adj_mat[INDEX(y, x, size)] = 1; }

void Pattern::del_edge(int x, int y) { adj_mat[INDEX(x, y, size)] = adj_mat[INDEX(y, x, size)] = 0; }

// Breadth-first reachability from vertex 0 (src/pattern.cpp:227-252).
bool Pattern::check_connected() const {
    std::vector<int> seen(size, 0), queue;
    queue.push_back(0);
    seen[0] = 1;
    for (size_t h = 0; h < queue.size(); ++h)
        for (int y = 0; y < size; ++y)
            if (adj_mat[INDEX(queue[h], y, size)] && !seen[y]) {
                seen[y] = 1;
                queue.push_back(y);
            }
    return (int)queue.size() == size;
}

std::vector<std::vector<int> > all_permutations(int n) {
    std::vector<std::vector<int> > out;
    std::vector<int> p(n);
    for (int i = 0; i < n; ++i) p[i] = i;
    do {
        out.push_back(p);
    } while (std::next_permutation(p.begin(), p.end()));
    return out;
}

// Edge set of every relabelling v -> perm[v] (src/pattern.cpp:255-285).
void Pattern::count_all_isomorphism(std::set<std::set<int> >& s) const {
    for (const std::vector<int>& perm : all_permutations(size)) {
        std::set<int> edges;
        for (int i = 0; i < size; ++i)
            for (int j = i + 1; j < size; ++j)
                if (adj_mat[INDEX(i, j, size)]) {
                    int a = std::min(perm[i], perm[j]), b = std::max(perm[i], perm[j]);
                    edges.insert(a * size + b);
                }
        s.insert(edges);
    }
}

void Pattern::print() const {
    for (int i = 0; i < size; ++i)
        for (int j = i + 1; j < size; ++j)
            if (adj_mat[INDEX(i, j, size)]) printf("(%d,%d) ", i, j);
    printf("\n");
}

// Kahn topological sort over the directed 0/1 matrix (src/pattern.cpp:287-317).
bool Pattern::is_dag() const {
    std::vector<int> indeg(size, 0), order;
    for (int i = 0; i < size; ++i)
        for (int j = 0; j < size; ++j)
            if (adj_mat[INDEX(i, j, size)] > 0) ++indeg[j];
    for (int i = 0; i < size; ++i)
        if (indeg[i] == 0) order.push_back(i);
    for (size_t h = 0; h < order.size(); ++h)
        for (int j = 0; j < size; ++j)
            if (adj_mat[INDEX(order[h], j, size)] > 0 && --indeg[j] == 0) order.push_back(j);
    return (int)order.size() == size;
}
