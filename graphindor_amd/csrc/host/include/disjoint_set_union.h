// Union-find with path compression (reference include/disjoint_set_union.h,
// src/disjoint_set_union.cpp:3-33).
#pragma once
#include <vector>

class DisjointSetUnion {
public:
    explicit DisjointSetUnion(int n) : size(n), set_size(n), father(n) {}
    void init() {
        for (int i = 0; i < size; ++i) father[i] = i;
        set_size = size;
    }
    void merge(int a, int b) {
        int ra = root(a), rb = root(b);
        if (ra != rb) {
            father[ra] = rb;
            --set_size;
        }
    }
    inline int get_set_size() const { return set_size; }
    inline int get_size() const { return size; }

private:
    int root(int a) { return father[a] == a ? a : (father[a] = root(father[a])); }
    int size, set_size;
    std::vector<int> father;
};
