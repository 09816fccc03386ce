// Pattern: an n-vertex undirected (or, for restriction DAGs, directed) graph
// held as a row-major n*n 0/1 matrix.  Same public interface as the reference
// (include/pattern.h:27-49).
#pragma once
#include <cstdio>
#include <set>
#include <vector>

#ifndef INDEX
#define INDEX(x, y, n) ((x) * (n) + (y))
#endif

class Pattern {
public:
    Pattern(int _size, bool clique = false);
    Pattern(int _size, const char* buffer);
    Pattern(const Pattern& p);
    Pattern& operator=(const Pattern& p);
    ~Pattern();

    void add_edge(int x, int y);
    void del_edge(int x, int y);
    inline void add_ordered_edge(int x, int y) { adj_mat[INDEX(x, y, size)] = 1; }
    inline int get_size() const { return size; }
    inline const int* get_adj_mat_ptr() const { return adj_mat; }
    bool check_connected() const;
    // Inserts into `s` the edge set (encoded min*n+max) of every relabelling.
    void count_all_isomorphism(std::set<std::set<int> >& s) const;
    void print() const;
    bool is_dag() const;

private:
    int* adj_mat;
    int size;
};

// All permutations of 0..n-1 in lexicographic order (the order the reference's
// recursive get_full_permutation produces, src/schedule_IEP.cpp:998-1014).
std::vector<std::vector<int> > all_permutations(int n);
