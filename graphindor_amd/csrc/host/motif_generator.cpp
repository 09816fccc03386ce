// Connected k-vertex patterns up to isomorphism (reference
// src/motif_generator.cpp:5-44): walk the upper-triangle edges in row-major
// order, deciding "use" before "skip", and keep a pattern when it is
// connected and no relabelling of it was kept before.
#include "motif_generator.h"

void MotifGenerator::generate_all_pattern(std::vector<Pattern>& vec, std::set<int> edge_set, Pattern& p, int depth,
                                          int i, int j) {
    if (i == size - 1) {
        if (s.count(edge_set) == 0 && p.check_connected()) {
            vec.push_back(p);
            p.count_all_isomorphism(s);
        }
        return;
    }
    int ni = (j == size - 1) ? i + 1 : i;
    int nj = (j == size - 1) ? i + 2 : j + 1;
    edge_set.insert(i * size + j);
    p.add_edge(i, j);
    generate_all_pattern(vec, edge_set, p, depth + 1, ni, nj);
    p.del_edge(i, j);
    edge_set.erase(i * size + j);
    generate_all_pattern(vec, edge_set, p, depth + 1, ni, nj);
}

std::vector<Pattern> MotifGenerator::generate() {
    s.clear();
    std::vector<Pattern> vec;
    Pattern p(size);
    generate_all_pattern(vec, std::set<int>(), p, 0, 0, 1);
    return vec;
}
