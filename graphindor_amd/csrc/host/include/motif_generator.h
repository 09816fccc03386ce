// All connected k-vertex patterns up to isomorphism, in the reference's
// generation order (include/motif_generator.h, src/motif_generator.cpp:5-44).
#pragma once
#include <set>
#include <vector>

#include "pattern.h"

class MotifGenerator {
public:
    explicit MotifGenerator(int input_size) : size(input_size) {}
    std::vector<Pattern> generate();

private:
    int size;
    std::set<std::set<int> > s;
    void generate_all_pattern(std::vector<Pattern>& vec, std::set<int> edge_set, Pattern& p, int depth, int i, int j);
};
