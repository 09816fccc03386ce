// VertexSet: only its process-wide bound survives in this build.  The
// reference sizes every materialised intersection by
// VertexSet::max_intersection_size, taken from the .g header
// (src/dataloader.cpp:100, include/vertex_set.h:54); the device engine uses
// the same bound for its per-wave scratch.
#pragma once

class VertexSet {
public:
    static int max_intersection_size;
};
