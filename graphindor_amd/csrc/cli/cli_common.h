// Shared pieces of the gpu_graph / gpu_clique / gpu_mc command lines.
#pragma once
#include <cstdio>

#include "graphindor_amd.h"

// "title: S.UUUUUU s." exactly as the reference's TimeInterval::print
// (include/timeinterval.h:15-21), for a duration measured elsewhere.
inline void print_seconds(const char* title, double s) {
    long sec = (long)s;
    long usec = (long)((s - (double)sec) * 1e6 + 0.5);
    if (usec >= 1000000) {
        sec += 1;
        usec -= 1000000;
    }
    printf("%s: %ld.%06ld s.\n", title, sec, usec);
    fflush(stdout);
}

inline int require_device() {
    int n = gm_device_count();
    if (n < 1) fprintf(stderr, "no HIP device visible\n");
    return n;
}
