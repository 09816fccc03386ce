#!/bin/sh
# Compile the reference GraphSet CPU path from its own sources, where they lie
# under /root/reference, into oracle/_ref/ (test infrastructure; never shipped).
#
# The reference's graph.cpp includes <mpi.h> for its MPI driver
# (src/graph.cpp:14).  The image carries a real MPICH 3.3.2 under /opt/conda;
# its header is used as-is and no MPI library is linked: every MPI-using
# function lives in sections that --gc-sections drops, since nothing on the
# counting path (Graph::pattern_matching, Schedule_IEP, DataLoader) calls them.
# src/graphmpi.cpp is left out for the same reason.  Everything is built with
# hidden visibility so that the shared library can drop those sections too;
# only the shim's gsref_* entry points are exported.
set -e
REF=${GRAPHSET_REF:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/src" ] || [ ! -f /opt/conda/include/mpi.h ]; then
    echo "build_ref: reference sources or mpi.h absent; skipping" >&2
    exit 0
fi
mkdir -p "$OUT/obj"
CXXFLAGS="-O3 -std=c++14 -fopenmp -msse4.2 -mpopcnt -fPIC -ffunction-sections -fdata-sections -fvisibility=hidden -w -I$REF/include -I/opt/conda/include"
SRCS="graph labeled_graph dataloader pattern prefix schedule_IEP vertex_set motif_generator common disjoint_set_union set_operation"
OBJS=""
for s in $SRCS; do
    g++ $CXXFLAGS -c "$REF/src/$s.cpp" -o "$OUT/obj/$s.o"
    OBJS="$OBJS $OUT/obj/$s.o"
done
g++ $CXXFLAGS -c "$HERE/ref_shim.cpp" -o "$OUT/obj/ref_shim.o"
g++ -shared -fopenmp -Wl,--gc-sections -o "$OUT/libgsref.so" $OBJS "$OUT/obj/ref_shim.o"
for t in pm_test clique_test motif_counting_test; do
    g++ $CXXFLAGS "$REF/src/$t.cpp" $OBJS -Wl,--gc-sections -o "$OUT/$t"
done
echo "build_ref: built $OUT/libgsref.so and reference CLIs"
