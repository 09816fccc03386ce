// One instantiation of the enumeration kernel, enumerate_kernel<MODE, NMAX>,
// with its launcher and occupancy query.  Built once per (GMK_MODE,
// GMK_NMAX) pair as its own object, so the eight instantiations compile in
// parallel.
#include "engine_kernels.h"

#ifndef GMK_MODE
#error "GMK_MODE (0 pattern, 1 clique) must be defined"
#endif
#ifndef GMK_NMAX
#error "GMK_NMAX (4..7) must be defined"
#endif

#define GMK_PASTE(a, m, n) a##m##_##n
#define GMK_NAME(a, m, n) GMK_PASTE(a, m, n)

namespace gmk {

hipError_t GMK_NAME(enum_occupancy_, GMK_MODE, GMK_NMAX)(int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, enumerate_kernel<GMK_MODE, GMK_NMAX>, kBlock, 0);
}

void GMK_NAME(enum_launch_, GMK_MODE, GMK_NMAX)(int blocks, hipStream_t s, const KernelArgs& a) {
    hipLaunchKernelGGL((enumerate_kernel<GMK_MODE, GMK_NMAX>), dim3(blocks), dim3(kBlock), 0, s, a);
}

}  // namespace gmk
