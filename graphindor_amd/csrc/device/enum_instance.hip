// One instantiation of the enumeration kernel, enumerate_kernel<MODE, NMAX,
// WPE>, with its launcher and occupancy query.  Built once per (GMK_MODE,
// GMK_NMAX, GMK_WPE) as its own object, so the instantiations compile in
// parallel.
#include "engine_kernels.h"

#ifndef GMK_MODE
#error "GMK_MODE (0 pattern, 1 clique) must be defined"
#endif
#ifndef GMK_NMAX
#error "GMK_NMAX (4..7) must be defined"
#endif
#ifndef GMK_WPE
#error "GMK_WPE (waves per SIMD the build must fit: 2 or 4) must be defined"
#endif

#define GMK_PASTE(a, m, n, w) a##m##_##n##_##w
#define GMK_NAME(a, m, n, w) GMK_PASTE(a, m, n, w)

namespace gmk {

hipError_t GMK_NAME(enum_occupancy_, GMK_MODE, GMK_NMAX, GMK_WPE)(int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, enumerate_kernel<GMK_MODE, GMK_NMAX, GMK_WPE>,
                                                        kBlock, 0);
}

void GMK_NAME(enum_launch_, GMK_MODE, GMK_NMAX, GMK_WPE)(int blocks, hipStream_t s, const KernelArgs& a) {
    hipLaunchKernelGGL((enumerate_kernel<GMK_MODE, GMK_NMAX, GMK_WPE>), dim3(blocks), dim3(kBlock), 0, s, a);
}

}  // namespace gmk
