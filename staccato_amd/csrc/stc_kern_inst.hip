// stc_kern_inst.hip -- one instantiation of the sweep kernel.
//
// Compiled once per (STC_INST_REFINE, STC_INST_MAXR, STC_INST_WIDE) by the build
// (staccato_amd/build.py), each into its own object, so the instantiations
// compile in parallel.  Exposes C++ functions with the triple in their names;
// stc_dispatch.hip selects among them.
#include "stc_kernels.h"

#ifndef STC_INST_REFINE
#error "STC_INST_REFINE (0 or 1) must be defined"
#endif
#ifndef STC_INST_MAXR
#error "STC_INST_MAXR (1, 2, 4, 8 or 16) must be defined"
#endif
#ifndef STC_INST_WIDE
#error "STC_INST_WIDE (0: register panel, 1: wide panel, 2: wide panel with 3M products, 3: wide panel with the two-level update) must be defined"
#endif

#define STC_CAT2(a, b, c, d) a##b##_##c##_##d
#define STC_CAT(a, b, c, d) STC_CAT2(a, b, c, d)
#define STC_FN_PTR STC_CAT(sweep_kernel_fn_, STC_INST_REFINE, STC_INST_MAXR, STC_INST_WIDE)
#define STC_FN_LAUNCH STC_CAT(sweep_kernel_launch_, STC_INST_REFINE, STC_INST_MAXR, STC_INST_WIDE)

namespace stc {

using InstT = typename std::conditional<STC_INST_REFINE != 0, float, double>::type;

void* STC_FN_PTR() { return reinterpret_cast<void*>(&sweep_kernel<InstT, STC_INST_MAXR, STC_INST_WIDE>); }

hipError_t STC_FN_LAUNCH(int grid, int lds, hipStream_t st, const KParams& kp) {
    hipLaunchKernelGGL((sweep_kernel<InstT, STC_INST_MAXR, STC_INST_WIDE>), dim3(grid), dim3(kThreads), lds,
                       st, kp);
    return hipGetLastError();
}

}  // namespace stc
