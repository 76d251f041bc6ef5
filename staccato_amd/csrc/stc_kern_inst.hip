// stc_kern_inst.hip -- one instantiation of the sweep kernel.
//
// Compiled once per (STC_INST_REFINE, STC_INST_MAXR) pair by the build
// (staccato_amd/build.py), each into its own object, so the instantiations
// compile in parallel.  Exposes C++ functions with the pair in their names;
// stc_dispatch.hip selects among them.
#include "stc_kernels.h"

#ifndef STC_INST_REFINE
#error "STC_INST_REFINE (0 or 1) must be defined"
#endif
#ifndef STC_INST_MAXR
#error "STC_INST_MAXR (1, 2, 4, 8 or 16) must be defined"
#endif

#define STC_CAT2(a, b, c) a##b##_##c
#define STC_CAT(a, b, c) STC_CAT2(a, b, c)
#define STC_FN_PTR STC_CAT(sweep_kernel_fn_, STC_INST_REFINE, STC_INST_MAXR)
#define STC_FN_LAUNCH STC_CAT(sweep_kernel_launch_, STC_INST_REFINE, STC_INST_MAXR)

namespace stc {

using InstT = typename std::conditional<STC_INST_REFINE != 0, float, double>::type;

void* STC_FN_PTR() { return reinterpret_cast<void*>(&sweep_kernel<InstT, STC_INST_MAXR>); }

hipError_t STC_FN_LAUNCH(int grid, int lds, hipStream_t st, const KParams& kp) {
    hipLaunchKernelGGL((sweep_kernel<InstT, STC_INST_MAXR>), dim3(grid), dim3(kThreads), lds, st, kp);
    return hipGetLastError();
}

}  // namespace stc
