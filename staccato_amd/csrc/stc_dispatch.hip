// stc_dispatch.hip -- selects the compiled sweep-kernel instantiation
// (declared in stc_launch.h; each instantiation is its own object, built from
// stc_kern_inst.hip).
#include "stc_launch.h"

namespace stc {

#define STC_DECL(R, M)                                                   \
    void* sweep_kernel_fn_##R##_##M();                                   \
    hipError_t sweep_kernel_launch_##R##_##M(int, int, hipStream_t, const KParams&);
STC_DECL(0, 1) STC_DECL(0, 2) STC_DECL(0, 4) STC_DECL(0, 8) STC_DECL(0, 16)
STC_DECL(1, 1) STC_DECL(1, 2) STC_DECL(1, 4) STC_DECL(1, 8) STC_DECL(1, 16)
#undef STC_DECL

void* sweep_kernel_fn(int refine, int maxr) {
#define STC_CASE(R, M) \
    if (refine == R && maxr == M) return sweep_kernel_fn_##R##_##M();
    STC_CASE(0, 1) STC_CASE(0, 2) STC_CASE(0, 4) STC_CASE(0, 8) STC_CASE(0, 16)
    STC_CASE(1, 1) STC_CASE(1, 2) STC_CASE(1, 4) STC_CASE(1, 8) STC_CASE(1, 16)
#undef STC_CASE
    return nullptr;
}

hipError_t sweep_kernel_launch(int refine, int maxr, int grid, int lds, hipStream_t st,
                               const KParams& kp) {
#define STC_CASE(R, M) \
    if (refine == R && maxr == M) return sweep_kernel_launch_##R##_##M(grid, lds, st, kp);
    STC_CASE(0, 1) STC_CASE(0, 2) STC_CASE(0, 4) STC_CASE(0, 8) STC_CASE(0, 16)
    STC_CASE(1, 1) STC_CASE(1, 2) STC_CASE(1, 4) STC_CASE(1, 8) STC_CASE(1, 16)
#undef STC_CASE
    return hipErrorInvalidValue;
}

}  // namespace stc
