// stc_dispatch.hip -- selects the compiled sweep-kernel instantiation
// (declared in stc_launch.h; each instantiation is its own object, built from
// stc_kern_inst.hip).
#include "stc_launch.h"

namespace stc {

// (refine, rows per lane, panel kind) triples that are built (panel kind 0:
// register panel, 1: wide LDS panel, 2: wide panel with Gauss 3M products,
// complex64 only, 3: wide panel with the two-level update); the wide-panel kernel is not built where its LDS panel
// cannot hold 16 columns (complex128 beyond n = 512) or where every system
// fits the LDS whole (n <= 64).
#define STC_INSTANCES(X)                                                         \
    X(0, 1, 0) X(0, 2, 0) X(0, 4, 0) X(0, 8, 0) X(0, 16, 0)                      \
    X(1, 1, 0) X(1, 2, 0) X(1, 4, 0) X(1, 8, 0) X(1, 16, 0)                      \
    X(0, 2, 1) X(0, 4, 1) X(0, 8, 1)                                             \
    X(1, 2, 1) X(1, 4, 1) X(1, 8, 1) X(1, 16, 1)                                \
    X(1, 2, 2) X(1, 4, 2) X(1, 8, 2) X(1, 16, 2)                                \
    X(0, 2, 3) X(0, 4, 3) X(0, 8, 3)                                             \
    X(1, 2, 3) X(1, 4, 3) X(1, 8, 3) X(1, 16, 3)

#define STC_DECL(R, M, W)                                                        \
    void* sweep_kernel_fn_##R##_##M##_##W();                                     \
    hipError_t sweep_kernel_launch_##R##_##M##_##W(int, int, hipStream_t, const KParams&);
STC_INSTANCES(STC_DECL)
#undef STC_DECL

void* sweep_kernel_fn(int refine, int maxr, int wide) {
#define STC_CASE(R, M, W) \
    if (refine == R && maxr == M && wide == W) return sweep_kernel_fn_##R##_##M##_##W();
    STC_INSTANCES(STC_CASE)
#undef STC_CASE
    return nullptr;
}

hipError_t sweep_kernel_launch(int refine, int maxr, int wide, int grid, int lds, hipStream_t st,
                               const KParams& kp) {
#define STC_CASE(R, M, W) \
    if (refine == R && maxr == M && wide == W) return sweep_kernel_launch_##R##_##M##_##W(grid, lds, st, kp);
    STC_INSTANCES(STC_CASE)
#undef STC_CASE
    return hipErrorInvalidValue;
}

}  // namespace stc
