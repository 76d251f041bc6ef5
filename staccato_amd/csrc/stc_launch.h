// stc_launch.h -- host entry points of the compiled sweep kernels.
//
// Every (factor type, rows per lane, panel kind) instantiation of sweep_kernel
// lives in its own translation unit (stc_kern_inst.hip compiled once per
// triple), so the library builds in parallel.  The host API reaches them only through these
// two functions.
#pragma once

#include "stc_internal.h"

namespace stc {

// Rows per lane the kernels are instantiated for: ceil(n_max / 64) rounded up
// to one of these.
constexpr int kMaxrList[] = {1, 2, 4, 8, 16};

// Device function pointer of sweep_kernel<T, maxr, wide> (T = float when
// refine, else double), for hipFuncSetAttribute / occupancy queries.  nullptr
// when the triple is not built.
void* sweep_kernel_fn(int refine, int maxr, int wide);

// Launch sweep_kernel<T, maxr, wide> with `grid` workgroups and `lds` bytes of
// dynamic LDS on `st`.  Returns hipSuccess or the launch error.
hipError_t sweep_kernel_launch(int refine, int maxr, int wide, int grid, int lds, hipStream_t st,
                               const KParams& kp);

}  // namespace stc
