// Capacity dispatch of the fused path kernel: the forward vector and the factor
// sums live in registers, so the kernel is compiled for a few capacities and the
// smallest one that holds (n, F) is launched.
#pragma once

#include "lmm_step.hpp"
#include "hip_check.hpp"

namespace fm {

template <int NCAP, int FCAP, typename real, int VAR>
void launch_fused_cap(const StepArgs& a, hipStream_t stream) {
    const int64_t blocks = ((int64_t)a.P + kStepThreads - 1) / kStepThreads;
    hipLaunchKernelGGL((lmm_fused_kernel<NCAP, FCAP, real, VAR>), dim3((unsigned)blocks), dim3(kStepThreads), 0,
                       stream, a, (const real*)a.rec, a.delta, a.times, a.first_alive, a.tenor_time_index, a.products,
                       a.periods, (const real*)a.dw, a.wave_sum, a.wave_m2, a.contrib, (real*)a.paths, a.row_of,
                       a.numeraire, (real*)a.ckpt);
    HIP_CHECK(hipGetLastError());
}

template <int NCAP, typename real, int VAR>
void launch_fused_n(const StepArgs& a, hipStream_t stream) {
    if (a.F <= 1) launch_fused_cap<NCAP, 1, real, VAR>(a, stream);
    else if (a.F <= 3) launch_fused_cap<NCAP, 3, real, VAR>(a, stream);
    else if (a.F <= 5) launch_fused_cap<NCAP, 5, real, VAR>(a, stream);
    else if (a.F <= 10) launch_fused_cap<NCAP, 10, real, VAR>(a, stream);
    else throw FmError(FM_ERR_UNSUPPORTED, "number of factors above 10 is not supported by the fused kernel");
}

template <typename real, int VAR = kVarFunctional>
void launch_fused(const StepArgs& a, hipStream_t stream) {
    if (a.n <= 8) launch_fused_n<8, real, VAR>(a, stream);
    else if (a.n <= 20) launch_fused_n<20, real, VAR>(a, stream);
    else if (a.n <= 40) launch_fused_n<40, real, VAR>(a, stream);
    else if (a.n <= 80) launch_fused_n<80, real, VAR>(a, stream);
    else throw FmError(FM_ERR_UNSUPPORTED, "more than 80 forward rates are not supported by the fused kernel");
}

} // namespace fm
