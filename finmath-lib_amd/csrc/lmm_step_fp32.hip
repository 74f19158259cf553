// FP32-mixed instantiations of the fused path kernel: the path arithmetic in float
// (fused multiply-add allowed), numeraire and payoff/numeraire in double.
#include "lmm_step_dispatch.hpp"
#include "lmm_spot.hpp"
#include "lmm_adjoint.hpp"

namespace fm {
// pricing and simulation: the spot kernel; the adjoint's forward pass: the fused kernel
void launch_lmm_step_fp32(const StepArgs& a, void* stream) {
    if (a.ckpt) launch_fused<float>(a, (hipStream_t)stream);
    else launch_spot<float>(a, (hipStream_t)stream);
}
void launch_lmm_adjoint_fp32(const StepArgs& a, const AdjArgs& g, void* stream) { launch_adjoint<float>(a, g, (hipStream_t)stream); }
}
