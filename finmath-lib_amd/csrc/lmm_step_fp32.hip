// FP32-mixed instantiations of the spot Euler path kernel and its reverse sweep: the path
// arithmetic in float (fused multiply-add allowed), numeraire and payoff/numeraire in double.
#include "lmm_spot.hpp"
#include "lmm_spot_adjoint.hpp"

namespace fm {
// pricing, simulation and the adjoint's forward pass: the spot kernel
void launch_lmm_step_fp32(const StepArgs& a, void* stream) {
    launch_spot<float>(a, (hipStream_t)stream);
}
void launch_lmm_adjoint_fp32(const StepArgs& a, const AdjArgs& g, void* stream) { launch_spot_adjoint<float>(a, g, (hipStream_t)stream); }
}
