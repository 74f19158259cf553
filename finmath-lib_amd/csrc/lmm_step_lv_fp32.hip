// FP32-mixed instantiations of the general path kernel with a local-volatility covariance
// wrapper, with and without the predictor-corrector.
#include "lmm_step_dispatch.hpp"

namespace fm {
void launch_lmm_step_lv_fp32(const StepArgs& a, void* stream) {
    if (a.pc) launch_fused<float, kVarGeneralPCLV>(a, (hipStream_t)stream);
    else launch_fused<float, kVarGeneralLV>(a, (hipStream_t)stream);
}
}
