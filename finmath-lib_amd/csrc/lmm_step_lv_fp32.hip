// FP32-mixed instantiations of the general path kernel with a local-volatility covariance
// wrapper, with and without the predictor-corrector.
#include "lmm_general.hpp"

namespace fm {
void launch_lmm_step_lv_fp32(const StepArgs& a, void* stream) {
    if (a.pc) launch_general<float, true, true>(a, (hipStream_t)stream);
    else launch_general<float, false, true>(a, (hipStream_t)stream);
}
}
