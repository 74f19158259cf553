// FP32-mixed instantiations of the general path kernel with the predictor-corrector.
#include "lmm_general.hpp"

namespace fm {
void launch_lmm_step_pc_fp32(const StepArgs& a, void* stream) { launch_general<float, true, false>(a, (hipStream_t)stream); }
}
