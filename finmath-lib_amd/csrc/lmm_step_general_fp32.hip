// FP32-mixed instantiations of the general path kernel (terminal or spot measure,
// carried or functional state).
#include "lmm_general.hpp"

namespace fm {
void launch_lmm_step_general_fp32(const StepArgs& a, void* stream) { launch_general<float, false, false>(a, (hipStream_t)stream); }
}
