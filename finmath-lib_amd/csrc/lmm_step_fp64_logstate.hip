// FP64 path kernel carrying the log-forwards between steps, with fused multiply-add:
// FP64 arithmetic throughout, rounding differs from the op-order-faithful FP64 kernel
// by a few units in the last place per step.
#include "lmm_step_dispatch.hpp"

namespace fm {
void launch_lmm_step_fp64_logstate(const StepArgs& a, void* stream) { launch_fused<double, true>(a, (hipStream_t)stream); }
}
