// FP64 path kernel carrying the log-normal state between steps (the EULER scheme's
// state), with fused multiply-add: FP64 arithmetic throughout, rounding differs from
// the op-order-faithful FP64 kernels by a few units in the last place per step.
#include "lmm_step_dispatch.hpp"

namespace fm {
void launch_lmm_step_fp64_logstate(const StepArgs& a, void* stream) { launch_fused<double, kVarLogState>(a, (hipStream_t)stream); }
}
