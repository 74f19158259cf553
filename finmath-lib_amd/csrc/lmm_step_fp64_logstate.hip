// FP64_LOGSTATE: the spot kernel (lmm_spot.hpp), which carries the log-forwards in FP64,
// compiled with fused multiply-add. FP64 arithmetic throughout; the rounding differs from the
// FP64 build, which keeps every multiply and add separate, by a few units in the last place
// per step.
#include "lmm_spot.hpp"

namespace fm {
void launch_lmm_step_fp64_logstate(const StepArgs& a, void* stream) { launch_spot<double, 1>(a, (hipStream_t)stream); }
}
