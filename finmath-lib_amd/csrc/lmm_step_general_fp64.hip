// FP64 instantiations of the general path kernel (terminal or spot measure, carried or
// functional state). Compiled with -ffp-contract=off like the FP64 kernel.
#include "lmm_general.hpp"

namespace fm {
void launch_lmm_step_general_fp64(const StepArgs& a, void* stream) { launch_general<double, false, false>(a, (hipStream_t)stream); }
}
