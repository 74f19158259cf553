// FP64 instantiations of the general path kernel with the predictor-corrector.
// Compiled with -ffp-contract=off like the FP64 kernel.
#include "lmm_general.hpp"

namespace fm {
void launch_lmm_step_pc_fp64(const StepArgs& a, void* stream) { launch_general<double, true, false>(a, (hipStream_t)stream); }
}
