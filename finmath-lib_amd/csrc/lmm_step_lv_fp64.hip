// FP64 instantiations of the general path kernel with a local-volatility covariance
// wrapper, with and without the predictor-corrector. Compiled with -ffp-contract=off like
// the FP64 kernel.
#include "lmm_general.hpp"

namespace fm {
void launch_lmm_step_lv_fp64(const StepArgs& a, void* stream) {
    if (a.pc) launch_general<double, true, true>(a, (hipStream_t)stream);
    else launch_general<double, false, true>(a, (hipStream_t)stream);
}
}
