// FP64 instantiations of the fused path kernel. Compiled with -ffp-contract=off:
// every product and sum is rounded separately, as the JVM evaluates the reference.
#include "lmm_step_dispatch.hpp"

namespace fm {
void launch_lmm_step_fp64(const StepArgs& a, void* stream) { launch_fused<double>(a, (hipStream_t)stream); }
}
