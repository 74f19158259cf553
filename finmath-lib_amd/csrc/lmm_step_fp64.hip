// FP64 instantiations of the fused path kernel. Compiled with -ffp-contract=off:
// every product and sum is rounded separately, as the JVM evaluates the reference.
#include "lmm_step_dispatch.hpp"
#include "lmm_adjoint.hpp"

namespace fm {
void launch_lmm_step_fp64(const StepArgs& a, void* stream) { launch_fused<double>(a, (hipStream_t)stream); }
void launch_lmm_adjoint_fp64(const StepArgs& a, const AdjArgs& g, void* stream) { launch_adjoint<double>(a, g, (hipStream_t)stream); }
}
