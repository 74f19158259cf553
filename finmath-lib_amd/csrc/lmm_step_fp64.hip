// FP64 instantiations of the spot Euler path kernel and of the fused kernel the adjoint's
// forward pass runs. Compiled with -ffp-contract=off: every product and sum is rounded
// separately, as the JVM evaluates the reference.
#include "lmm_step_dispatch.hpp"
#include "lmm_spot.hpp"
#include "lmm_adjoint.hpp"

namespace fm {
// The pricing and simulation passes run the spot kernel (state carried as the log-forward);
// the adjoint's forward pass (checkpoints) runs the fused kernel whose steps the reverse sweep
// recomputes from the checkpointed values.
void launch_lmm_step_fp64(const StepArgs& a, void* stream) {
    if (a.ckpt) launch_fused<double>(a, (hipStream_t)stream);
    else launch_spot<double>(a, (hipStream_t)stream);
}
void launch_lmm_adjoint_fp64(const StepArgs& a, const AdjArgs& g, void* stream) { launch_adjoint<double>(a, g, (hipStream_t)stream); }
}
