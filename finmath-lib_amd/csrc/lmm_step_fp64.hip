// FP64 instantiations of the spot Euler path kernel and its reverse sweep. Compiled with
// -ffp-contract=off: every product and sum is rounded separately, as the JVM evaluates the
// reference.
#include "lmm_spot.hpp"
#include "lmm_spot_adjoint.hpp"

namespace fm {
// Pricing, simulation and the adjoint's forward pass (with checkpoints) run the spot kernel;
// the reverse sweep recomputes its steps from the checkpointed states.
void launch_lmm_step_fp64(const StepArgs& a, void* stream) {
    launch_spot<double>(a, (hipStream_t)stream);
}
void launch_lmm_adjoint_fp64(const StepArgs& a, const AdjArgs& g, void* stream) { launch_spot_adjoint<double>(a, g, (hipStream_t)stream); }
}
