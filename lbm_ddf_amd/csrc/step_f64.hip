// step_f64.hip -- FP64 instantiation of the one-pass step kernel.
// Compiled with -ffp-contract=off so that the arithmetic keeps the
// reference's rounding sequence (no fused multiply-adds).
#include "step_kernel.hpp"

namespace lbm {

hipError_t launch_step_f64(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s) {
    return launch_step_impl<PolicyF64>(a, mode, reduce, bx, flags, s);
}

}  // namespace lbm
