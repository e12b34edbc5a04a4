// step_lp.hip -- 16-bit storage (FP16, BF16) instantiations of the one-pass
// step kernel; FP32 arithmetic with fused multiply-adds allowed.
#include "step_kernel.hpp"

namespace lbm {

hipError_t launch_step_f16(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s) {
    return launch_step_impl<PolicyF16>(a, mode, reduce, bx, flags, s);
}

hipError_t launch_step_bf16(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s) {
    return launch_step_impl<PolicyBF16>(a, mode, reduce, bx, flags, s);
}

}  // namespace lbm
