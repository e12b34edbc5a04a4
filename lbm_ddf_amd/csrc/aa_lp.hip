// aa_lp.hip -- 16-bit storage (FP16, BF16) instantiations of the in-place
// (AA) step kernel; FP32 arithmetic with fused multiply-adds allowed.
#include "aa_kernel.hpp"

namespace lbm {

hipError_t launch_step_aa_f16(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s) {
    return launch_aa_impl<PolicyF16>(a, mode, reduce, bx, flags, s);
}

hipError_t launch_step_aa_bf16(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s) {
    return launch_aa_impl<PolicyBF16>(a, mode, reduce, bx, flags, s);
}

}  // namespace lbm
