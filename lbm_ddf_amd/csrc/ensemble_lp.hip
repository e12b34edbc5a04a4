// ensemble_lp.hip -- 16-bit storage instantiations of the ensemble kernel
// (FP32 arithmetic, FMA contraction allowed).
#include "ensemble_kernel.hpp"

namespace lbm {

hipError_t launch_ens_lp(int prec, const EnsCavity *cav, const uint32_t *ctrl, const uint32_t *tmap, int parity,
                         bool reduce, bool aa, int bx, int flags, int grid, hipStream_t s) {
    if (prec == LBM_PREC_FP16) return launch_prec<PolicyF16>(cav, ctrl, tmap, parity, reduce, aa, bx, flags, grid, s);
    if (prec == LBM_PREC_BF16) return launch_prec<PolicyBF16>(cav, ctrl, tmap, parity, reduce, aa, bx, flags, grid, s);
    return hipErrorInvalidValue;
}

}  // namespace lbm
