// ensemble_kernel.hpp -- the persistent multi-cavity step kernel (see
// ensemble.hip for the design).  Instantiated per precision in ensemble.hip
// (FP64, no FMA contraction) and ensemble_lp.hip (16-bit storage).
#pragma once

#include "step_kernel.hpp"

namespace lbm {

struct EnsCavity {
    StepArgs a[2];  // [0]: current buffers -> other; [1]: other -> current
    DiagArgs d;
    uint32_t tile0, ntiles;
    int32_t active;
    int32_t slot;
};

template <class P, int MODE, bool REDUCE, int BX>
__global__ __launch_bounds__(BX) void ens_step_kernel(const EnsCavity *__restrict__ cav,
                                                      const uint32_t *__restrict__ ctrl, int parity) {
    const uint32_t ncav = ctrl[0], ntiles = ctrl[1];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint32_t lo = 0, hi = ncav;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cav[mid].tile0 <= t) lo = mid;
            else hi = mid;
        }
        const EnsCavity &c = cav[lo];
        if (!c.active) continue;
        const StepArgs &a = c.a[parity];
        const uint32_t lt = t - c.tile0;
        const uint32_t cell = lt * BX + threadIdx.x;
        double q[NQ];
        if constexpr (REDUCE) {
#pragma unroll
            for (int k = 0; k < NQ; ++k) q[k] = 0.0;
        }
        if (cell < a.n) {
            const uint32_t lx = static_cast<uint32_t>(a.lx), ly = static_cast<uint32_t>(a.ly);
            const uint32_t r = cell / lx;
            const int ix = static_cast<int>(cell - r * lx);
            const int iy = static_cast<int>(r % ly);
            const int iz = static_cast<int>(r / ly);
            step_cell<P, MODE, REDUCE, 0>(a, ix, iy, iz, q);
        }
        if constexpr (REDUCE) block_reduce_diag<BX>(q, a.partials + static_cast<uint64_t>(lt) * NQ);
    }
}

template <class P, int BX>
hipError_t launch_bx(const EnsCavity *cav, const uint32_t *ctrl, int parity, bool reduce, int grid,
                     hipStream_t s) {
    if (reduce)
        hipLaunchKernelGGL((ens_step_kernel<P, MODE_FUSED, true, BX>), dim3(grid), dim3(BX), 0, s, cav, ctrl, parity);
    else
        hipLaunchKernelGGL((ens_step_kernel<P, MODE_FUSED, false, BX>), dim3(grid), dim3(BX), 0, s, cav, ctrl, parity);
    return hipGetLastError();
}

template <class P>
hipError_t launch_prec(const EnsCavity *cav, const uint32_t *ctrl, int parity, bool reduce, int bx, int grid,
                       hipStream_t s) {
    switch (bx) {
        case 64: return launch_bx<P, 64>(cav, ctrl, parity, reduce, grid, s);
        case 128: return launch_bx<P, 128>(cav, ctrl, parity, reduce, grid, s);
        case 256: return launch_bx<P, 256>(cav, ctrl, parity, reduce, grid, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lbm
