// ensemble_kernel.hpp -- the persistent multi-cavity step kernel (see
// ensemble.hip for the design).  Instantiated per precision in ensemble.hip
// (FP64, no FMA contraction) and ensemble_lp.hip (16-bit storage).
#pragma once

#include "aa_kernel.hpp"
#include "div_magic.hpp"
#include "tile_order.hpp"

namespace lbm {

struct EnsCavity {
    // A/B pattern: [0] current buffers -> other, [1] other -> current.
    // In-place pattern: both the one buffer.
    StepArgs a[2];
    DiagArgs d;
    uint32_t tile0, ntiles;
    int32_t active;
    int32_t slot;
    // In-place pattern: 1 when the cavity's pass at launch parity 0 is the
    // odd one (its buffer holds the exchanged layout), else 0.
    int32_t aa_odd0;
    // Division of a cell index (< 2^31) by lx and by ly as a multiply and a
    // shift (div_magic.hpp).
    uint32_t mlx, mly;
    uint32_t shlx, shly;
};

// One launch of the ensemble step: every resident cavity advances by one
// pass.  AA selects the in-place passes (even or odd per cavity, from
// parity and aa_odd0); otherwise the A/B pass from buffer set `parity`.
template <class P, int MODE, bool REDUCE, int BX, int FLAGS, bool AA>
__global__ __launch_bounds__(BX) void ens_step_kernel(const EnsCavity *__restrict__ cav,
                                                      const uint32_t *__restrict__ ctrl,
                                                      const uint32_t *__restrict__ tmap, int parity) {
    const uint32_t ntiles = ctrl[1];
    const uint32_t cpt = ctrl[2];  // cells per thread: a tile is cpt * BX consecutive cells
    // Tile order (tile_order.hpp): plain, or chunks of neighbouring tiles
    // dealt to the XCDs in turn.  (One contiguous eighth of the tiles per
    // XCD fetches as little but finishes unevenly, because the XCDs then hold
    // different cavities: 16-bit 64-cavity launch 1.79 ms against 1.76 ms in
    // the plain order.)
    const TileOrder order = tile_order((FLAGS & kXcdSwizzle) != 0, blockIdx.x, gridDim.x);
    uint32_t kl = order.first;
    uint64_t t = tile_of(order, kl), t_next = 0;
    // A tile's cavity comes from the tile map, loaded one tile ahead, so
    // that finding it costs no dependent chain of table reads.
    uint32_t ci = t < ntiles ? tmap[t] : 0u, ci_next = 0u;
    for (; t < ntiles; kl += order.step, t = t_next, ci = ci_next) {
        t_next = tile_of(order, kl + order.step);
        ci_next = t_next < ntiles ? tmap[t_next] : 0u;
        const EnsCavity &c = cav[ci];
        if (!c.active) continue;
        const StepArgs &a = c.a[AA ? 0 : parity];
        const bool odd = AA && (parity ^ c.aa_odd0) != 0;
        // the cell update of this cavity's pass
        auto cell_update = [&](int ix, int iy, int iz, double(&qc)[NQ]) {
            constexpr int FL = FLAGS & (kNtStore | kNtLoad | kOff32);
            if constexpr (AA) {
                if (odd) aa_cell<P, MODE_AA_ODD, REDUCE, FL>(a, ix, iy, iz, qc);
                else aa_cell<P, MODE_AA_EVEN, REDUCE, FL>(a, ix, iy, iz, qc);
            } else {
                step_cell<P, MODE, REDUCE, FL>(a, ix, iy, iz, qc);
            }
        };
        const uint32_t lt = t - c.tile0;
        double q[NQ];
        if constexpr (REDUCE) {
#pragma unroll
            for (int k = 0; k < NQ; ++k) q[k] = 0.0;
        }
        for (uint32_t k = 0; k < cpt; ++k) {
            const uint32_t cell = (lt * cpt + k) * BX + threadIdx.x;
            if (cell >= a.n) break;
            const uint32_t lx = static_cast<uint32_t>(a.lx), ly = static_cast<uint32_t>(a.ly);
            const uint32_t r = div_by(cell, c.mlx, c.shlx);
            const uint32_t iz32 = div_by(r, c.mly, c.shly);
            const int ix = static_cast<int>(cell - r * lx);
            const int iy = static_cast<int>(r - iz32 * ly);
            const int iz = static_cast<int>(iz32);
            if constexpr (REDUCE) {
                double qc[NQ];
                cell_update(ix, iy, iz, qc);
#pragma unroll
                for (int j = 0; j < NQ; ++j) q[j] = (j == Q_MAXUSQ) ? fmax(q[j], qc[j]) : q[j] + qc[j];
            } else {
                cell_update(ix, iy, iz, q);
            }
        }
        if constexpr (REDUCE) block_reduce_diag<BX>(q, a.partials + static_cast<uint64_t>(lt) * NQ);
    }
}

template <class P, int BX, int FL, bool AA>
hipError_t launch_fl(const EnsCavity *cav, const uint32_t *ctrl, const uint32_t *tmap, int parity, int grid,
                     hipStream_t s) {
    hipLaunchKernelGGL((ens_step_kernel<P, MODE_FUSED, false, BX, FL, AA>), dim3(grid), dim3(BX), 0, s, cav, ctrl,
                       tmap, parity);
    return hipGetLastError();
}

// Flags of the plain launch: 0, 3 (non-temporal), 4 (XCD order), 7 (both),
// each also with kOff32 (16-bit members whose buffers span under 4 GiB; in
// FP64 it gained nothing in place and cost 2% in A/B).
// The launch that folds the diagnostics uses the XCD order only.
template <class P, int BX, bool AA>
hipError_t launch_bx(const EnsCavity *cav, const uint32_t *ctrl, const uint32_t *tmap, int parity, bool reduce,
                     int flags, int grid, hipStream_t s) {
    if (reduce) {
        hipLaunchKernelGGL((ens_step_kernel<P, MODE_FUSED, true, BX, kXcdSwizzle, AA>), dim3(grid), dim3(BX), 0, s,
                           cav, ctrl, tmap, parity);
        return hipGetLastError();
    }
    if constexpr (!P::exact) {
        switch (flags) {
            case kOff32 | 0: return launch_fl<P, BX, kOff32 | 0, AA>(cav, ctrl, tmap, parity, grid, s);
            case kOff32 | 3: return launch_fl<P, BX, kOff32 | 3, AA>(cav, ctrl, tmap, parity, grid, s);
            case kOff32 | 4: return launch_fl<P, BX, kOff32 | 4, AA>(cav, ctrl, tmap, parity, grid, s);
            case kOff32 | 7: return launch_fl<P, BX, kOff32 | 7, AA>(cav, ctrl, tmap, parity, grid, s);
            default: break;
        }
    }
    switch (flags) {
        case 0: return launch_fl<P, BX, 0, AA>(cav, ctrl, tmap, parity, grid, s);
        case 3: return launch_fl<P, BX, 3, AA>(cav, ctrl, tmap, parity, grid, s);
        case 4: return launch_fl<P, BX, 4, AA>(cav, ctrl, tmap, parity, grid, s);
        case 7: return launch_fl<P, BX, 7, AA>(cav, ctrl, tmap, parity, grid, s);
        default: return hipErrorInvalidValue;
    }
}

template <class P, bool AA>
hipError_t launch_pattern(const EnsCavity *cav, const uint32_t *ctrl, const uint32_t *tmap, int parity, bool reduce,
                          int bx, int flags, int grid, hipStream_t s) {
    switch (bx) {
        case 64: return launch_bx<P, 64, AA>(cav, ctrl, tmap, parity, reduce, flags, grid, s);
        case 128: return launch_bx<P, 128, AA>(cav, ctrl, tmap, parity, reduce, flags, grid, s);
        case 256: return launch_bx<P, 256, AA>(cav, ctrl, tmap, parity, reduce, flags, grid, s);
        default: return hipErrorInvalidValue;
    }
}

template <class P>
hipError_t launch_prec(const EnsCavity *cav, const uint32_t *ctrl, const uint32_t *tmap, int parity, bool reduce,
                       bool aa, int bx, int flags, int grid, hipStream_t s) {
    return aa ? launch_pattern<P, true>(cav, ctrl, tmap, parity, reduce, bx, flags, grid, s)
              : launch_pattern<P, false>(cav, ctrl, tmap, parity, reduce, bx, flags, grid, s);
}

}  // namespace lbm
