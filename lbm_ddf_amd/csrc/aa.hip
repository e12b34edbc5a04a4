// aa.hip -- FP64 instantiation of the in-place (AA) step kernel, and the
// in-place data movements of the AA layouts:
//
//   exchange (E -> N): completes a pending streaming without a second buffer.
//     In layout E, slot (opp(i), y) holds f*_i(y) and slot (i, y + c_i) holds
//     f*_opp(i)(y + c_i) whenever y + c_i is inside; in layout N each belongs
//     in the other's slot.  The map is a product of disjoint swaps, one per
//     (y, i) with i < opp(i) and y + c_i inside; wall values are already in
//     place.
//   flip (reference post-collision state -> E): the streaming of the stage
//     interface (src/streaming.cu:3-251) applied to one distribution in
//     place: each cell moves f*_j to slot opp(j), or the wall value of
//     direction opp(j) where x + c_j is outside; an exchange then finishes it.
//
// Compiled with -ffp-contract=off (FP64 arithmetic in the reference's order).
#include "aa_kernel.hpp"

namespace lbm {

namespace {

constexpr int kFlatBlock = 256;

__device__ __forceinline__ void cell_of(uint64_t idx, int lx, int ly, int &ix, int &iy, int &iz) {
    const uint64_t row = idx / static_cast<uint64_t>(lx);
    ix = static_cast<int>(idx - row * lx);
    iy = static_cast<int>(row % static_cast<uint64_t>(ly));
    iz = static_cast<int>(row / static_cast<uint64_t>(ly));
}

template <typename E>
__global__ void aa_exchange_kernel(E *f, E *g, uint64_t ps, uint64_t rs, uint64_t ncell, int lx, int ly, int lz) {
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= ncell) return;
    int ix, iy, iz;
    cell_of(idx, lx, ly, ix, iy, iz);
    const uint64_t at = (static_cast<uint64_t>(iz) * ly + iy) * rs + ix;
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        constexpr int cx = CX[i], cy = CY[i], cz = CZ[i], o = OPP[i];
        if constexpr (i < o) {
            const int x = ix + cx, y = iy + cy, z = iz + cz;
            if (x >= 0 && x < lx && y >= 0 && y < ly && z >= 0 && z < lz) {
                const uint64_t a = at + o * ps;
                const uint64_t b = at + i * ps + static_cast<uint64_t>(static_cast<int64_t>(cx) +
                                                                       static_cast<int64_t>(cy + cz * ly) * static_cast<int64_t>(rs));
                if (f) {
                    const E t = f[a];
                    f[a] = f[b];
                    f[b] = t;
                }
                if (g) {
                    const E t = g[a];
                    g[a] = g[b];
                    g[b] = t;
                }
            }
        }
    });
}

template <bool SCALAR>
__global__ void aa_flip_kernel(StepArgs a) {
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= a.n) return;
    int ix, iy, iz;
    cell_of(idx, a.lx, a.ly, ix, iy, iz);
    const int ly = a.ly, lz = a.lz;
    const uint64_t ps = a.ps;
    const uint64_t at = (static_cast<uint64_t>(iz) * ly + iy) * a.rs + ix;
    double *p = static_cast<double *>(SCALAR ? a.gb : a.fb);
    double v[Q];
    static_for<Q>([&](auto I) {
        constexpr int j = I;
        v[j] = p[at + j * ps];
    });
    const double density = (iy == 0 || iy == ly - 1) ? a.rho[idx] : 0.0;
    const double tw = iy == 0 ? a.t_hot : a.t_cold;
    static_for<Q>([&](auto I) {
        constexpr int j = I;
        constexpr int cy = CY[j], cz = CZ[j], o = OPP[j];
        double out = v[j];
        if constexpr (SCALAR && cy != 0) {
            const bool dy = cy > 0 ? iy < ly - 1 : iy > 0;
            const bool dz = cz == 0 ? true : (cz > 0 ? iz < lz - 1 : iz > 0);
            // direction opp(j) enters from beyond a y wall (and not also from
            // beyond a z wall, which the reference applies last)
            if (!dy && dz) out = -v[j] + 2.0 * weight(j) * tw * (1.0 + density);
        }
        p[at + o * ps] = out;
    });
}

}  // namespace

hipError_t launch_aa_exchange(int precision, void *f, void *g, uint64_t ps, uint64_t rs, uint64_t ncell, int lx,
                              int ly, int lz, hipStream_t s) {
    const dim3 grid(static_cast<uint32_t>((ncell + kFlatBlock - 1) / kFlatBlock));
    if (precision == LBM_PREC_FP64)
        hipLaunchKernelGGL(aa_exchange_kernel<double>, grid, dim3(kFlatBlock), 0, s, static_cast<double *>(f),
                           static_cast<double *>(g), ps, rs, ncell, lx, ly, lz);
    else
        hipLaunchKernelGGL(aa_exchange_kernel<uint16_t>, grid, dim3(kFlatBlock), 0, s, static_cast<uint16_t *>(f),
                           static_cast<uint16_t *>(g), ps, rs, ncell, lx, ly, lz);
    return hipGetLastError();
}

hipError_t launch_aa_flip(const StepArgs &a, bool scalar, hipStream_t s) {
    const dim3 grid(static_cast<uint32_t>((a.n + kFlatBlock - 1) / kFlatBlock));
    if (scalar) hipLaunchKernelGGL(aa_flip_kernel<true>, grid, dim3(kFlatBlock), 0, s, a);
    else hipLaunchKernelGGL(aa_flip_kernel<false>, grid, dim3(kFlatBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_step_aa_f64(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s) {
    return launch_aa_impl<PolicyF64>(a, mode, reduce, bx, flags, s);
}

}  // namespace lbm
