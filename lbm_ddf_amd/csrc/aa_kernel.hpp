// aa_kernel.hpp -- the step kernel for one population buffer per
// distribution, updated in place (AA pattern).
//
// The pull A/B scheme keeps two copies of f and g.  Here one copy holds them,
// and two layouts of it alternate.  Slot (i, x) is population i of cell x
// (see pop_layout).
//
//   N ("natural"):   slot (i, x) holds f_i(x) after streaming -- what the
//                    reference stores after src/main.cu:41.
//   E ("exchanged"): slot (opp(j), x) holds the post-collision f*_j(x) when
//                    x + c_j lies inside the domain; when it does not, the
//                    slot already holds the wall value f_opp(j)(x) that the
//                    next streaming would produce there.
//
// An even pass reads N at the cell's own slots, collides, and writes E back
// to the same slots.  An odd pass gathers f_i(x) from slot (opp(i), x - c_i),
// or from (i, x) where x - c_i lies outside, collides, and writes f*_j(x) to
// slot (j, x + c_j), or the wall value to (opp(j), x) where x + c_j lies
// outside -- which is layout N.  In both passes a cell reads and writes the
// same set of slots and the sets of different cells are disjoint, so the
// update is race-free in place.
//
// The wall rules (src/streaming.cu:3-251) are applied by the writer: the
// value streamed into (opp(j), x) from outside is a function of the cell's
// own f*_j(x) and, for the isothermal walls, of rho' of the step just
// completed -- both known when the cell writes.  The arithmetic is the one of
// the A/B kernel, so FP64 results are the same bit for bit; no wall-rho side
// array is needed.
//
// 16-bit storage encodes g with the reference temperature of the writer's
// row, as the A/B scheme does, so the stored deviations stay small: a slot
// holding a value streamed from inside the domain decodes with its source
// row, a wall value with the cell's own row (writer_row in tref.hpp); both
// passes and layout N of the reference state follow that rule.
#pragma once

#include "step_kernel.hpp"

namespace lbm {

template <class P, int MODE, bool REDUCE, int FLAGS>
__device__ __forceinline__ void aa_cell(const StepArgs &a, int ix, int iy, int iz, double (&q)[NQ]) {
    constexpr bool NTS = (FLAGS & kNtStore) != 0;
    constexpr bool NTL = (FLAGS & kNtLoad) != 0;
    constexpr bool O32 = (FLAGS & kOff32) != 0;
    constexpr bool EX = P::exact;
    using S = typename P::S;
    using T = typename P::T;
    using L = typename P::L;
    const int lx = a.lx, ly = a.ly, lz = a.lz;
    const uint64_t ps = a.ps;  // population stride within a row
    const uint32_t idx = static_cast<uint32_t>((static_cast<uint64_t>(iz) * ly + iy) * lx + ix);
    // this cell's slot of population 0 (see pop_layout)
    const uint64_t at = (static_cast<uint64_t>(iz) * ly + iy) * a.rs + ix;
    // one buffer per distribution, read and written in place (no __restrict__)
    S *fs = static_cast<S *>(a.fb);
    S *gs = static_cast<S *>(a.gb);
    // kOff32: the same slots as 32-bit byte offsets from the buffers' starts
    constexpr uint32_t SZB = sizeof(S);
    [[maybe_unused]] const uint32_t atb = static_cast<uint32_t>(at) * SZB;
    [[maybe_unused]] const uint32_t psb = static_cast<uint32_t>(ps) * SZB;
    [[maybe_unused]] const uint32_t rsb = static_cast<uint32_t>(a.rs) * SZB;
    [[maybe_unused]] const uint32_t zsb = rsb * static_cast<uint32_t>(ly);
    // a slot given by its byte offset
    [[maybe_unused]] auto fslot = [&](uint32_t off) { return reinterpret_cast<S *>(reinterpret_cast<char *>(fs) + off); };
    [[maybe_unused]] auto gslot = [&](uint32_t off) { return reinterpret_cast<S *>(reinterpret_cast<char *>(gs) + off); };
    const T t0 = static_cast<T>(a.t0);
    // 16-bit g storage: this row's reference temperature and the offsets of
    // the rows that wrote what this cell reads (ignored in FP64).  In
    // registers g is carried as deviations from w_i*T_c (see step_cell).
    const T tref_y = tref_row<T>(a.tref0, a.dtref, iy);
    const T dref_up = tref_row<T>(a.tref0, a.dtref, source_row(iy, 1)) - tref_y;   // streamed with c_y = +1
    const T dref_dn = tref_row<T>(a.tref0, a.dtref, source_row(iy, -1)) - tref_y;  // streamed with c_y = -1
    // neighbour x - c lies inside for c = +1 (lo) / c = -1 (hi); x + c inside
    // for c = +1 is the hi test and for c = -1 the lo test
    const bool xlo = ix > 0, xhi = ix < lx - 1;
    const bool ylo = iy > 0, yhi = iy < ly - 1;
    const bool zlo = iz > 0, zhi = iz < lz - 1;
    L *lag = static_cast<L *>(a.lag);

    T f[Q], g[Q];
    T rho, ux, uy, uz, phi, dsum = T(0);
    {
        S fr[Q], gr[Q];
        bool streamed[Q];  // the value came from the source cell x - c_i (inside the domain)
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            constexpr int cx = CX[i], cy = CY[i], cz = CZ[i], o = OPP[i];
            const bool inx = cx == 0 ? true : (cx > 0 ? xlo : xhi);
            const bool iny = cy == 0 ? true : (cy > 0 ? ylo : yhi);
            const bool inz = cz == 0 ? true : (cz > 0 ? zlo : zhi);
            streamed[i] = i != 0 && inx && iny && inz;
            if constexpr (O32) {
                // c_i in bytes (mod 2^32; the chosen slot lies in the buffer)
                const uint32_t db = static_cast<uint32_t>(cx * static_cast<int>(SZB)) + static_cast<uint32_t>(cy) * rsb +
                                    static_cast<uint32_t>(cz) * zsb;
                const uint32_t own = atb + static_cast<uint32_t>(i) * psb;
                uint32_t off = own;
                if constexpr (MODE == MODE_AA_ODD && i != 0)
                    off = streamed[i] ? atb + static_cast<uint32_t>(o) * psb - db : own;
                fr[i] = ld<NTL && (MODE != MODE_AA_ODD || cx == 0)>(fslot(off));
                gr[i] = ld<NTL && (MODE != MODE_AA_ODD || cx == 0)>(gslot(off));
            } else if constexpr (MODE == MODE_AA_ODD && i != 0) {
                const int64_t d = static_cast<int64_t>(cx) + static_cast<int64_t>(cy + cz * ly) * static_cast<int64_t>(a.rs);
                const uint64_t off = streamed[i] ? at + o * ps - static_cast<uint64_t>(d) : at + i * ps;
                fr[i] = ld<NTL && cx == 0>(&fs[off]);
                gr[i] = ld<NTL && cx == 0>(&gs[off]);
            } else {
                fr[i] = ld<NTL>(&fs[at + i * ps]);
                gr[i] = ld<NTL>(&gs[at + i * ps]);
            }
        });
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            constexpr int cy = CY[i];
            f[i] = P::dec_f(fr[i]);
            if constexpr (EX) {
                g[i] = P::template dec_g<i>(gr[i], tref_y);
            } else {
                // written by the source row when streamed, else by this cell
                g[i] = P::dec_f(gr[i]);
                if constexpr (cy != 0) {
                    constexpr T w = static_cast<T>(weight(i));
                    if (streamed[i]) g[i] = g[i] + w * (cy > 0 ? dref_up : dref_dn);
                }
            }
        });
    }

    if constexpr (MODE == MODE_AA_PROLOGUE) {
        rho = static_cast<T>(ld<false>(&a.rho[idx]));
        ux = static_cast<T>(ld<false>(&a.ux[idx]));
        uy = static_cast<T>(ld<false>(&a.uy[idx]));
        uz = static_cast<T>(ld<false>(&a.uz[idx]));
        const double phid = ld<false>(&a.phi[idx]);
        phi = static_cast<T>(phid);
        if constexpr (!EX) dsum = static_cast<T>(phid - static_cast<double>(tref_y));
    } else {
        // moments, summed in direction order (src/macrovar.cu:17-27)
        T jx = T(0), jy = T(0), jz = T(0), sg = T(0);
        rho = T(0);
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            rho = rho + f[i];
            if constexpr (CX[i] > 0) jx = jx + f[i];
            if constexpr (CX[i] < 0) jx = jx - f[i];
            if constexpr (CY[i] > 0) jy = jy + f[i];
            if constexpr (CY[i] < 0) jy = jy - f[i];
            if constexpr (CZ[i] > 0) jz = jz + f[i];
            if constexpr (CZ[i] < 0) jz = jz - f[i];
            sg = sg + g[i];
        });
        if constexpr (EX) {
            phi = sg;
        } else {
            dsum = sg;
            phi = tref_y + dsum;
        }
        const T fprev = static_cast<T>(a.grav_beta) * P::dec_side(ld<false>(&lag[idx]));
        ux = jx;
        uy = jy + T(0.5) * fprev;
        uz = jz;
        if constexpr (MODE == MODE_AA_MOMENTS) {
            st<false>(&a.rho[idx], static_cast<double>(rho));
            st<false>(&a.ux[idx], static_cast<double>(ux));
            st<false>(&a.uy[idx], static_cast<double>(uy));
            st<false>(&a.uz[idx], static_cast<double>(uz));
            st<false>(&a.phi[idx], EX ? static_cast<double>(phi) : static_cast<double>(tref_y) + static_cast<double>(dsum));
            st<false>(&a.fy[idx], static_cast<double>(fprev));
        }
    }

    if constexpr (MODE != MODE_AA_MOMENTS) {
        // forcing (src/macrovar.cu:45) and collision
        const T dphi = EX ? phi - t0 : (tref_y - t0) + dsum;
        const T F = static_cast<T>(a.grav_beta) * dphi;
        collide_cell<P>(f, g, rho, ux, uy, uz, phi, dsum, F, a);
        const T rw = rho;  // rho' of this step: read by the isothermal walls of the next streaming
        const double tw_d = iy == 0 ? a.t_hot : a.t_cold;
        const T tw = static_cast<T>(tw_d);
        const T twc = EX ? tw : static_cast<T>(tw_d - static_cast<double>(tref_y));  // T_w - T_c
        static_for<Q>([&](auto I) {
            constexpr int j = I;
            constexpr int cx = CX[j], cy = CY[j], cz = CZ[j], o = OPP[j];
            if constexpr (j == 0) {
                if constexpr (O32) {
                    st<NTS>(fslot(atb), P::enc_f(f[0]));
                    st<NTS>(gslot(atb), EX ? P::template enc_g<0>(g[0], tref_y) : P::enc_f(g[0]));
                } else {
                    st<NTS>(&fs[at], P::enc_f(f[0]));
                    st<NTS>(&gs[at], EX ? P::template enc_g<0>(g[0], tref_y) : P::enc_f(g[0]));
                }
            } else {
                const bool dx = cx == 0 ? true : (cx > 0 ? xhi : xlo);
                const bool dy = cy == 0 ? true : (cy > 0 ? yhi : ylo);
                const bool dz = cz == 0 ? true : (cz > 0 ? zhi : zlo);
                const bool inside = dx && dy && dz;
                T gv = g[j];
                if constexpr (cy != 0) {
                    // Direction opp(j) streams into this cell from beyond a y
                    // wall: isothermal anti-bounce-back, unless its source is
                    // also beyond a z wall (z walls are applied last).
                    if (!dy && dz) {
                        constexpr T w2 = static_cast<T>(2.0 * weight(j));
                        if constexpr (EX) gv = -gv + w2 * tw * (T(1) + rw);
                        else gv = -gv + w2 * (twc + tw * rw);
                    }
                }
                // Non-temporal hint only for stores without an x shift: an
                // odd pass's x-shifted store leaves part lines at wave edges,
                // which the L2 completes with the neighbouring wave's part if
                // it keeps them (a streamed part line goes to memory alone).
                constexpr bool NTSJ = NTS && (MODE != MODE_AA_ODD || cx == 0);
                // 16-bit g is stored against this (the writer's) row
                if constexpr (O32) {
                    const uint32_t own = atb + static_cast<uint32_t>(o) * psb;
                    uint32_t dst = own;
                    if constexpr (MODE == MODE_AA_ODD) {
                        const uint32_t db = static_cast<uint32_t>(cx * static_cast<int>(SZB)) +
                                            static_cast<uint32_t>(cy) * rsb + static_cast<uint32_t>(cz) * zsb;
                        dst = inside ? atb + static_cast<uint32_t>(j) * psb + db : own;
                    }
                    st<NTSJ>(fslot(dst), P::enc_f(f[j]));
                    st<NTSJ>(gslot(dst), EX ? P::template enc_g<j>(gv, tref_y) : P::enc_f(gv));
                } else {
                    uint64_t dst = at + o * ps;
                    if constexpr (MODE == MODE_AA_ODD) {
                        const int64_t d = static_cast<int64_t>(cx) +
                                          static_cast<int64_t>(cy + cz * ly) * static_cast<int64_t>(a.rs);
                        if (inside) dst = at + j * ps + static_cast<uint64_t>(d);
                    }
                    st<NTSJ>(&fs[dst], P::enc_f(f[j]));
                    st<NTSJ>(&gs[dst], EX ? P::template enc_g<j>(gv, tref_y) : P::enc_f(gv));
                }
            }
        });
        st<false>(&lag[idx], P::enc_side(dphi));
    }

    if constexpr (REDUCE && MODE != MODE_AA_PROLOGUE) {
        const double phid = EX ? static_cast<double>(phi) : static_cast<double>(tref_y) + static_cast<double>(dsum);
        diag_terms(q, true, iy, ly, static_cast<double>(rho), static_cast<double>(ux), static_cast<double>(uy),
                   static_cast<double>(uz), phid);
    }
}

template <class P, int MODE, bool REDUCE, int BX, int FLAGS>
__global__ __launch_bounds__(BX) void aa_kernel(const StepArgs a) {
    const int ly = a.ly;
    const uint32_t b = (FLAGS & kXcdSwizzle) ? xcd_swizzle(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t row = b / static_cast<uint32_t>(a.ntx);
    const int bxi = static_cast<int>(b - row * static_cast<uint32_t>(a.ntx));
    const int iy = static_cast<int>(row % static_cast<uint32_t>(ly));
    const int iz = static_cast<int>(row / static_cast<uint32_t>(ly));
    const int ix = bxi * BX + static_cast<int>(threadIdx.x);

    double q[NQ];
    if constexpr (REDUCE) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) q[k] = 0.0;
    }
    if (ix < a.lx) aa_cell<P, MODE, REDUCE, FLAGS>(a, ix, iy, iz, q);
    if constexpr (REDUCE && MODE != MODE_AA_PROLOGUE)
        block_reduce_diag<BX>(q, a.partials + static_cast<uint64_t>(blockIdx.x) * NQ);
}

template <class P, int MODE, int BXV, int FL>
hipError_t launch_aa_variant(const StepArgs &a, dim3 grid, hipStream_t s) {
    hipLaunchKernelGGL((aa_kernel<P, MODE, false, BXV, FL>), grid, dim3(BXV), 0, s, a);
    return hipGetLastError();
}

// Launch-shape variants of the even and odd passes: flags 0, 3, 4, 7.
template <class P, int MODE, int BXV>
hipError_t launch_aa_flags(const StepArgs &a, int flags, dim3 grid, hipStream_t s) {
    switch (flags) {
        case 0: return launch_aa_variant<P, MODE, BXV, 0>(a, grid, s);
        case 3: return launch_aa_variant<P, MODE, BXV, 3>(a, grid, s);
        case 4: return launch_aa_variant<P, MODE, BXV, 4>(a, grid, s);
        case 7: return launch_aa_variant<P, MODE, BXV, 7>(a, grid, s);
        default: return hipErrorInvalidValue;
    }
}

template <class P, int BXV>
hipError_t launch_aa_bx(const StepArgs &a, int mode, bool reduce, int flags, dim3 grid, hipStream_t s) {
    const dim3 block(BXV);
    switch (mode) {
        case MODE_AA_EVEN:
            if (!reduce) return launch_aa_flags<P, MODE_AA_EVEN, BXV>(a, flags, grid, s);
            hipLaunchKernelGGL((aa_kernel<P, MODE_AA_EVEN, true, BXV, 0>), grid, block, 0, s, a);
            break;
        case MODE_AA_ODD:
            if (!reduce) return launch_aa_flags<P, MODE_AA_ODD, BXV>(a, flags, grid, s);
            hipLaunchKernelGGL((aa_kernel<P, MODE_AA_ODD, true, BXV, 0>), grid, block, 0, s, a);
            break;
        case MODE_AA_PROLOGUE:
            hipLaunchKernelGGL((aa_kernel<P, MODE_AA_PROLOGUE, false, BXV, 0>), grid, block, 0, s, a);
            break;
        case MODE_AA_MOMENTS:
            if (reduce) hipLaunchKernelGGL((aa_kernel<P, MODE_AA_MOMENTS, true, BXV, 0>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((aa_kernel<P, MODE_AA_MOMENTS, false, BXV, 0>), grid, block, 0, s, a);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class P>
hipError_t launch_aa_impl(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s) {
    const int64_t nb = step_grid_blocks(a.lx, a.ly, a.lz, bx);
    if (nb <= 0 || nb > 0x7fffffffLL) return hipErrorInvalidValue;
    if (a.fb == nullptr || a.gb == nullptr || a.lag == nullptr) return hipErrorInvalidValue;
    const dim3 grid(static_cast<uint32_t>(nb));
    switch (bx) {
        case 64: return launch_aa_bx<P, 64>(a, mode, reduce, flags, grid, s);
        case 128: return launch_aa_bx<P, 128>(a, mode, reduce, flags, grid, s);
        case 256: return launch_aa_bx<P, 256>(a, mode, reduce, flags, grid, s);
        case 512: return launch_aa_bx<P, 512>(a, mode, reduce, flags, grid, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lbm
