// step_kernel.hpp -- the one-pass D3Q27 double-distribution step kernel.
//
// One thread owns one lattice cell.  A fused launch does, for every cell x,
// what the reference does in five kernels and two device copies
// (src/main.cu:27-48):
//
//   gather  f_i(x) = f*_i(x - c_i), or the wall rule        (src/streaming.cu:3-125)
//           g_i(x) = g*_i(x - c_i), or the wall rule        (src/streaming.cu:127-251)
//   moments rho', j, phi; u = j + F(phi of step k-1)/2       (src/macrovar.cu:3-34)
//   forcing F = g*beta*(phi - t0)                            (src/macrovar.cu:36-47)
//   collide f (BGK + buoyancy), g (BGK)                      (src/collision.cu:3-81)
//   store   f*, g* to the other buffer (pull A/B scheme)
//
// plus, optionally, the per-cell terms of the diagnostics (src/output.cu).
//
// Two per-cell side values carry what the reference keeps in full fields:
//   lag[x]      phi - t0 of the previous completed step, because the
//               reference adds half of the force computed from the previous
//               phi to u (src/macrovar.cu:31 with src/main.cu:27);
//   wall_rho[x] rho' of the previous completed step on the planes iy = 0 and
//               iy = LY-1, read by the isothermal anti-bounce-back
//               (src/streaming.cu:206-227).
// Each is read and rewritten only by the thread that owns the cell, so the
// update is in place.
//
// FP64 storage keeps the reference's operation order (divisions by RT, tau
// and tauc included) and is compiled with -ffp-contract=off; the result is
// then the same as the FP64 oracle bit for bit.  16-bit storage keeps
// deviations (f as stored by the reference; g - w_i*T_ref(iy) with T_ref the
// conductive wall-to-wall profile of the cell's row) and computes in FP32.
#pragma once

#include "internal.hpp"
#include "lattice.hpp"
#include "tref.hpp"

namespace lbm {

// ---------------------------------------------------------------------------
// Storage policies
// ---------------------------------------------------------------------------

struct PolicyF64 {
    using S = double;  // stored population
    using T = double;  // arithmetic
    using L = double;  // side values (lag phi, wall rho')
    static constexpr bool exact = true;
    __device__ __forceinline__ static T dec_f(S s) { return s; }
    __device__ __forceinline__ static S enc_f(T t) { return t; }
    __device__ __forceinline__ static T dec_side(L s) { return s; }
    __device__ __forceinline__ static L enc_side(T t) { return t; }
    template <int I>
    __device__ __forceinline__ static T dec_g(S s, T) { return s; }
    template <int I>
    __device__ __forceinline__ static S enc_g(T t, T) { return t; }
};

// FP16: value * 2^15 stored as IEEE binary16 (round to nearest even).  The
// power-of-two scale keeps the small deviations out of the subnormal range.
// dec_g/enc_g take the row's reference temperature: g is stored as
// g_i - w_i*T_ref so that near the isothermal walls, where the Nusselt
// gradients live, the stored values are small deviations, not |T| ~ 0.5.
// The side values (lag phi - t0, bounded by |T_hot - T_cold| / 2, and wall
// rho') use the same scaled binary16.  FP16 storage is accepted only for
// |T_hot|, |T_cold| <= 1: then the stored deviations and side values stay
// below 2 in magnitude while |u| is below the sound speed, and their scaled
// values below the binary16 maximum (65504).
struct PolicyF16 {
    using S = _Float16;
    using T = float;
    using L = _Float16;
    static constexpr bool exact = false;
    static constexpr float kScale = 32768.0f;
    static constexpr float kInv = 1.0f / 32768.0f;
    __device__ __forceinline__ static T dec_f(S s) { return static_cast<float>(s) * kInv; }
    __device__ __forceinline__ static S enc_f(T t) { return static_cast<_Float16>(t * kScale); }
    __device__ __forceinline__ static T dec_side(L s) { return static_cast<float>(s) * kInv; }
    __device__ __forceinline__ static L enc_side(T t) { return static_cast<_Float16>(t * kScale); }
    template <int I>
    __device__ __forceinline__ static T dec_g(S s, T tref) {
        return static_cast<float>(s) * kInv + static_cast<float>(weight(I)) * tref;
    }
    template <int I>
    __device__ __forceinline__ static S enc_g(T t, T tref) {
        return static_cast<_Float16>((t - static_cast<float>(weight(I)) * tref) * kScale);
    }
};

// BF16: the upper 16 bits of the binary32 value, rounded to nearest even.
__device__ __forceinline__ float bf16_bits_to_float(uint16_t b) {
    return __uint_as_float(static_cast<uint32_t>(b) << 16);
}
__device__ __forceinline__ uint16_t float_to_bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu) != 0u) return static_cast<uint16_t>((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

struct PolicyBF16 {
    using S = uint16_t;
    using T = float;
    using L = float;
    static constexpr bool exact = false;
    __device__ __forceinline__ static T dec_f(S s) { return bf16_bits_to_float(s); }
    __device__ __forceinline__ static S enc_f(T t) { return float_to_bf16_bits(t); }
    __device__ __forceinline__ static T dec_side(L s) { return s; }
    __device__ __forceinline__ static L enc_side(T t) { return t; }
    template <int I>
    __device__ __forceinline__ static T dec_g(S s, T tref) {
        return bf16_bits_to_float(s) + static_cast<float>(weight(I)) * tref;
    }
    template <int I>
    __device__ __forceinline__ static S enc_g(T t, T tref) {
        return float_to_bf16_bits(t - static_cast<float>(weight(I)) * tref);
    }
};

// Loads and stores of device memory.  The pointers are cast to the global
// address space: an ensemble kernel reads them from its descriptor table, so
// the compiler cannot tell their address space and would otherwise use flat
// instructions, which also count as scalar-memory operations and make each
// wait for a descriptor field wait for every population load in flight.
template <typename S>
using global_t = __attribute__((address_space(1))) S;
template <bool NT, typename S>
__device__ __forceinline__ S ld(const S *p) {
    const global_t<S> *g = (const global_t<S> *)p;
    if constexpr (NT) return __builtin_nontemporal_load(g);
    else return *g;
}
template <bool NT, typename S>
__device__ __forceinline__ void st(S *p, S v) {
    global_t<S> *g = (global_t<S> *)p;
    if constexpr (NT) __builtin_nontemporal_store(v, g);
    else *g = v;
}

// ---------------------------------------------------------------------------
// Block reduction of the diagnostics (deterministic: fixed shuffle tree, then
// waves summed in order by one thread per quantity).
// ---------------------------------------------------------------------------

template <int BX>
__device__ __forceinline__ void block_reduce_diag(double (&q)[NQ], double *out) {
    static_assert(BX % 64 == 0, "block must be whole waves");
    constexpr int NW = BX / 64;
    __shared__ double red[NW][NQ];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        double v = q[k];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            double o = __shfl_xor(v, off, 64);
            v = (k == Q_MAXUSQ) ? fmax(v, o) : v + o;
        }
        q[k] = v;
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) red[wave][k] = q[k];
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        const int k = threadIdx.x;
        double v = red[0][k];
        for (int w = 1; w < NW; ++w) v = (k == Q_MAXUSQ) ? fmax(v, red[w][k]) : v + red[w][k];
        st<false>(&out[k], v);
    }
    __syncthreads();  // red[][] may be rewritten by the caller's next tile
}

// Per-cell diagnostic terms of a completed step.
__device__ __forceinline__ void diag_terms(double (&q)[NQ], bool active, int iy, int ly, double rho,
                                           double ux, double uy, double uz, double phi) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) q[k] = 0.0;
    if (!active) return;
    const double usq = ux * ux + uy * uy + uz * uz;
    q[Q_UYPHI] = uy * phi;
    q[Q_UX] = ux;
    q[Q_UY] = uy;
    q[Q_UZ] = uz;
    q[Q_PHI] = phi;
    q[Q_USQ] = usq;
    q[Q_RHO] = rho;
    q[Q_B0] = iy == 0 ? phi : 0.0;
    q[Q_B1] = iy == 1 ? phi : 0.0;
    q[Q_B2] = iy == 2 ? phi : 0.0;
    q[Q_T0] = iy == ly - 1 ? phi : 0.0;
    q[Q_T1] = iy == ly - 2 ? phi : 0.0;
    q[Q_T2] = iy == ly - 3 ? phi : 0.0;
    const bool finite = isfinite(rho) && isfinite(usq) && isfinite(phi);
    q[Q_MAXUSQ] = finite ? usq : 0.0;
    q[Q_NONFINITE] = finite ? 0.0 : 1.0;
}

// Sums per-block partials in a fixed order (one block of kFinishThreads) and
// applies the formulas of src/output.cu:228-256 (Nu) and :31-34 (means).
constexpr int kFinishThreads = 256;
// Fixed-order sum of n partial records by one block of kFinishThreads:
// thread t takes records t, t+256, ...; then a fixed tree.  Totals end in
// sh[0][*].
__device__ __forceinline__ void sum_records_block(const double *partials, int64_t nblocks,
                                                  double (&sh)[kFinishThreads][NQ + 1]) {
    const int t = threadIdx.x;
    double acc[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) acc[k] = 0.0;
    for (int64_t j = t; j < nblocks; j += kFinishThreads) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            const double v = ld<false>(&partials[j * NQ + k]);
            acc[k] = (k == Q_MAXUSQ) ? fmax(acc[k], v) : acc[k] + v;
        }
    }
#pragma unroll
    for (int k = 0; k < NQ; ++k) sh[t][k] = acc[k];
    __syncthreads();
    for (int s = kFinishThreads / 2; s >= 1; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int k = 0; k < NQ; ++k)
                sh[t][k] = (k == Q_MAXUSQ) ? fmax(sh[t][k], sh[t + s][k]) : sh[t][k] + sh[t + s][k];
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void diag_finish_block(const double *partials, int64_t nblocks, const DiagArgs &d,
                                                  lbm_diag *dst) {
    __shared__ double sh[kFinishThreads][NQ + 1];
    sum_records_block(partials, nblocks, sh);
    if (threadIdx.x == 0) {
        const double ncell = static_cast<double>(d.lx) * d.ly * d.lz;
        const double nplane = static_cast<double>(d.lx) * d.lz;
        const double *S = sh[0];
        lbm_diag r;
        r.step = d.step;
        const double uyT_avg = S[Q_UYPHI] / ncell;
        r.nu_bulk = 1.0 + uyT_avg * d.ly / (d.diff * d.dT);
        const double dtdy_bottom = ((-3.0 * S[Q_B0] + 4.0 * S[Q_B1] - S[Q_B2]) / 2.0) / nplane;
        const double dtdy_top = ((3.0 * S[Q_T0] - 4.0 * S[Q_T1] + S[Q_T2]) / 2.0) / nplane;
        r.nu_top = -dtdy_top * d.ly / d.dT;
        r.nu_bottom = -dtdy_bottom * d.ly / d.dT;
        r.umean = S[Q_UX] / ncell;
        r.vmean = S[Q_UY] / ncell;
        r.wmean = S[Q_UZ] / ncell;
        r.tmean = S[Q_PHI] / ncell;
        r.kinetic_energy = 0.5 * S[Q_USQ] / ncell;
        r.mass = S[Q_RHO];
        r.max_speed = sqrt(S[Q_MAXUSQ]);
        r.nonfinite = S[Q_NONFINITE];
        *dst = r;
    }
}


// ---------------------------------------------------------------------------
// Collision of one cell, f and g (src/collision.cu:22-32 and :74-81).
//
// FP64: the reference's operation order on full values.  16-bit storage:
// FP32 arithmetic on deviations.  f is a deviation already (the reference
// stores f with rho0 = 1 taken out).  g arrives as g_i - w_i*T_c with T_c the
// cell's row reference temperature (tref_row), and phi is passed split as
// T_c + dsum, dsum = sum of the g deviations.  The collision is then
//   g_i - w_i T_c  <-  d_i - (d_i - w_i (dsum + phi E_i)) / tau_c,
//   E_i = e_i + (e_i^2 - |u|^2/c_s^2) / 2,
// which is the reference's g_i - (g_i - w_i phi (1 + E_i)) / tau_c minus
// w_i T_c: every FP32 operation acts on quantities of the size of the
// deviations, so no rounding of a number of size |T| ~ 0.5 enters the
// stored values (rounding w_i T_c, a constant per row and direction, would
// add a heat source coherent along each row).
// ---------------------------------------------------------------------------

template <class P>
__device__ __forceinline__ void collide_cell(typename P::T (&f)[Q], typename P::T (&g)[Q],
                                             typename P::T rho, typename P::T ux, typename P::T uy,
                                             typename P::T uz, typename P::T phi, typename P::T dsum,
                                             typename P::T F, const StepArgs &a) {
    using T = typename P::T;
    if constexpr (P::exact) {
        // Reference operation order, FP64.
        const double RT = 1.0 / 3.0;
        const double tau = a.tau, tauc = a.tauc;
        const double uv = (ux * ux + uy * uy + uz * uz) / RT;
        const double c1F = a.c1 * F;
        // ((1 - 0.5/tau) * F_y * (c_y - v)) / RT for c_y = -1, 0, +1
        const double pf[3] = {(c1F * (-1.0 - uy)) / RT, (c1F * (0.0 - uy)) / RT, (c1F * (1.0 - uy)) / RT};
        double eu[Q];
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            if constexpr (i == 0) {
                eu[0] = 0.0;
            } else if constexpr (is_canonical(i)) {
                eu[i] = cdotu<i>(ux, uy, uz) / RT;
            } else {
                eu[i] = -eu[OPP[i]];
            }
        });
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            constexpr double w = weight(i);
            const double e = eu[i];
            const double A = 0.5 * (e * e - uv);
            const double feq = w * ((rho + e) + A);
            const double feq0 = w * ((1.0 + e) + A);
            double fc = -(f[i] - feq) / tau;
            fc = fc + pf[CY[i] + 1] * feq0;
            f[i] = f[i] + fc;
            const double geq = (w * phi) * ((1.0 + e) + A);
            g[i] = g[i] + (-(g[i] - geq) / tauc);
        });
        (void)dsum;
    } else {
        // FP32 arithmetic with reciprocals; same model, g as deviations.
        const float omega = static_cast<float>(1.0 / a.tau);
        const float omegac = static_cast<float>(1.0 / a.tauc);
        const float uv = (ux * ux + uy * uy + uz * uz) * 3.0f;
        const float c1F3 = static_cast<float>(a.c1) * F * 3.0f;
        const float pf[3] = {c1F3 * (-1.0f - uy), c1F3 * (0.0f - uy), c1F3 * (1.0f - uy)};
        float eu[Q];
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            if constexpr (i == 0) {
                eu[0] = 0.0f;
            } else if constexpr (is_canonical(i)) {
                eu[i] = cdotu<i>(ux, uy, uz) * 3.0f;
            } else {
                eu[i] = -eu[OPP[i]];
            }
        });
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            constexpr float w = static_cast<float>(weight(i));
            const float e = eu[i];
            const float A = 0.5f * (e * e - uv);
            const float feq = w * ((rho + e) + A);
            const float feq0 = w * ((1.0f + e) + A);
            f[i] = f[i] - (f[i] - feq) * omega + pf[CY[i] + 1] * feq0;
            const float geq_dev = w * (dsum + phi * (e + A));  // g_eq - w T_c
            g[i] = g[i] - (g[i] - geq_dev) * omegac;
        });
    }
    (void)sizeof(T);
}

// ---------------------------------------------------------------------------
// Launch variants.  All variants compute identical values; they differ only in
// cache hints and occupancy, and the autotuner picks among them.
// ---------------------------------------------------------------------------
enum StepFlags : int {
    kNtStore = 1,  // non-temporal stores of the new populations (streamed out once)
    kNtLoad = 2,   // non-temporal loads of populations without an x shift
    kXcdSwizzle = 4,  // XCD-aware block order (see xcd_swizzle)
    kOff32 = 16,   // population addresses as 32-bit byte offsets from the buffers'
                   // starts (ensemble members whose buffers span under 4 GiB)
};

// The dispatcher hands workgroups to the 8 XCDs of an MI355X round-robin
// (block b runs on XCD b % 8), and each XCD has its own L2.  Adjacent
// segments of one x-row share the cache lines that the x-shifted loads
// straddle; without remapping those lines are fetched once per XCD.  This
// bijection on [0, nb) gives each XCD a contiguous range of logical blocks,
// so a row's segments run on one XCD, close together in time.
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t b, uint32_t nb) {
    constexpr uint32_t kXcds = 8;
    const uint32_t q = nb / kXcds, r = nb % kXcds;
    const uint32_t xcd = b % kXcds, k = b / kXcds;
    return xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

// ---------------------------------------------------------------------------
// The step kernel.
// ---------------------------------------------------------------------------

// One lattice cell of a step pass: gather (or plain load in the prologue),
// moments, forcing, collision, stores, side values, and the cell's terms of
// the diagnostics.  Shared by the single-cavity and the ensemble kernels.
//
// With 16-bit storage g is carried as deviations from w_i*T_c, T_c the row
// reference temperature of this cell (see collide_cell): a value pulled
// from the row below or above, stored against that row's temperature, is
// moved to this row's by adding w_i*(T_src - T_c), a binary32 difference of
// two row temperatures that is exact; phi is T_c + (sum of the deviations).
template <class P, int MODE, bool REDUCE, int FLAGS>
__device__ __forceinline__ void step_cell(const StepArgs &a, int ix, int iy, int iz, double (&q)[NQ]) {
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
    // kOff32: the same slots as 32-bit byte offsets
    constexpr uint32_t SZB = sizeof(S);
    [[maybe_unused]] const uint32_t atb = static_cast<uint32_t>(at) * SZB;
    [[maybe_unused]] const uint32_t psb = static_cast<uint32_t>(ps) * SZB;
    [[maybe_unused]] const uint32_t rsb = static_cast<uint32_t>(a.rs) * SZB;
    [[maybe_unused]] const uint32_t zsb = rsb * static_cast<uint32_t>(ly);
    const S *__restrict__ fa = static_cast<const S *>(a.fa);
    const S *__restrict__ ga = static_cast<const S *>(a.ga);
    S *__restrict__ fb = static_cast<S *>(a.fb);
    S *__restrict__ gb = static_cast<S *>(a.gb);
    const T t0 = static_cast<T>(a.t0);
    // 16-bit g storage: this row's reference temperature and the offsets of
    // the source rows' (ignored in FP64)
    const T tref_y = tref_row<T>(a.tref0, a.dtref, iy);
    const T dref_up = tref_row<T>(a.tref0, a.dtref, source_row(iy, 1)) - tref_y;   // pulled with c_y = +1
    const T dref_dn = tref_row<T>(a.tref0, a.dtref, source_row(iy, -1)) - tref_y;  // pulled with c_y = -1

    T f[Q], g[Q];
    // rho', u, phi; dsum = phi - T_c with 16-bit storage
    T rho, ux, uy, uz, phi, dsum = T(0);
    const bool wall_lo = iy == 0, wall_hi = iy == ly - 1;
    const uint64_t wall_idx = (static_cast<uint64_t>(wall_hi ? 1 : 0) * lz + iz) * lx + ix;
    L *__restrict__ wall_rho = static_cast<L *>(a.wall_rho);
    L *__restrict__ lag = static_cast<L *>(a.lag);

    if constexpr (MODE == MODE_PROLOGUE) {
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            f[i] = P::dec_f(ld<NTL>(&fa[at + i * ps]));
            g[i] = EX ? P::template dec_g<i>(ld<NTL>(&ga[at + i * ps]), tref_y) : P::dec_f(ld<NTL>(&ga[at + i * ps]));
        });
        rho = static_cast<T>(ld<false>(&a.rho[idx]));
        ux = static_cast<T>(ld<false>(&a.ux[idx]));
        uy = static_cast<T>(ld<false>(&a.uy[idx]));
        uz = static_cast<T>(ld<false>(&a.uz[idx]));
        const double phid = ld<false>(&a.phi[idx]);
        phi = static_cast<T>(phid);
        if constexpr (!EX) dsum = static_cast<T>(phid - static_cast<double>(tref_y));
    } else {
        // ---- gather (pull) with the wall rules -------------------------
        const bool xlo = ix > 0, xhi = ix < lx - 1;
        const bool ylo = iy > 0, yhi = iy < ly - 1;
        const bool zlo = iz > 0, zhi = iz < lz - 1;
        const bool wall = wall_lo || wall_hi;
        L rw_raw = L(0);
        if (wall) rw_raw = ld<false>(&wall_rho[wall_idx]);
        // All 54 loads are issued before any value is used, so their
        // latencies overlap; a direction whose source lies outside the domain
        // reads the cell's own opposite population instead (half-way
        // bounce-back), selected by address, not by branch.
        S fr[Q], gr[Q];
        bool pulled[Q];
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            constexpr int cx = CX[i], cy = CY[i], cz = CZ[i], o = OPP[i];
            if constexpr (i == 0) {
                fr[0] = ld<NTL>(&fa[at]);
                gr[0] = ld<NTL>(&ga[at]);
                pulled[0] = true;
            } else if constexpr (O32) {
                const bool inx = cx == 0 ? true : (cx > 0 ? xlo : xhi);
                const bool iny = cy == 0 ? true : (cy > 0 ? ylo : yhi);
                const bool inz = cz == 0 ? true : (cz > 0 ? zlo : zhi);
                pulled[i] = inx && iny && inz;
                // byte offsets (mod 2^32; the chosen one lies in the buffer)
                const uint32_t src = atb + (static_cast<uint32_t>(i) * psb -
                                            (static_cast<uint32_t>(cx * SZB) + static_cast<uint32_t>(cy) * rsb +
                                             static_cast<uint32_t>(cz) * zsb));
                const uint32_t own = atb + static_cast<uint32_t>(o) * psb;
                const uint32_t off = pulled[i] ? src : own;
                fr[i] = ld<NTL && cx == 0>(reinterpret_cast<const S *>(reinterpret_cast<const char *>(fa) + off));
                gr[i] = ld<NTL && cx == 0>(reinterpret_cast<const S *>(reinterpret_cast<const char *>(ga) + off));
            } else {
                const bool inx = cx == 0 ? true : (cx > 0 ? xlo : xhi);
                const bool iny = cy == 0 ? true : (cy > 0 ? ylo : yhi);
                const bool inz = cz == 0 ? true : (cz > 0 ? zlo : zhi);
                // slot of population i at the source cell x - c_i
                const uint64_t src = at + i * ps - static_cast<uint64_t>(static_cast<int64_t>(cx) +
                                                                         static_cast<int64_t>(cy + cz * ly) * static_cast<int64_t>(a.rs));
                pulled[i] = inx && iny && inz;
                const uint64_t off = pulled[i] ? src : at + o * ps;
                fr[i] = ld<NTL && cx == 0>(&fa[off]);
                gr[i] = ld<NTL && cx == 0>(&ga[off]);
            }
        });
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            constexpr int cy = CY[i];
            f[i] = P::dec_f(fr[i]);
            // the value was stored by its source row: iy - c_y when pulled,
            // the cell's own row when bounced back (w_opp(i) = w_i)
            if constexpr (EX) {
                g[i] = P::template dec_g<i>(gr[i], tref_y);
            } else {
                g[i] = P::dec_f(gr[i]);
                if constexpr (cy != 0) {
                    constexpr T w = static_cast<T>(weight(i));
                    if (pulled[i]) g[i] = g[i] + w * (cy > 0 ? dref_up : dref_dn);
                }
            }
        });
        if (wall) {
            // Isothermal anti-bounce-back at the y walls for g,
            // -g_opp + 2 w_i T_w (1 + rho').  The reference applies the z walls
            // after the y walls, so a direction whose source is also outside in
            // z keeps the adiabatic bounce-back.  As deviations from w_i T_c:
            // -d_opp + 2 w_i ((T_w - T_c) + T_w rho').
            const T rw = P::dec_side(rw_raw);
            const double tw_d = wall_lo ? a.t_hot : a.t_cold;
            const T tw = static_cast<T>(tw_d);
            const T twc = EX ? tw : static_cast<T>(tw_d - static_cast<double>(tref_y));
            static_for<Q>([&](auto I) {
                constexpr int i = I;
                constexpr int cy = CY[i], cz = CZ[i];
                if constexpr (cy != 0) {
                    const bool iny = cy > 0 ? ylo : yhi;
                    const bool inz = cz == 0 ? true : (cz > 0 ? zlo : zhi);
                    if (!iny && inz) {
                        constexpr T w2 = static_cast<T>(2.0 * weight(i));
                        if constexpr (EX) g[i] = -g[i] + w2 * tw * (T(1) + rw);
                        else g[i] = -g[i] + w2 * (twc + tw * rw);
                    }
                }
            });
        }
        // ---- moments, summed in direction order (src/macrovar.cu:17-27)
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
        if constexpr (MODE == MODE_FINALIZE) {
            static_for<Q>([&](auto I) {
                constexpr int i = I;
                st<NTS>(&fb[at + i * ps], P::enc_f(f[i]));
                st<NTS>(&gb[at + i * ps], EX ? P::template enc_g<i>(g[i], tref_y) : P::enc_f(g[i]));
            });
            st<false>(&a.rho[idx], static_cast<double>(rho));
            st<false>(&a.ux[idx], static_cast<double>(ux));
            st<false>(&a.uy[idx], static_cast<double>(uy));
            st<false>(&a.uz[idx], static_cast<double>(uz));
            st<false>(&a.phi[idx], EX ? static_cast<double>(phi) : static_cast<double>(tref_y) + static_cast<double>(dsum));
            st<false>(&a.fy[idx], static_cast<double>(fprev));
        }
    }

    if constexpr (MODE != MODE_FINALIZE) {
        // ---- forcing (src/macrovar.cu:45) and collision -----------------
        const T dphi = EX ? phi - t0 : (tref_y - t0) + dsum;
        const T F = static_cast<T>(a.grav_beta) * dphi;
        collide_cell<P>(f, g, rho, ux, uy, uz, phi, dsum, F, a);
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            if constexpr (O32) {
                const uint32_t off = atb + static_cast<uint32_t>(i) * psb;
                st<NTS>(reinterpret_cast<S *>(reinterpret_cast<char *>(fb) + off), P::enc_f(f[i]));
                st<NTS>(reinterpret_cast<S *>(reinterpret_cast<char *>(gb) + off),
                        EX ? P::template enc_g<i>(g[i], tref_y) : P::enc_f(g[i]));
            } else {
                st<NTS>(&fb[at + i * ps], P::enc_f(f[i]));
                st<NTS>(&gb[at + i * ps], EX ? P::template enc_g<i>(g[i], tref_y) : P::enc_f(g[i]));
            }
        });
        st<false>(&lag[idx], P::enc_side(dphi));
        if (wall_lo || wall_hi) st<false>(&wall_rho[wall_idx], P::enc_side(rho));
    }

    if constexpr (REDUCE && MODE != MODE_PROLOGUE) {
        const double phid = EX ? static_cast<double>(phi) : static_cast<double>(tref_y) + static_cast<double>(dsum);
        diag_terms(q, true, iy, ly, static_cast<double>(rho), static_cast<double>(ux),
                   static_cast<double>(uy), static_cast<double>(uz), phid);
    }
}

template <class P, int MODE, bool REDUCE, int BX, int FLAGS>
__global__ __launch_bounds__(BX) void step_kernel(const StepArgs a) {
    const int lx = a.lx, ly = a.ly;
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
    if (ix < lx) step_cell<P, MODE, REDUCE, FLAGS>(a, ix, iy, iz, q);
    if constexpr (REDUCE && MODE != MODE_PROLOGUE)
        block_reduce_diag<BX>(q, a.partials + static_cast<uint64_t>(blockIdx.x) * NQ);
}

// Variant flags used by the plain fused pass; the other passes use FLAGS=0.
template <class P, int BXV, int FL>
hipError_t launch_fused_variant(const StepArgs &a, dim3 grid, hipStream_t s) {
    hipLaunchKernelGGL((step_kernel<P, MODE_FUSED, false, BXV, FL>), grid, dim3(BXV), 0, s, a);
    return hipGetLastError();
}

template <class P, int BXV>
hipError_t launch_bx(const StepArgs &a, int mode, bool reduce, int flags, dim3 grid, hipStream_t s) {
    if (mode == MODE_FUSED && !reduce) {
        switch (flags) {
            case 0: return launch_fused_variant<P, BXV, 0>(a, grid, s);
            case 1: return launch_fused_variant<P, BXV, 1>(a, grid, s);
            case 2: return launch_fused_variant<P, BXV, 2>(a, grid, s);
            case 3: return launch_fused_variant<P, BXV, 3>(a, grid, s);
            case 4: return launch_fused_variant<P, BXV, 4>(a, grid, s);
            case 5: return launch_fused_variant<P, BXV, 5>(a, grid, s);
            case 6: return launch_fused_variant<P, BXV, 6>(a, grid, s);
            case 7: return launch_fused_variant<P, BXV, 7>(a, grid, s);
            default: return hipErrorInvalidValue;
        }
    }
    const dim3 block(BXV);
    if (mode == MODE_PROLOGUE)
        hipLaunchKernelGGL((step_kernel<P, MODE_PROLOGUE, false, BXV, 0>), grid, block, 0, s, a);
    else if (mode == MODE_FUSED)
        hipLaunchKernelGGL((step_kernel<P, MODE_FUSED, true, BXV, 0>), grid, block, 0, s, a);
    else if (!reduce)
        hipLaunchKernelGGL((step_kernel<P, MODE_FINALIZE, false, BXV, 0>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((step_kernel<P, MODE_FINALIZE, true, BXV, 0>), grid, block, 0, s, a);
    return hipGetLastError();
}

template <class P>
hipError_t launch_step_impl(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s) {
    const int64_t nb = step_grid_blocks(a.lx, a.ly, a.lz, bx);
    if (nb <= 0 || nb > 0x7fffffffLL) return hipErrorInvalidValue;
    const dim3 grid(static_cast<uint32_t>(nb));
    switch (bx) {
        case 64: return launch_bx<P, 64>(a, mode, reduce, flags, grid, s);
        case 128: return launch_bx<P, 128>(a, mode, reduce, flags, grid, s);
        case 256: return launch_bx<P, 256>(a, mode, reduce, flags, grid, s);
        case 512: return launch_bx<P, 512>(a, mode, reduce, flags, grid, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lbm
