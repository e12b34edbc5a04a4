// stage.hip -- reference-state kernels: initial state, the six stage kernels
// of the reference time step (one launch each, as src/main.cu:27-45 issues
// them), the diagnostics over the moment fields, the per-plane profile and
// the population-plane codecs.  Compiled with -ffp-contract=off.
//
// These kernels are written for clarity, one thread per cell over a flat
// index; the time-critical path is step_kernel.hpp.
#include "step_kernel.hpp"

namespace lbm {

namespace {

constexpr int kFlatBlock = 256;

__device__ __forceinline__ void cell_coords(uint64_t idx, int lx, int ly, int &ix, int &iy, int &iz) {
    const uint64_t row = idx / static_cast<uint64_t>(lx);
    ix = static_cast<int>(idx - row * lx);
    iy = static_cast<int>(row % static_cast<uint64_t>(ly));
    iz = static_cast<int>(row / static_cast<uint64_t>(ly));
}

inline dim3 flat_grid(uint64_t n) { return dim3(static_cast<uint32_t>((n + kFlatBlock - 1) / kFlatBlock)); }

// Slot of population 0 of the cell with flat index idx (x fastest) in the
// row-blocked population layout; population ip is ip*ps further.
__device__ __forceinline__ uint64_t slot_of(uint64_t idx, int lx, uint64_t rs) {
    const uint64_t row = idx / static_cast<uint64_t>(lx);
    return row * rs + (idx - row * lx);
}

// Counter-based draw (splitmix64 finaliser); matches the oracle's
// orc_perturbation bit for bit.
__device__ __forceinline__ double perturbation(uint64_t seed, uint64_t idx, double amp) {
    uint64_t z = seed + (idx + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    const double u = static_cast<double>(z >> 11) * 0x1.0p-53;
    return amp * (2.0 * u - 1.0);
}

// init_arrays, src/initial.cu:208-242 (conductive: phi from the linear
// profile of Parallel/initial.cu:89 plus the seeded perturbation).
// writer_rows: 16-bit in-place storage, whose reference state encodes each
// slot against the row of the cell that streamed it in (tref.hpp).
template <class P>
__global__ void init_kernel(StepArgs a, int conductive, uint64_t seed, double amp, int writer_rows) {
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= a.n) return;
    int ix, iy, iz;
    cell_coords(idx, a.lx, a.ly, ix, iy, iz);
    double phi = 0.5;
    if (conductive) {
        const double t = a.t_hot - (a.t_hot - a.t_cold) * iy / (a.ly - 1.0);
        phi = t + perturbation(seed, idx, amp);
    }
    const uint64_t ps = a.ps;
    const uint64_t at = slot_of(idx, a.lx, a.rs);
    a.rho[idx] = 0.0;
    a.ux[idx] = 0.0;
    a.uy[idx] = 0.0;
    a.uz[idx] = 0.0;
    a.phi[idx] = phi;
    a.fy[idx] = a.grav_beta * (phi - a.t0);
    const double rho = 0.0, usq = 0.0;
    const double fy = a.fy[idx];
    auto *f = static_cast<typename P::S *>(a.fb);
    auto *g = static_cast<typename P::S *>(a.gb);
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        const double feq = weight(i) * rho * (1.0 - 1.5 * usq);
        const double fv = feq - 0.5 * fy * CY[i] / (1.0 / 3.0) * feq;
        const double geq = weight(i) * phi * (1.0 - 1.5 * usq);
        const int row = writer_rows ? writer_row(ix, iy, iz, a.lx, a.ly, a.lz, CX[i], CY[i], CZ[i]) : iy;
        const typename P::T tref = tref_row<typename P::T>(a.tref0, a.dtref, row);
        f[at + i * ps] = P::enc_f(static_cast<typename P::T>(fv));
        g[at + i * ps] = P::template enc_g<i>(static_cast<typename P::T>(geq), tref);
    });
}

// forcing, src/macrovar.cu:36-47 (F_x = F_z = 0 are not stored).
__global__ void forcing_kernel(StepArgs a) {
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= a.n) return;
    a.fy[idx] = a.grav_beta * (a.phi[idx] - a.t0);
}

// collision_BGK, src/collision.cu:3-32, in place on f.
__global__ void collide_f_kernel(StepArgs a) {
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= a.n) return;
    const uint64_t ps = a.ps;
    const uint64_t at = slot_of(idx, a.lx, a.rs);
    double *f = static_cast<double *>(a.fb);
    const double fy9 = a.fy[idx], rho9 = a.rho[idx];
    const double u9 = a.ux[idx], v9 = a.uy[idx], w9 = a.uz[idx];
    const double tau = a.tau;
    static_for<Q>([&](auto I) {
        constexpr int ip = I;
        const double RT = 1.0 / 3.0;
        const double eu = (CX[ip] * u9 + CY[ip] * v9 + CZ[ip] * w9) / RT;
        const double uv = (u9 * u9 + v9 * v9 + w9 * w9) / RT;
        const double feq = weight(ip) * (rho9 + eu + 0.5 * (eu * eu - uv));
        double fchange = -(f[at + ip * ps] - feq) / tau;
        const double feq_rho0 = weight(ip) * (1.0 + eu + 0.5 * (eu * eu - uv));
        fchange += (1.0 - 0.5 / tau) * fy9 * (CY[ip] - v9) / RT * feq_rho0;
        f[at + ip * ps] += fchange;
    });
}

// collision_BGK_scalar, src/collision.cu:52-81, in place on g.
__global__ void collide_g_kernel(StepArgs a) {
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= a.n) return;
    const uint64_t ps = a.ps;
    const uint64_t at = slot_of(idx, a.lx, a.rs);
    double *g = static_cast<double *>(a.gb);
    const double phi9 = a.phi[idx];
    const double u9 = a.ux[idx], v9 = a.uy[idx], w9 = a.uz[idx];
    const double RT = 1.0 / 3.0;
    const double tauc = a.tauc;
    static_for<Q>([&](auto I) {
        constexpr int ip = I;
        const double g9 = g[at + ip * ps];
        const double eu = (CX[ip] * u9 + CY[ip] * v9 + CZ[ip] * w9) / RT;
        const double uv = (u9 * u9 + v9 * v9 + w9 * w9) / RT;
        const double geq = weight(ip) * phi9 * (1.0 + eu + 0.5 * (eu * eu - uv));
        const double gchange = -(g9 - geq) / tauc;
        g[at + ip * ps] += gchange;
    });
}

// streaming / streaming_scalar, src/streaming.cu:3-251: pull from a (fa or
// ga) into b, then the wall blocks in the reference order x, y, z.
template <bool SCALAR>
__global__ void stream_kernel(StepArgs a) {
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= a.n) return;
    int ix, iy, iz;
    cell_coords(idx, a.lx, a.ly, ix, iy, iz);
    const int lx = a.lx, ly = a.ly, lz = a.lz;
    const uint64_t ps = a.ps;
    const uint64_t at = slot_of(idx, lx, a.rs);
    const double *src = static_cast<const double *>(SCALAR ? a.ga : a.fa);
    double *dst = static_cast<double *>(SCALAR ? a.gb : a.fb);
    static_for<Q>([&](auto I) {
        constexpr int ip = I;
        const int sx = ix - CX[ip], sy = iy - CY[ip], sz = iz - CZ[ip];
        if (sx >= 0 && sx < lx && sy >= 0 && sy < ly && sz >= 0 && sz < lz) {
            const uint64_t sat = (static_cast<uint64_t>(sz) * ly + sy) * a.rs + sx;
            dst[at + ip * ps] = src[sat + ip * ps];
        }
    });
    // Wall blocks: each writes the directions pointing into the domain from
    // that wall.  x walls, y walls, z walls, in that order.
    auto bb = [&](auto I) {
        constexpr int ip = I;
        dst[at + ip * ps] = src[at + OPP[ip] * ps];
    };
    static_for<Q>([&](auto I) {
        constexpr int ip = I;
        if constexpr (CX[ip] > 0) if (ix == 0) bb(I);
    });
    static_for<Q>([&](auto I) {
        constexpr int ip = I;
        if constexpr (CX[ip] < 0) if (ix == lx - 1) bb(I);
    });
    if constexpr (SCALAR) {
        const double density = (iy == 0 || iy == ly - 1) ? a.rho[idx] : 0.0;
        static_for<Q>([&](auto I) {
            constexpr int ip = I;
            if constexpr (CY[ip] > 0)
                if (iy == 0) dst[at + ip * ps] = -src[at + OPP[ip] * ps] + 2.0 * weight(ip) * a.t_hot * (1.0 + density);
        });
        static_for<Q>([&](auto I) {
            constexpr int ip = I;
            if constexpr (CY[ip] < 0)
                if (iy == ly - 1) dst[at + ip * ps] = -src[at + OPP[ip] * ps] + 2.0 * weight(ip) * a.t_cold * (1.0 + density);
        });
    } else {
        static_for<Q>([&](auto I) {
            constexpr int ip = I;
            if constexpr (CY[ip] > 0) if (iy == 0) bb(I);
        });
        static_for<Q>([&](auto I) {
            constexpr int ip = I;
            if constexpr (CY[ip] < 0) if (iy == ly - 1) bb(I);
        });
    }
    static_for<Q>([&](auto I) {
        constexpr int ip = I;
        if constexpr (CZ[ip] > 0) if (iz == 0) bb(I);
    });
    static_for<Q>([&](auto I) {
        constexpr int ip = I;
        if constexpr (CZ[ip] < 0) if (iz == lz - 1) bb(I);
    });
}

// macrovar, src/macrovar.cu:3-34 (F_x = F_z = 0).
__global__ void macrovar_kernel(StepArgs a) {
    const uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= a.n) return;
    const uint64_t ps = a.ps;
    const uint64_t at = slot_of(idx, a.lx, a.rs);
    const double *f = static_cast<const double *>(a.fa);
    const double *g = static_cast<const double *>(a.ga);
    double rho = 0.0, ux = 0.0, uy = 0.0, uz = 0.0, phi = 0.0;
    static_for<Q>([&](auto I) {
        constexpr int ip = I;
        const double fv = f[at + ip * ps];
        const double gv = g[at + ip * ps];
        rho += fv;
        ux += CX[ip] * fv;
        uy += CY[ip] * fv;
        uz += CZ[ip] * fv;
        phi += gv;
    });
    a.rho[idx] = rho;
    a.ux[idx] = ux + 0.5 * 0.0;
    a.uy[idx] = uy + 0.5 * a.fy[idx];
    a.uz[idx] = uz + 0.5 * 0.0;
    a.phi[idx] = phi;
}

// Diagnostics over the moment fields; same partial layout as the fused
// kernel (one NQ record per block of BX cells along x).
template <int BX>
__global__ __launch_bounds__(BX) void reduce_fields_kernel(StepArgs a) {
    const uint32_t b = blockIdx.x;
    const uint32_t row = b / static_cast<uint32_t>(a.ntx);
    const int bxi = static_cast<int>(b - row * static_cast<uint32_t>(a.ntx));
    const int iy = static_cast<int>(row % static_cast<uint32_t>(a.ly));
    const int iz = static_cast<int>(row / static_cast<uint32_t>(a.ly));
    const int ix = bxi * BX + static_cast<int>(threadIdx.x);
    const bool active = ix < a.lx;
    double q[NQ];
    if (active) {
        const uint64_t idx = (static_cast<uint64_t>(iz) * a.ly + iy) * a.lx + ix;
        diag_terms(q, true, iy, a.ly, a.rho[idx], a.ux[idx], a.uy[idx], a.uz[idx], a.phi[idx]);
    } else {
        diag_terms(q, false, iy, a.ly, 0.0, 0.0, 0.0, 0.0, 0.0);
    }
    block_reduce_diag<BX>(q, a.partials + static_cast<uint64_t>(blockIdx.x) * NQ);
}

// Sums the per-block partials in a fixed order and applies the formulas of
// src/output.cu:228-256 (Nu) and :31-34 (means).  The one-sided wall
// gradients are linear in phi, so the plane sums give the same plane means:
//   <dT/dy>_bottom = (-3 S0 + 4 S1 - S2) / 2 / (LX*LZ)      (:211, :236-253)
//   <dT/dy>_top    = ( 3 T0 - 4 T1 + T2) / 2 / (LX*LZ)      (:216)
__global__ __launch_bounds__(kFinishThreads) void diag_finish_kernel(const double *partials, int64_t nblocks,
                                                                    DiagArgs d, lbm_diag *dst) {
    diag_finish_block(partials, nblocks, d, dst);
}

// First level for large grids: block b sums records [b*per, (b+1)*per) into
// out[b].  The split depends only on nblocks, so the result is deterministic.
__global__ __launch_bounds__(kFinishThreads) void partial_sum_kernel(const double *partials, int64_t nblocks,
                                                                    int64_t per, double *out) {
    __shared__ double sh[kFinishThreads][NQ + 1];
    const int64_t first = static_cast<int64_t>(blockIdx.x) * per;
    const int64_t count = first < nblocks ? (nblocks - first < per ? nblocks - first : per) : 0;
    sum_records_block(partials + first * NQ, count, sh);
    if (threadIdx.x < NQ) out[static_cast<int64_t>(blockIdx.x) * NQ + threadIdx.x] = sh[0][threadIdx.x];
}

// output_profile, src/output.cu:336-358: one block per y-plane.
constexpr int kProfileThreads = 256;
__global__ __launch_bounds__(kProfileThreads) void profile_kernel(StepArgs a, double *out) {
    __shared__ double sh[kProfileThreads][5];
    const int iy = blockIdx.x;
    const int t = threadIdx.x;
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    const int64_t plane = static_cast<int64_t>(a.lx) * a.lz;
    for (int64_t j = t; j < plane; j += kProfileThreads) {
        const int iz = static_cast<int>(j / a.lx);
        const int ix = static_cast<int>(j - static_cast<int64_t>(iz) * a.lx);
        const uint64_t idx = (static_cast<uint64_t>(iz) * a.ly + iy) * a.lx + ix;
        const double p = a.phi[idx];
        s[0] += a.ux[idx];
        s[1] += a.uy[idx];
        s[2] += a.uz[idx];
        s[3] += p;
        s[4] += p * p;
    }
    for (int k = 0; k < 5; ++k) sh[t][k] = s[k];
    __syncthreads();
    for (int st = kProfileThreads / 2; st >= 1; st >>= 1) {
        if (t < st)
            for (int k = 0; k < 5; ++k) sh[t][k] += sh[t + st][k];
        __syncthreads();
    }
    if (t == 0) {
        const double count = static_cast<double>(plane);
        const double ap = sh[0][3] / count;
        out[iy * 5 + 0] = sh[0][0] / count;
        out[iy * 5 + 1] = sh[0][1] / count;
        out[iy * 5 + 2] = sh[0][2] / count;
        out[iy * 5 + 3] = ap;
        out[iy * 5 + 4] = sqrt(sh[0][4] / count - ap * ap);
    }
}

// Reference temperature of the 16-bit g storage of population IP at flat
// cell index i: the binary32 value the step kernels encode and decode that
// slot with (its own row; with writer_rows the row of the cell that wrote
// the slot), in binary64.
template <int IP>
__device__ __forceinline__ double codec_tref(uint64_t i, const CodecArgs &c) {
    const uint64_t r = i / static_cast<uint64_t>(c.lx);
    const int ix = static_cast<int>(i - r * static_cast<uint64_t>(c.lx));
    const int iy = static_cast<int>(r % static_cast<uint64_t>(c.ly));
    const int iz = static_cast<int>(r / static_cast<uint64_t>(c.ly));
    const int row = c.writer_rows ? writer_row(ix, iy, iz, c.lx, c.ly, c.lz, CX[IP], CY[IP], CZ[IP]) : iy;
    return static_cast<double>(tref_row<float>(c.tref0, c.dtref, row));
}

// Host copies convert in binary64: a decoded value is the stored value
// exactly (the stored deviation, unscaled, plus w_i * T_ref(iy) for g), and
// encoding a decoded value gives back the same bits, so reads, writes and
// checkpoints round-trip.  A new value is rounded to nearest through binary32
// (the formats' own rounding, as in the step kernels).
template <class P>
__device__ __forceinline__ double stored_deviation(typename P::S s) {
    if constexpr (std::is_same_v<P, PolicyF64>) return s;
    else if constexpr (std::is_same_v<P, PolicyF16>) return static_cast<double>(static_cast<float>(s)) * (1.0 / 32768.0);
    else return static_cast<double>(bf16_bits_to_float(s));
}
template <class P>
__device__ __forceinline__ typename P::S store_deviation(double d) {
    if constexpr (std::is_same_v<P, PolicyF64>) return d;
    else if constexpr (std::is_same_v<P, PolicyF16>) return static_cast<_Float16>(static_cast<float>(d * 32768.0));
    else return float_to_bf16_bits(static_cast<float>(d));
}

template <class P, int WHICH, int IP>
__global__ void decode_kernel(const void *src, double *dst, uint64_t n, CodecArgs c) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const auto s = static_cast<const typename P::S *>(src)[slot_of(i, c.lx, c.rs)];
    if constexpr (WHICH == 0 || P::exact) dst[i] = stored_deviation<P>(s);
    else dst[i] = stored_deviation<P>(s) + weight(IP) * codec_tref<IP>(i, c);
}

template <class P, int WHICH, int IP>
__global__ void encode_kernel(const double *src, void *dst, uint64_t n, CodecArgs c) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    auto *d = static_cast<typename P::S *>(dst);
    const uint64_t at = slot_of(i, c.lx, c.rs);
    if constexpr (WHICH == 0 || P::exact) d[at] = store_deviation<P>(src[i]);
    else d[at] = store_deviation<P>(src[i] - weight(IP) * codec_tref<IP>(i, c));
}

template <class P>
hipError_t codec_plane(bool encode, int which, int ip, const void *src, void *dst, uint64_t n, const CodecArgs &t0,
                       hipStream_t s) {
    hipError_t err = hipErrorInvalidValue;
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        if (i != ip) return;
        if (encode) {
            if (which == 0)
                hipLaunchKernelGGL((encode_kernel<P, 0, i>), flat_grid(n), dim3(kFlatBlock), 0, s,
                                   static_cast<const double *>(src), dst, n, t0);
            else
                hipLaunchKernelGGL((encode_kernel<P, 1, i>), flat_grid(n), dim3(kFlatBlock), 0, s,
                                   static_cast<const double *>(src), dst, n, t0);
        } else {
            if (which == 0)
                hipLaunchKernelGGL((decode_kernel<P, 0, i>), flat_grid(n), dim3(kFlatBlock), 0, s, src,
                                   static_cast<double *>(dst), n, t0);
            else
                hipLaunchKernelGGL((decode_kernel<P, 1, i>), flat_grid(n), dim3(kFlatBlock), 0, s, src,
                                   static_cast<double *>(dst), n, t0);
        }
        err = hipGetLastError();
    });
    return err;
}

// Plain device copy, 16 bytes per thread, one thread per element: the
// achievable-bandwidth reference and the calibration kernel for HBM counters.
// (Measured on MI355X over 16 GiB: 6.19 TB/s, against 4.85 TB/s for a
// grid-stride loop and 5.6-5.9 TB/s for 4 or 8 elements per thread.)
__global__ __launch_bounds__(256) void copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n16) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n16) dst[i] = src[i];
}

}  // namespace

hipError_t launch_init(const StepArgs &a, int precision, int conductive, uint64_t seed, double amp,
                       bool writer_rows, hipStream_t s) {
    const int wr = writer_rows ? 1 : 0;
    switch (precision) {
        case LBM_PREC_FP64:
            hipLaunchKernelGGL(init_kernel<PolicyF64>, flat_grid(a.n), dim3(kFlatBlock), 0, s, a, conductive, seed, amp, wr);
            break;
        case LBM_PREC_FP16:
            hipLaunchKernelGGL(init_kernel<PolicyF16>, flat_grid(a.n), dim3(kFlatBlock), 0, s, a, conductive, seed, amp, wr);
            break;
        case LBM_PREC_BF16:
            hipLaunchKernelGGL(init_kernel<PolicyBF16>, flat_grid(a.n), dim3(kFlatBlock), 0, s, a, conductive, seed, amp, wr);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_forcing(const StepArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(forcing_kernel, flat_grid(a.n), dim3(kFlatBlock), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_collide_f(const StepArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(collide_f_kernel, flat_grid(a.n), dim3(kFlatBlock), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_collide_g(const StepArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(collide_g_kernel, flat_grid(a.n), dim3(kFlatBlock), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_stream_f(const StepArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(stream_kernel<false>, flat_grid(a.n), dim3(kFlatBlock), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_stream_g(const StepArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(stream_kernel<true>, flat_grid(a.n), dim3(kFlatBlock), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_macrovar(const StepArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(macrovar_kernel, flat_grid(a.n), dim3(kFlatBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_reduce_fields(const StepArgs &a, int bx, hipStream_t s) {
    const int64_t nb = step_grid_blocks(a.lx, a.ly, a.lz, bx);
    switch (bx) {
        case 64: hipLaunchKernelGGL(reduce_fields_kernel<64>, dim3(nb), dim3(64), 0, s, a); break;
        case 128: hipLaunchKernelGGL(reduce_fields_kernel<128>, dim3(nb), dim3(128), 0, s, a); break;
        case 256: hipLaunchKernelGGL(reduce_fields_kernel<256>, dim3(nb), dim3(256), 0, s, a); break;
        case 512: hipLaunchKernelGGL(reduce_fields_kernel<512>, dim3(nb), dim3(512), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_diag_finish(const double *partials, int64_t nblocks, const DiagArgs &d, lbm_diag *dst,
                              double *scratch, hipStream_t s) {
    if (scratch && nblocks > 8 * kPartialBlocks) {
        const int64_t per = (nblocks + kPartialBlocks - 1) / kPartialBlocks;
        hipLaunchKernelGGL(partial_sum_kernel, dim3(kPartialBlocks), dim3(kFinishThreads), 0, s, partials, nblocks, per,
                           scratch);
        hipLaunchKernelGGL(diag_finish_kernel, dim3(1), dim3(kFinishThreads), 0, s, scratch,
                           static_cast<int64_t>(kPartialBlocks), d, dst);
    } else {
        hipLaunchKernelGGL(diag_finish_kernel, dim3(1), dim3(kFinishThreads), 0, s, partials, nblocks, d, dst);
    }
    return hipGetLastError();
}

hipError_t launch_profile(const StepArgs &a, double *out, hipStream_t s) {
    hipLaunchKernelGGL(profile_kernel, dim3(a.ly), dim3(kProfileThreads), 0, s, a, out);
    return hipGetLastError();
}

hipError_t launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t s) {
    const uint64_t n16 = bytes / 16;
    const uint64_t blocks = (n16 + 255) / 256;
    if (blocks == 0 || blocks > 0x7fffffffull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(copy_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, s,
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n16);
    return hipGetLastError();
}

hipError_t launch_decode_plane(int precision, int which, int ip, const void *src, double *dst, uint64_t n,
                               const CodecArgs &t0, hipStream_t s) {
    switch (precision) {
        case LBM_PREC_FP64: return codec_plane<PolicyF64>(false, which, ip, src, dst, n, t0, s);
        case LBM_PREC_FP16: return codec_plane<PolicyF16>(false, which, ip, src, dst, n, t0, s);
        case LBM_PREC_BF16: return codec_plane<PolicyBF16>(false, which, ip, src, dst, n, t0, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_encode_plane(int precision, int which, int ip, const double *src, void *dst, uint64_t n,
                               const CodecArgs &t0, hipStream_t s) {
    switch (precision) {
        case LBM_PREC_FP64: return codec_plane<PolicyF64>(true, which, ip, src, dst, n, t0, s);
        case LBM_PREC_FP16: return codec_plane<PolicyF16>(true, which, ip, src, dst, n, t0, s);
        case LBM_PREC_BF16: return codec_plane<PolicyBF16>(true, which, ip, src, dst, n, t0, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace lbm
