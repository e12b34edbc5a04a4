// pattern.hip -- copy kernels with the step kernels' memory access patterns
// and no arithmetic: the achievable-bandwidth reference of each pass.
//
// A step pass moves the 54 populations of a cell (27 f, then 27 g) in the
// row-blocked layout (pop_layout: stream s of the cell (ix, row) at
// row * 54 * LX + s * LX + ix).  How fast HBM serves those bytes depends on
// the addresses, not only on their number: an x shift of one element makes a
// wave-wide access straddle one more cache line than an aligned one, and the
// in-place passes read and write one buffer.  Each kernel here moves exactly
// the population bytes of one pass along exactly that pass's addresses (the
// same gather and scatter rules, the wall substitutions included), one cell
// per thread, with the step kernels' launch shape and cache hints, so that
// "kernel time / pattern time" says how close the arithmetic-carrying kernel
// comes to the bandwidth its access pattern allows.
//
//   COPY           own slots of A -> own slots of B
//   PULL           the A/B pass: population i from slot (i, x - c_i) of A, or
//                  (opp(i), x) when x - c_i is outside -> own slots of B
//   PUSH           own slots of A -> slot (j, x + c_j) of B, or (opp(j), x)
//                  when x + c_j is outside (the scatter form of streaming)
//   INPLACE_EVEN   the in-place even pass: own slots, population j written
//                  to the slot of opp(j); one buffer
//   INPLACE_ODD    the in-place odd pass: population i from slot
//                  (opp(i), x - c_i), or (i, x) when x - c_i is outside;
//                  written to (j, x + c_j), or (opp(j), x) when x + c_j is
//                  outside; one buffer (a cell reads and writes the same set
//                  of slots, and the sets of different cells are disjoint)
//
// The rest populations (s = 0 and 27) never move.  Stores are non-temporal
// and so are the loads without an x shift (step flags bits 0 and 1); flag
// bit 2 selects the XCD-aware block order of the step kernels.
#include "internal.hpp"
#include "step_kernel.hpp"

namespace lbm {
namespace {

constexpr int kStreams = 2 * Q;

template <bool NT, typename E>
__device__ __forceinline__ E ld_pat(const E *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <typename E, int PAT, int BX, int FLAGS>
__global__ __launch_bounds__(BX) void pattern_kernel(const E *src, E *dst, int lx, int ly, int lz, int ntx) {
    constexpr bool XCD = (FLAGS & kXcdSwizzle) != 0;
    const uint32_t b = XCD ? xcd_swizzle(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t row = b / static_cast<uint32_t>(ntx);
    const int ix = static_cast<int>(b - row * static_cast<uint32_t>(ntx)) * BX + static_cast<int>(threadIdx.x);
    if (ix >= lx) return;
    const int iy = static_cast<int>(row % static_cast<uint32_t>(ly));
    const int iz = static_cast<int>(row / static_cast<uint32_t>(ly));
    const uint64_t ps = static_cast<uint64_t>(lx);
    const uint64_t rs = kStreams * ps;
    const uint64_t at = static_cast<uint64_t>(row) * rs + ix;
    const bool xlo = ix > 0, xhi = ix < lx - 1;
    const bool ylo = iy > 0, yhi = iy < ly - 1;
    const bool zlo = iz > 0, zhi = iz < lz - 1;
    constexpr bool GATHER = PAT == LBM_ACCESS_PULL || PAT == LBM_ACCESS_INPLACE_ODD;
    constexpr bool SCATTER = PAT == LBM_ACCESS_PUSH || PAT == LBM_ACCESS_INPLACE_ODD;
    constexpr bool ODD = PAT == LBM_ACCESS_INPLACE_ODD;

    E v[kStreams];
    static_for<kStreams>([&](auto S) {
        constexpr int s = S, i = s % Q, base = s - i, cx = CX[i], cy = CY[i], cz = CZ[i], o = OPP[i];
        if constexpr (!GATHER || i == 0) {
            v[s] = ld_pat<true>(&src[at + s * ps]);
        } else {
            const bool inx = cx == 0 ? true : (cx > 0 ? xlo : xhi);
            const bool iny = cy == 0 ? true : (cy > 0 ? ylo : yhi);
            const bool inz = cz == 0 ? true : (cz > 0 ? zlo : zhi);
            const int64_t d = static_cast<int64_t>(cx) + static_cast<int64_t>(cy + cz * ly) * static_cast<int64_t>(rs);
            // streamed from x - c_i: the same stream (pull) or the opposite one
            // (in-place odd); otherwise the cell's own opposite (pull) or own
            // slot (in-place odd)
            constexpr int from = ODD ? base + o : s;
            constexpr int bounce = ODD ? s : base + o;
            const uint64_t off = (inx && iny && inz) ? at + from * ps - static_cast<uint64_t>(d) : at + bounce * ps;
            v[s] = ld_pat<cx == 0>(&src[off]);
        }
    });
    static_for<kStreams>([&](auto S) {
        constexpr int s = S, j = s % Q, base = s - j, cx = CX[j], cy = CY[j], cz = CZ[j], o = OPP[j];
        uint64_t off = at + s * ps;
        if constexpr (PAT == LBM_ACCESS_INPLACE_EVEN && j != 0) off = at + (base + o) * ps;  // exchanged slots
        if constexpr (SCATTER && j != 0) {
            const bool dx = cx == 0 ? true : (cx > 0 ? xhi : xlo);
            const bool dy = cy == 0 ? true : (cy > 0 ? yhi : ylo);
            const bool dz = cz == 0 ? true : (cz > 0 ? zhi : zlo);
            const int64_t d = static_cast<int64_t>(cx) + static_cast<int64_t>(cy + cz * ly) * static_cast<int64_t>(rs);
            off = (dx && dy && dz) ? at + s * ps + static_cast<uint64_t>(d) : at + (base + o) * ps;
        }
        __builtin_nontemporal_store(v[s], &dst[off]);
    });
}

template <typename E, int PAT, int BX>
hipError_t launch_pat_flags(const E *src, E *dst, int lx, int ly, int lz, int flags, hipStream_t s) {
    const int ntx = (lx + BX - 1) / BX;
    const dim3 grid(static_cast<uint32_t>(step_grid_blocks(lx, ly, lz, BX)));
    if (flags == 3) hipLaunchKernelGGL((pattern_kernel<E, PAT, BX, 3>), grid, dim3(BX), 0, s, src, dst, lx, ly, lz, ntx);
    else hipLaunchKernelGGL((pattern_kernel<E, PAT, BX, 7>), grid, dim3(BX), 0, s, src, dst, lx, ly, lz, ntx);
    return hipGetLastError();
}

template <typename E, int PAT>
hipError_t launch_pat_bx(const E *src, E *dst, int lx, int ly, int lz, int bx, int flags, hipStream_t s) {
    switch (bx) {
        case 128: return launch_pat_flags<E, PAT, 128>(src, dst, lx, ly, lz, flags, s);
        case 256: return launch_pat_flags<E, PAT, 256>(src, dst, lx, ly, lz, flags, s);
        default: return launch_pat_flags<E, PAT, 512>(src, dst, lx, ly, lz, flags, s);
    }
}

template <typename E>
hipError_t launch_pat(int pattern, const void *src, void *dst, int lx, int ly, int lz, int bx, int flags, hipStream_t s) {
    const E *sp = static_cast<const E *>(src);
    E *dp = static_cast<E *>(dst);
    switch (pattern) {
        case LBM_ACCESS_COPY: return launch_pat_bx<E, LBM_ACCESS_COPY>(sp, dp, lx, ly, lz, bx, flags, s);
        case LBM_ACCESS_PULL: return launch_pat_bx<E, LBM_ACCESS_PULL>(sp, dp, lx, ly, lz, bx, flags, s);
        case LBM_ACCESS_PUSH: return launch_pat_bx<E, LBM_ACCESS_PUSH>(sp, dp, lx, ly, lz, bx, flags, s);
        case LBM_ACCESS_INPLACE_EVEN: return launch_pat_bx<E, LBM_ACCESS_INPLACE_EVEN>(sp, dp, lx, ly, lz, bx, flags, s);
        default: return launch_pat_bx<E, LBM_ACCESS_INPLACE_ODD>(sp, dp, lx, ly, lz, bx, flags, s);
    }
}

// Argument checks shared by both entry points; returns LBM_OK or an error.
int check_pattern_args(int32_t pattern, int32_t elem_bytes, int32_t lx, int32_t ly, int32_t lz, int32_t block_x,
                       int32_t flags) {
    if (pattern < LBM_ACCESS_COPY || pattern > LBM_ACCESS_INPLACE_ODD)
        return ctx_fail(LBM_E_INVALID, "access pattern must be LBM_ACCESS_COPY..LBM_ACCESS_INPLACE_ODD");
    if (elem_bytes != 8 && elem_bytes != 2) return ctx_fail(LBM_E_INVALID, "elem_bytes must be 8 or 2");
    if (lx < 1 || ly < 1 || lz < 1) return ctx_fail(LBM_E_INVALID, "need lx, ly, lz >= 1");
    const uint64_t n = static_cast<uint64_t>(lx) * ly * lz;
    if (n >= (1ull << 32)) return ctx_fail(LBM_E_INVALID, "cell count must be below 2^32");
    if (block_x != 128 && block_x != 256 && block_x != 512)
        return ctx_fail(LBM_E_UNSUPPORTED, "block_x must be 128, 256 or 512");
    if (flags != 3 && flags != 7) return ctx_fail(LBM_E_UNSUPPORTED, "flags must be 3 or 7");
    if (step_grid_blocks(lx, ly, lz, block_x) > 0x7fffffffLL) return ctx_fail(LBM_E_INVALID, "grid too large");
    return LBM_OK;
}

}  // namespace

hipError_t launch_pattern(int pattern, int elem_bytes, const void *src, void *dst, int lx, int ly, int lz, int bx,
                          int flags, hipStream_t s) {
    return elem_bytes == 8 ? launch_pat<double>(pattern, src, dst, lx, ly, lz, bx, flags, s)
                           : launch_pat<uint16_t>(pattern, src, dst, lx, ly, lz, bx, flags, s);
}

}  // namespace lbm

extern "C" int lbm_pattern_copy(int32_t pattern, int32_t elem_bytes, int32_t lx, int32_t ly, int32_t lz,
                                int32_t block_x, int32_t flags, const void *src, void *dst, void *stream) {
    if (int rc = lbm::check_pattern_args(pattern, elem_bytes, lx, ly, lz, block_x, flags)) return rc;
    if (!src || !dst) return lbm::ctx_fail(LBM_E_INVALID, "NULL buffer");
    const bool inplace = pattern == LBM_ACCESS_INPLACE_EVEN || pattern == LBM_ACCESS_INPLACE_ODD;
    if (inplace != (src == dst))
        return lbm::ctx_fail(LBM_E_INVALID, inplace ? "in-place patterns need dst == src" : "dst must differ from src");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const hipError_t e = lbm::launch_pattern(pattern, elem_bytes, src, dst, lx, ly, lz, block_x, flags, s);
    if (e != hipSuccess) return lbm::ctx_fail(LBM_E_HIP, std::string("pattern copy launch: ") + hipGetErrorString(e));
    return LBM_OK;
}

extern "C" int lbm_pattern_bandwidth(int32_t pattern, int32_t elem_bytes, int32_t lx, int32_t ly, int32_t lz,
                                     int32_t block_x, int32_t flags, int32_t reps, double *ms_per_pass) {
    if (int rc = lbm::check_pattern_args(pattern, elem_bytes, lx, ly, lz, block_x, flags)) return rc;
    if (!ms_per_pass || reps < 1) return lbm::ctx_fail(LBM_E_INVALID, "need reps >= 1 and an output pointer");
    const bool inplace = pattern == LBM_ACCESS_INPLACE_EVEN || pattern == LBM_ACCESS_INPLACE_ODD;
    const uint64_t bytes = 54ull * static_cast<uint64_t>(lx) * ly * lz * static_cast<uint64_t>(elem_bytes);
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LBM_OK;
    float ms = 0.0f;
    auto run = [&](int r) {
        // out of place: alternate the buffers so that every pass reads what the previous one wrote
        const void *src = inplace || r % 2 == 0 ? a : b;
        void *dst = inplace ? a : (r % 2 == 0 ? b : a);
        return lbm_pattern_copy(pattern, elem_bytes, lx, ly, lz, block_x, flags, src, dst, st);
    };
    if (hipMalloc(&a, bytes) != hipSuccess || (!inplace && hipMalloc(&b, bytes) != hipSuccess)) {
        rc = lbm::ctx_fail(LBM_E_NOMEM, "pattern buffers of " + std::to_string(bytes) + " bytes");
        goto done;
    }
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = lbm::ctx_fail(LBM_E_HIP, "stream/event creation");
        goto done;
    }
    if (hipMemsetAsync(a, 0, bytes, st) != hipSuccess || (b && hipMemsetAsync(b, 0, bytes, st) != hipSuccess)) {
        rc = lbm::ctx_fail(LBM_E_HIP, "pattern buffer clear");
        goto done;
    }
    for (int r = 0; r < 2 && rc == LBM_OK; ++r) rc = run(r);  // warm-up
    if (rc != LBM_OK) goto done;
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps && rc == LBM_OK; ++r) rc = run(r);
    (void)hipEventRecord(e1, st);
    if (rc != LBM_OK) goto done;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = lbm::ctx_fail(LBM_E_HIP, "pattern timing");
        goto done;
    }
    *ms_per_pass = static_cast<double>(ms) / reps;
done:
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
