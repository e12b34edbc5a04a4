// noalloc.hip -- lbm_reduce_fields_noalloc: the diagnostics of lbm_reduce
// over moment fields the caller holds in device memory, with every scratch
// byte in a fixed device array, so that the call allocates nothing.  It is
// for callers that cannot allocate: code under hipGraph capture, serving
// loops that must not touch the allocator, harnesses that forbid it.
//
// The kernels are lbm_reduce's own (reduce_fields_kernel, then the fixed-order
// finish, two-level above 8 * kPartialBlocks records) with the same block
// width a context of the same shape uses by default, so the record equals
// lbm_reduce's for the same fields bit for bit.
//
// The array `scratch` below holds the per-block partial records and the
// finish's first-level sums.  Each GPU the library runs on holds its own
// copy until the process ends.  Host code never passes the array's name to a
// kernel (on the host it is a stand-in, not a device address): the first call
// on a device resolves the device address with hipGetSymbolAddress, with that
// device current, and keeps it per device.  Every record the finish reads is
// written earlier in the same call, so nothing needs clearing; calls that
// overlap on one GPU share the array and race.
//
// The array has external linkage (in namespace lbm::noalloc): with internal
// linkage (static, or an anonymous namespace) it is a local symbol of the
// device code object, which the runtime cannot resolve, and
// hipGetSymbolAddress aborts the process.
#include <mutex>

#include "internal.hpp"

namespace lbm {
namespace noalloc {

// Partial records of up to kRecords blocks plus the first-level sums.
constexpr size_t kRecords = LBM_NOALLOC_MAX_RECORDS;
constexpr size_t kScratchDoubles = (kRecords + kPartialBlocks) * NQ;

__device__ __align__(256) double scratch[kScratchDoubles];

}  // namespace noalloc
}  // namespace lbm

namespace {

using lbm::noalloc::kRecords;
using lbm::noalloc::scratch;

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double *g_scratch_dev[kMaxDevices] = {};

int default_block(int lx) {
    if (lx <= 64) return 64;
    if (lx <= 128) return 128;
    return 256;
}

}  // namespace

extern "C" int lbm_reduce_fields_noalloc(const lbm_params *p, const double *rho, const double *ux, const double *uy,
                                         const double *uz, const double *phi, int64_t step, lbm_diag *out,
                                         void *stream) {
    if (!p) return lbm::ctx_fail(LBM_E_INVALID, "params is NULL");
    if (p->lx < 1 || p->ly < 3 || p->lz < 1) return lbm::ctx_fail(LBM_E_INVALID, "need lx >= 1, ly >= 3, lz >= 1");
    const uint64_t n = static_cast<uint64_t>(p->lx) * p->ly * p->lz;
    if (n >= (1ull << 32)) return lbm::ctx_fail(LBM_E_INVALID, "cell count must be below 2^32");
    const int bx = default_block(p->lx);
    const int64_t records = lbm::step_grid_blocks(p->lx, p->ly, p->lz, bx);
    if (records > static_cast<int64_t>(kRecords))
        return lbm::ctx_fail(LBM_E_INVALID, "lbm_reduce_fields_noalloc: ceil(lx/block)*ly*lz = " +
                                                std::to_string(records) + " partial records exceed the fixed scratch of " +
                                                std::to_string(kRecords) + " (block 64/128/256 for lx <= 64/128/more)");
    if (!rho || !ux || !uy || !uz || !phi || !out) return lbm::ctx_fail(LBM_E_INVALID, "NULL field or output pointer");
    if (!(p->t_hot != p->t_cold) || !(p->prandtl > 0.0) || !(p->rayleigh > 0.0))
        return lbm::ctx_fail(LBM_E_INVALID, "need t_hot != t_cold, Ra > 0 and Pr > 0");
    lbm_derived d{};
    if (int rc = lbm_derive(p, &d)) return rc;

    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return lbm::ctx_fail(LBM_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return lbm::ctx_fail(LBM_E_UNSUPPORTED, "device index beyond 63");
    double *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if (!g_scratch_dev[dev]) {
            void *ptr = nullptr;
            e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr)
                return lbm::ctx_fail(LBM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            g_scratch_dev[dev] = static_cast<double *>(ptr);
        }
        base = g_scratch_dev[dev];
    }
    double *partials = base;
    double *first_level = base + kRecords * lbm::NQ;

    lbm::StepArgs a{};
    a.n = n;
    a.lx = p->lx;
    a.ly = p->ly;
    a.lz = p->lz;
    a.ntx = (p->lx + bx - 1) / bx;
    a.rho = const_cast<double *>(rho);
    a.ux = const_cast<double *>(ux);
    a.uy = const_cast<double *>(uy);
    a.uz = const_cast<double *>(uz);
    a.phi = const_cast<double *>(phi);
    a.partials = partials;
    lbm::DiagArgs da{};
    da.lx = p->lx;
    da.ly = p->ly;
    da.lz = p->lz;
    da.diff = d.diff;
    da.dT = p->t_hot - p->t_cold;
    da.step = step;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    e = lbm::launch_reduce_fields(a, bx, s);
    if (e == hipSuccess) e = lbm::launch_diag_finish(partials, records, da, out, first_level, s);
    if (e != hipSuccess) return lbm::ctx_fail(LBM_E_HIP, std::string("reduction launch: ") + hipGetErrorString(e));
    return LBM_OK;
}
