// internal.hpp -- shared declarations of the HIP library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "lbm_ddf.h"

namespace lbm {

// Kernel modes of the one-pass step kernel.
enum StepMode : int {
    MODE_PROLOGUE = 0,  // forcing + collide f, g of the reference state (no streaming)
    MODE_FUSED = 1,     // stream + moments of step k, forcing + collide of step k+1
    MODE_FINALIZE = 2,  // stream + moments of step k; writes the reference state
    // In-place (AA) passes over one buffer per distribution:
    MODE_AA_PROLOGUE = 3,  // layout N + reference fields -> collide -> layout E
    MODE_AA_EVEN = 4,      // layout N: moments, collide -> layout E (own slots)
    MODE_AA_ODD = 5,       // layout E: gather, moments, collide -> layout N (neighbour slots)
    MODE_AA_MOMENTS = 6,   // layout N: moments -> reference fields (no population writes)
};

// Number of per-block partial sums of the folded diagnostics.
constexpr int NQ = 16;
enum DiagQ : int {
    Q_UYPHI = 0,  // sum u_y*phi                     (src/output.cu:199-201,228-231)
    Q_UX, Q_UY, Q_UZ, Q_PHI,  // sums for the means  (src/output.cu:24-30)
    Q_USQ,        // sum |u|^2
    Q_RHO,        // sum rho'
    Q_B0, Q_B1, Q_B2,  // sum phi over the planes iy = 0, 1, 2
    Q_T0, Q_T1, Q_T2,  // sum phi over the planes iy = LY-1, LY-2, LY-3
    Q_MAXUSQ,     // max |u|^2
    Q_NONFINITE,  // count of cells with a non-finite moment
    Q_PAD
};

// Arguments of the step kernels (one cavity).  Pointers are device pointers.
struct StepArgs {
    const void *fa, *ga;  // source populations (layout: see pop_offset)
    void *fb, *gb;        // destination populations
    uint64_t n;           // cells
    uint64_t ps;          // population stride in elements (within a row)
    uint64_t rs;          // row stride in elements
    int32_t lx, ly, lz;
    int32_t ntx;          // blocks per x-row
    double tau, tauc;     // relaxation times (src/globals.cu:14,17)
    double c1;            // 1 - 0.5/tau (src/collision.cu:31)
    double grav_beta;     // grav0*beta (src/macrovar.cu:45)
    double t_hot, t_cold, t0;
    // 16-bit storage keeps g_i - w_i*T_ref(iy), with T_ref the conductive
    // profile between the walls: T_ref(iy) = tref0 - iy*dtref.
    double tref0, dtref;
    void *lag;            // phi - t0 of the previous completed step [n]
    void *wall_rho;       // rho' of the previous completed step at iy=0 and iy=LY-1 [2][lz][lx]
    // reference-state fields (prologue reads, finalize writes)
    double *rho, *ux, *uy, *uz, *phi, *fy;
    double *partials;     // [gridDim.x][NQ] when reducing
};

// Geometry/physics of one cavity for the diagnostics finish kernel.
struct DiagArgs {
    int32_t lx, ly, lz;
    double diff, dT;
    int64_t step;
};

// Launchers (defined in step_f64.hip / step_lp.hip / stage_f64.hip / reduce.hip).
hipError_t launch_step_f64(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s);
hipError_t launch_step_f16(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s);
hipError_t launch_step_bf16(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s);

// In-place (AA) pattern: step passes, the E -> N exchange (any precision; f
// or g may be null) and the flip of one distribution's post-collision
// reference state to layout E (FP64, the stage interface's streaming).
hipError_t launch_step_aa_f64(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s);
hipError_t launch_step_aa_f16(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s);
hipError_t launch_step_aa_bf16(const StepArgs &a, int mode, bool reduce, int bx, int flags, hipStream_t s);
hipError_t launch_aa_exchange(int precision, void *f, void *g, uint64_t ps, uint64_t rs, uint64_t ncell, int lx,
                              int ly, int lz, hipStream_t s);
hipError_t launch_aa_flip(const StepArgs &a, bool scalar, hipStream_t s);

// Reference-state stage kernels (FP64).
// writer_rows: 16-bit g of the in-place pattern, encoded against the row of
// the cell that wrote each slot (tref.hpp writer_row).
hipError_t launch_init(const StepArgs &a, int precision, int conductive, uint64_t seed, double amp,
                       bool writer_rows, hipStream_t s);
hipError_t launch_forcing(const StepArgs &a, hipStream_t s);
hipError_t launch_collide_f(const StepArgs &a, hipStream_t s);
hipError_t launch_collide_g(const StepArgs &a, hipStream_t s);
hipError_t launch_stream_f(const StepArgs &a, hipStream_t s);
hipError_t launch_stream_g(const StepArgs &a, hipStream_t s);
hipError_t launch_macrovar(const StepArgs &a, hipStream_t s);

// Diagnostics over the reference-state fields; the finish kernel turns the
// per-block partials into an lbm_diag at dst.
hipError_t launch_reduce_fields(const StepArgs &a, int bx, hipStream_t s);
// Large partial arrays are first summed by kPartialBlocks blocks into
// scratch (kPartialBlocks * NQ doubles); nullptr scratch: one pass.
constexpr int kPartialBlocks = 1024;
hipError_t launch_diag_finish(const double *partials, int64_t nblocks, const DiagArgs &d,
                              lbm_diag *dst, double *scratch, hipStream_t s);
hipError_t launch_profile(const StepArgs &a, double *out, hipStream_t s);

// Plain device copy of bytes (multiple of 16) for the bandwidth reference.
hipError_t launch_copy(const void *src, void *dst, uint64_t bytes, hipStream_t s);
// One access-pattern copy (LBM_ACCESS_*) over 54 row-blocked streams of
// elem_bytes 8 or 2; block_x 128/256/512, flags 3 or 7 (pattern.hip).
hipError_t launch_pattern(int pattern, int elem_bytes, const void *src, void *dst, int lx, int ly, int lz, int bx,
                          int flags, hipStream_t s);

// Population plane conversion between the storage format and FP64.
struct CodecArgs {
    int32_t lx, ly, lz;
    int32_t writer_rows;  // 16-bit in-place storage: g slots encoded against their writer's row
    uint64_t rs;          // row stride of the population layout
    double tref0, dtref;  // row reference temperature of the g storage
};
hipError_t launch_decode_plane(int precision, int which, int ip, const void *src, double *dst,
                               uint64_t n, const CodecArgs &c, hipStream_t s);
hipError_t launch_encode_plane(int precision, int which, int ip, const double *src, void *dst,
                               uint64_t n, const CodecArgs &c, hipStream_t s);

// Population layout in HBM.  Element (population ip, cell (ix, row)) of f
// lies at buf + row*rs + ip*ps + ix and of g at 27*ps further, with
// row = iz*LY + iy the cell's x-line.  Two layouts share this formula:
//
//   row-blocked (rs = 54*LX, ps = LX): each x-line holds its 27 f and 27 g
//     populations side by side, so the 54 streams that a block of one line
//     reads and writes lie within 54*LX elements and their HBM channels
//     follow from the low address bits.  With one plane per population the
//     concurrent streams lie in 54 distinct physical pages and their channel
//     spread depended on where the driver placed each page: measured on
//     MI355X, the same FP64 512^3 kernel ran between 19.2 and 21 ms from one
//     allocation to the next, against 18.6-19.4 ms row-blocked, and 128^3
//     and 96x128x96 cavities ran 10-14% faster row-blocked.
//   planes (rs = LX, ps >= N): the reference's [ip][z][y][x] with each plane
//     staggered by 64 elements.  Used only for the cavities of an ensemble
//     whose x-line is not a whole number of 128-byte lines (LX * element
//     size % 128 != 0): an ensemble's flat tiles let a 64-lane wave cover the
//     end of one x-line and the start of the next, which are contiguous only
//     in this layout (row-blocked they would split lines: measured, 48 cubes
//     of side 33-256, FP16, 19.9k MLUPS in planes against 13.7k).  The single
//     cavity launches one block per x-line segment, so no wave spans lines.
//
// Host copies convert to the reference's order.
struct PopLayout {
    uint64_t ps, rs, elems;  // population stride, row stride, elements of one f+g buffer
};
inline PopLayout pop_layout(int lx, int ly, int lz, size_t elem_bytes, bool flat_tiles) {
    const uint64_t n = static_cast<uint64_t>(lx) * ly * lz;
    if (!flat_tiles || (static_cast<uint64_t>(lx) * elem_bytes) % 128 == 0)
        return {static_cast<uint64_t>(lx), 54ull * lx, 54ull * n};
    const uint64_t ps = (n + 63) / 64 * 64 + 64;
    return {ps, static_cast<uint64_t>(lx), 54ull * ps};
}

inline int64_t step_grid_blocks(int lx, int ly, int lz, int bx) {
    int64_t ntx = (lx + bx - 1) / bx;
    return ntx * (int64_t)ly * (int64_t)lz;
}

enum Phase { PHASE_EMPTY = 0, PHASE_REF = 1, PHASE_FUSED = 2 };
// lbm_step replays runs of this many fused passes (an even number, so the
// buffer parity returns to where it started) from a captured hipGraph.
constexpr int kGraphPasses = 32;
enum AaLayout { AA_LAYOUT_N = 0, AA_LAYOUT_E = 1 };

}  // namespace lbm

// One cavity's device state (opaque to C callers).  Each buffer buf[k]
// holds f and g (see pop_layout); f[k] and g[k] point at their first
// elements.  In REF phase f[cf], g[cg] are the reference's post-streaming
// populations, in FUSED phase its post-collision ones.  With the in-place
// pattern only buf[0] exists (cf = cg = 0) and aa_layout tells which of the
// two layouts it holds (always N in REF phase).
struct lbm_ctx {
    lbm_params p{};
    lbm_derived d{};
    double beta = 0.0;
    uint64_t n = 0;
    uint64_t ps = 0;  // population stride (elements)
    uint64_t rs = 0;  // row stride (elements)
    int prec = 0;
    void *buf[2] = {nullptr, nullptr};
    void *f[2] = {nullptr, nullptr};
    void *g[2] = {nullptr, nullptr};
    int cf = 0, cg = 0;
    double *rho = nullptr, *ux = nullptr, *uy = nullptr, *uz = nullptr, *phi = nullptr, *fy = nullptr;
    void *lag = nullptr;
    void *wall_rho = nullptr;
    double *partials = nullptr;
    double *partials2 = nullptr;  // first-level sums (kPartialBlocks x NQ)
    lbm_diag *diag_dev = nullptr;
    int64_t diag_cap = 0;
    double *plane = nullptr;  // one population plane in FP64 (codec staging)
    hipStream_t stream = nullptr;
    bool own_stream = true;  // false when an ensemble lends its stream
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Captured sequences of kGraphPasses fused passes, one per starting
    // parity state; key = cf | cg << 1 | aa_layout << 2 | flags << 3 | bx << 6.
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_key[2] = {-1, -1};
    bool aa = false;    // in-place pattern (LBM_PATTERN_AA)
    int aa_layout = 0;  // AA_LAYOUT_N or AA_LAYOUT_E
    int phase = lbm::PHASE_EMPTY;
    int64_t nstep = 0;
    int bx = 256;
    int flags = 0;
    int device = 0;
    // owned by an ensemble: readable through the single-cavity calls, but
    // stepping or destroying it on its own would desynchronise the
    // ensemble's descriptor table (the public calls refuse with LBM_E_STATE)
    bool member = false;
};


namespace lbm {
// Helpers shared by capi.hip and ensemble.hip (defined in capi.hip).
int ctx_create(const lbm_params *p, hipStream_t stream, bool flat_tiles, lbm_ctx **out);
StepArgs ctx_args(const lbm_ctx *c, const void *fa, const void *ga, void *fb, void *gb);
int ctx_prologue(lbm_ctx *c);    // REF -> FUSED (collides the next step)
int ctx_ensure_ref(lbm_ctx *c);  // FUSED -> REF (finalize)
int ctx_forcing(lbm_ctx *c);     // forcing kernel on the reference state (any precision)
int ctx_destroy(lbm_ctx *c);     // frees a context, ensemble member or not
int ctx_fail(int code, const std::string &msg);
}  // namespace lbm
