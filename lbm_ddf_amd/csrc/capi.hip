// capi.hip -- the C ABI of include/lbm_ddf.h: context, phases, host copies.
//
// A context holds two population buffers, each with f and g (pull A/B
// scheme; population layout: internal.hpp pop_layout), the reference-state
// moment fields, and the two per-cell side values of the fused kernel.  It
// is in one of three phases:
//   EMPTY  allocated, no state yet;
//   REF    populations are the reference's post-streaming f, g and the
//          fields rho', u, phi, F_y are the reference's (as after
//          src/main.cu:45);
//   FUSED  populations are post-collision (the streaming half of the last
//          step is pending), with lag phi and wall rho' valid.
// Calls that need the reference state run the finalize pass first.
//
// With the in-place pattern (LBM_PATTERN_AA) there is one buffer and the
// passes alternate between its two slot layouts
// (aa_kernel.hpp): REF phase always holds layout N; in FUSED phase aa_layout
// says whether the next pass is odd (layout E) or even (layout N).  Finalize
// is the moments pass, preceded by the in-place exchange when the layout is E.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(LBM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define LBM_HIP(call)                                      \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return hip_fail(e_, #call);  \
    } while (0)


size_t elem_size(int prec) { return prec == LBM_PREC_FP64 ? 8 : 2; }
// lag phi and wall rho': FP64, scaled FP16 (FP16 storage) or FP32 (BF16 storage)
size_t side_size(int prec) { return prec == LBM_PREC_FP64 ? 8 : (prec == LBM_PREC_FP16 ? 2 : 4); }

int default_block(int lx) {
    if (lx <= 64) return 64;
    if (lx <= 128) return 128;
    return 256;
}

int validate(const lbm_params *p) {
    if (!p) return fail(LBM_E_INVALID, "params is NULL");
    if (p->lx < 1 || p->ly < 3 || p->lz < 1) return fail(LBM_E_INVALID, "need lx >= 1, ly >= 3, lz >= 1");
    const uint64_t n = static_cast<uint64_t>(p->lx) * p->ly * p->lz;
    if (n >= (1ull << 32)) return fail(LBM_E_INVALID, "cell count must be below 2^32");
    if (p->precision < LBM_PREC_FP64 || p->precision > LBM_PREC_BF16)
        return fail(LBM_E_INVALID, "unknown precision");
    if (p->pattern != LBM_PATTERN_AB && p->pattern != LBM_PATTERN_AA)
        return fail(LBM_E_INVALID, "unknown pattern");
    if (p->reserved != 0) return fail(LBM_E_INVALID, "reserved must be 0");
    if (!(p->rayleigh > 0.0) || !(p->prandtl > 0.0)) return fail(LBM_E_INVALID, "Ra and Pr must be > 0");
    // the scaled binary16 code holds magnitudes below 2 (step_kernel.hpp, PolicyF16)
    if (p->precision == LBM_PREC_FP16 && !(std::fabs(p->t_hot) <= 1.0 && std::fabs(p->t_cold) <= 1.0))
        return fail(LBM_E_INVALID, "FP16 storage needs |t_hot|, |t_cold| <= 1");
    return LBM_OK;
}

}  // namespace

using lbm::PHASE_EMPTY;
using lbm::PHASE_FUSED;
using lbm::PHASE_REF;

namespace {

lbm::StepArgs make_args(const lbm_ctx *c, const void *fa, const void *ga, void *fb, void *gb) {
    lbm::StepArgs a{};
    a.fa = fa;
    a.ga = ga;
    a.fb = fb;
    a.gb = gb;
    a.n = c->n;
    a.ps = c->ps;
    a.rs = c->rs;
    a.lx = c->p.lx;
    a.ly = c->p.ly;
    a.lz = c->p.lz;
    a.ntx = (c->p.lx + c->bx - 1) / c->bx;
    a.tau = c->d.tau;
    a.tauc = c->d.tauc;
    a.c1 = 1.0 - 0.5 / c->d.tau;
    a.grav_beta = c->p.gravity * c->beta;
    a.t_hot = c->p.t_hot;
    a.t_cold = c->p.t_cold;
    a.t0 = c->d.t0;
    a.dtref = (c->p.t_hot - c->p.t_cold) / c->p.ly;
    a.tref0 = c->p.t_hot - 0.5 * a.dtref;
    a.lag = c->lag;
    a.wall_rho = c->wall_rho;
    a.rho = c->rho;
    a.ux = c->ux;
    a.uy = c->uy;
    a.uz = c->uz;
    a.phi = c->phi;
    a.fy = c->fy;
    a.partials = c->partials;
    return a;
}

lbm::CodecArgs codec_args(const lbm_ctx *c) {
    const lbm::StepArgs a = make_args(c, nullptr, nullptr, nullptr, nullptr);
    return lbm::CodecArgs{a.lx, a.ly, a.lz, c->aa ? 1 : 0, a.rs, a.tref0, a.dtref};
}

hipError_t launch_step(const lbm_ctx *c, const lbm::StepArgs &a, int mode, bool reduce) {
    if (mode >= lbm::MODE_AA_PROLOGUE) {
        switch (c->prec) {
            case LBM_PREC_FP64: return lbm::launch_step_aa_f64(a, mode, reduce, c->bx, c->flags, c->stream);
            case LBM_PREC_FP16: return lbm::launch_step_aa_f16(a, mode, reduce, c->bx, c->flags, c->stream);
            case LBM_PREC_BF16: return lbm::launch_step_aa_bf16(a, mode, reduce, c->bx, c->flags, c->stream);
            default: return hipErrorInvalidValue;
        }
    }
    switch (c->prec) {
        case LBM_PREC_FP64: return lbm::launch_step_f64(a, mode, reduce, c->bx, c->flags, c->stream);
        case LBM_PREC_FP16: return lbm::launch_step_f16(a, mode, reduce, c->bx, c->flags, c->stream);
        case LBM_PREC_BF16: return lbm::launch_step_bf16(a, mode, reduce, c->bx, c->flags, c->stream);
        default: return hipErrorInvalidValue;
    }
}

// One in-place pass: the AA counterpart of the given A/B mode.
int run_pass_aa(lbm_ctx *c, int mode, bool reduce) {
    lbm::StepArgs a = make_args(c, nullptr, nullptr, c->f[0], c->g[0]);
    int aa_mode = lbm::MODE_AA_MOMENTS;
    int next = lbm::AA_LAYOUT_N;
    if (mode == lbm::MODE_PROLOGUE) {
        aa_mode = lbm::MODE_AA_PROLOGUE;
        next = lbm::AA_LAYOUT_E;
    } else if (mode == lbm::MODE_FUSED) {
        aa_mode = c->aa_layout == lbm::AA_LAYOUT_E ? lbm::MODE_AA_ODD : lbm::MODE_AA_EVEN;
        next = c->aa_layout == lbm::AA_LAYOUT_E ? lbm::AA_LAYOUT_N : lbm::AA_LAYOUT_E;
    } else if (c->aa_layout == lbm::AA_LAYOUT_E) {
        LBM_HIP(lbm::launch_aa_exchange(c->prec, c->f[0], c->g[0], c->ps, c->rs, c->n, c->p.lx, c->p.ly, c->p.lz,
                                        c->stream));
    }
    LBM_HIP(launch_step(c, a, aa_mode, reduce));
    c->aa_layout = next;
    return LBM_OK;
}

// One pass of the given mode from the current buffers into the others.
int run_pass(lbm_ctx *c, int mode, bool reduce) {
    if (c->aa) return run_pass_aa(c, mode, reduce);
    lbm::StepArgs a = make_args(c, c->f[c->cf], c->g[c->cg], c->f[1 - c->cf], c->g[1 - c->cg]);
    LBM_HIP(launch_step(c, a, mode, reduce));
    c->cf = 1 - c->cf;
    c->cg = 1 - c->cg;
    return LBM_OK;
}

int graph_key(const lbm_ctx *c) {
    return c->cf | c->cg << 1 | c->aa_layout << 2 | c->flags << 3 | c->bx << 6;
}

// kGraphPasses plain fused passes from the current state, replayed from a
// graph captured once per starting state (two states alternate at most:
// the parities after an even and after an odd number of passes).  The
// passes launch no memset and capture their arguments by value.
int run_fused_graph(lbm_ctx *c) {
    const int key = graph_key(c);
    int k = c->graph_key[0] == key ? 0 : (c->graph_key[1] == key ? 1 : -1);
    if (k < 0) {
        k = c->graph[0] == nullptr ? 0 : (c->graph[1] == nullptr ? 1 : 0);
        if (c->graph[k]) {
            LBM_HIP(hipGraphExecDestroy(c->graph[k]));
            c->graph[k] = nullptr;
            c->graph_key[k] = -1;
        }
        hipGraph_t g = nullptr;
        // capture records the passes without running them; their buffer
        // bookkeeping is restored if the capture fails part-way
        const int cf = c->cf, cg = c->cg, layout = c->aa_layout;
        LBM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = LBM_OK;
        for (int i = 0; i < lbm::kGraphPasses && rc == LBM_OK; ++i) rc = run_pass(c, lbm::MODE_FUSED, false);
        const hipError_t ec = hipStreamEndCapture(c->stream, &g);
        if (rc != LBM_OK || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            c->cf = cf;
            c->cg = cg;
            c->aa_layout = layout;
            return rc != LBM_OK ? rc : hip_fail(ec, "hipStreamEndCapture");
        }
        // the captured passes flipped the bookkeeping an even number of times
        const hipError_t ei = hipGraphInstantiate(&c->graph[k], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            c->graph[k] = nullptr;
            return hip_fail(ei, "hipGraphInstantiate");
        }
        c->graph_key[k] = key;
    }
    LBM_HIP(hipGraphLaunch(c->graph[k], c->stream));
    return LBM_OK;
}

lbm::DiagArgs diag_args(const lbm_ctx *c, int64_t step) {
    lbm::DiagArgs d{};
    d.lx = c->p.lx;
    d.ly = c->p.ly;
    d.lz = c->p.lz;
    d.diff = c->d.diff;
    d.dT = c->p.t_hot - c->p.t_cold;
    d.step = step;
    return d;
}

int ensure_diag_cap(lbm_ctx *c, int64_t cap) {
    if (cap <= c->diag_cap) return LBM_OK;
    if (c->diag_dev) {
        LBM_HIP(hipStreamSynchronize(c->stream));
        LBM_HIP(hipFree(c->diag_dev));
        c->diag_dev = nullptr;
        c->diag_cap = 0;
    }
    if (hipMalloc(&c->diag_dev, sizeof(lbm_diag) * cap) != hipSuccess) return fail(LBM_E_NOMEM, "diag buffer");
    c->diag_cap = cap;
    return LBM_OK;
}

int64_t nblocks(const lbm_ctx *c) { return lbm::step_grid_blocks(c->p.lx, c->p.ly, c->p.lz, c->bx); }

// Completes a pending streaming half-step (FUSED -> REF), optionally with the
// diagnostics of that step written to diag_dev[slot].
int finalize(lbm_ctx *c, bool reduce, int64_t slot) {
    if (c->phase != PHASE_FUSED) return LBM_OK;
    int rc = run_pass(c, lbm::MODE_FINALIZE, reduce);
    if (rc) return rc;
    if (reduce) LBM_HIP(lbm::launch_diag_finish(c->partials, nblocks(c), diag_args(c, c->nstep), c->diag_dev + slot, c->partials2, c->stream));
    c->phase = PHASE_REF;
    return LBM_OK;
}

int ensure_ref(lbm_ctx *c) {
    if (c->phase == PHASE_EMPTY) return fail(LBM_E_STATE, "context has no state: call lbm_init_* first");
    return finalize(c, false, 0);
}

int check_ctx(const lbm_ctx *c) {
    if (!c) return fail(LBM_E_INVALID, "ctx is NULL");
    return LBM_OK;
}

// Calls that advance the state (or free it) are the ensemble's to make for
// its members.
int check_own(const lbm_ctx *c, const char *what) {
    if (int rc = check_ctx(c)) return rc;
    if (c->member)
        return fail(LBM_E_STATE, std::string(what) + ": the context is an ensemble member (step it through lbm_ensemble_*)");
    return LBM_OK;
}

int require_fp64(const lbm_ctx *c, const char *what) {
    if (c->prec != LBM_PREC_FP64)
        return fail(LBM_E_UNSUPPORTED, std::string(what) + " is available for FP64 contexts only");
    return LBM_OK;
}

double *field_ptr(lbm_ctx *c, int field) {
    switch (field) {
        case LBM_FIELD_RHO: return c->rho;
        case LBM_FIELD_UX: return c->ux;
        case LBM_FIELD_UY: return c->uy;
        case LBM_FIELD_UZ: return c->uz;
        case LBM_FIELD_PHI: return c->phi;
        case LBM_FIELD_FY: return c->fy;
        default: return nullptr;
    }
}

void derive(const lbm_params *p, lbm_derived *d, double *beta_out) {
    // src/globals.cu:11-23; BETA default src/parameters.h:26.
    const double beta = p->beta > 0.0 ? p->beta : 0.1 / ((double)p->ly * 5.0);
    const double grav0 = p->gravity, tHot = p->t_hot, tCold = p->t_cold;
    d->visc = (double)p->ly * std::sqrt(grav0 * beta * (tHot - tCold) * (double)p->ly * p->prandtl / p->rayleigh);
    d->tau = 3.0 * d->visc + 0.5;
    d->u0 = std::sqrt(grav0 * beta * (tHot - tCold) * (double)p->ly);
    d->diff = d->visc / p->prandtl;
    d->tauc = 3.0 * d->diff + 0.5;
    d->t0 = 0.5 * (tCold + tHot);
    if (beta_out) *beta_out = beta;
}

// Partial-sum records for the largest grid any block width can launch.
size_t max_partials_bytes(const lbm_params *p) {
    return static_cast<size_t>(lbm::step_grid_blocks(p->lx, p->ly, p->lz, 64)) * lbm::NQ * 8;
}

bool valid_block(int lx, int bx) {
    if (bx != 64 && bx != 128 && bx != 256 && bx != 512) return false;
    return bx == 64 || bx / 2 < lx;  // no block more than half idle
}

// The in-place passes are built for flags 0, 3, 4 and 7.
bool valid_flags(bool aa, int flags) {
    if (flags < 0 || flags > 7) return false;
    return !aa || flags == 0 || flags == 3 || flags == 4 || flags == 7;
}

// What ctx_create allocates besides the arrays counted in bytes_for: the
// number of device allocations (two population buffers, six fields, lag, wall
// rho', partials, partials2, diagnostics records) and the granule each is
// rounded up to.
constexpr uint64_t kAllocations = 13;
constexpr uint64_t kAllocGranule = 4096;
constexpr uint64_t kDiagRecords = 16;

uint64_t bytes_for(const lbm_params *p, int bx) {
    const uint64_t n = static_cast<uint64_t>(p->lx) * p->ly * p->lz;
    const bool aa = p->pattern == LBM_PATTERN_AA;
    // the larger of the two layouts: a single cavity's or an ensemble member's
    const size_t es = elem_size(p->precision);
    const uint64_t elems = std::max(lbm::pop_layout(p->lx, p->ly, p->lz, es, false).elems,
                                    lbm::pop_layout(p->lx, p->ly, p->lz, es, true).elems);
    const uint64_t pops = (aa ? 1ull : 2ull) * elems * es;
    const uint64_t fields = 6ull * n * 8ull;
    const uint64_t side = n * side_size(p->precision) + (aa ? 0ull : 2ull * p->lx * p->lz * side_size(p->precision));
    // diagnostics: per-block partials, their first-level sums, 16 records
    const uint64_t diag = max_partials_bytes(p) + static_cast<uint64_t>(lbm::kPartialBlocks) * lbm::NQ * 8ull +
                          kDiagRecords * sizeof(lbm_diag);
    // each allocation rounded up to the allocator's granule
    const uint64_t slack = kAllocations * kAllocGranule;
    (void)bx;
    return pops + fields + side + diag + slack;
}

}  // namespace

extern "C" {

const char *lbm_last_error(void) { return g_last_error.c_str(); }

int lbm_derive(const lbm_params *p, lbm_derived *out) {
    if (!p || !out) return fail(LBM_E_INVALID, "NULL argument");
    if (p->ly < 1) return fail(LBM_E_INVALID, "ly must be >= 1");
    derive(p, out, nullptr);
    return LBM_OK;
}

int lbm_bytes_required(const lbm_params *p, uint64_t *bytes) {
    int rc = validate(p);
    if (rc) return rc;
    if (!bytes) return fail(LBM_E_INVALID, "bytes is NULL");
    *bytes = bytes_for(p, default_block(p->lx));
    return LBM_OK;
}

int lbm_destroy(lbm_ctx *c) {
    if (!c) return LBM_OK;
    if (int rc = check_own(c, "lbm_destroy")) return rc;
    return lbm::ctx_destroy(c);
}

int lbm_create(const lbm_params *p, lbm_ctx **out) { return lbm::ctx_create(p, nullptr, false, out); }

}  // extern "C"

namespace lbm {

int ctx_destroy(lbm_ctx *c) {
    if (!c) return LBM_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipGraphExec_t &g : c->graph)
        if (g) (void)hipGraphExecDestroy(g);
    void *bufs[] = {c->buf[0], c->buf[1], c->rho, c->ux, c->uy, c->uz, c->phi, c->fy,
                    c->lag, c->wall_rho, c->partials, c->partials2, c->diag_dev, c->plane};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LBM_OK;
}

int ctx_fail(int code, const std::string &msg) { return fail(code, msg); }

StepArgs ctx_args(const lbm_ctx *c, const void *fa, const void *ga, void *fb, void *gb) {
    return make_args(c, fa, ga, fb, gb);
}

int ctx_prologue(lbm_ctx *c) {
    if (c->phase == PHASE_EMPTY) return fail(LBM_E_STATE, "context has no state: call lbm_init_* first");
    if (c->phase != PHASE_REF) return LBM_OK;
    if (int rc = run_pass(c, lbm::MODE_PROLOGUE, false)) return rc;
    c->phase = PHASE_FUSED;
    c->nstep += 1;
    return LBM_OK;
}

int ctx_ensure_ref(lbm_ctx *c) { return ensure_ref(c); }

int ctx_forcing(lbm_ctx *c) {
    if (int rc = ensure_ref(c)) return rc;
    LBM_HIP(lbm::launch_forcing(make_args(c, nullptr, nullptr, nullptr, nullptr), c->stream));
    LBM_HIP(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

// Creates a context; with stream != nullptr the context borrows that stream
// (an ensemble's) instead of creating its own.  flat_tiles: the context is an
// ensemble member, stepped in flat tiles (selects its population layout).
int ctx_create(const lbm_params *p, hipStream_t stream, bool flat_tiles, lbm_ctx **out) {
    int rc = validate(p);
    if (rc) return rc;
    if (!out) return fail(LBM_E_INVALID, "out is NULL");
    *out = nullptr;
    lbm_ctx *c = new lbm_ctx();
    c->p = *p;
    c->prec = p->precision;
    c->aa = p->pattern == LBM_PATTERN_AA;
    c->n = static_cast<uint64_t>(p->lx) * p->ly * p->lz;
    const lbm::PopLayout lay = lbm::pop_layout(p->lx, p->ly, p->lz, elem_size(p->precision), flat_tiles);
    c->ps = lay.ps;
    c->rs = lay.rs;
    derive(p, &c->d, &c->beta);
    c->bx = default_block(p->lx);
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) {
        ctx_destroy(c);
        return hip_fail(e, "hipGetDevice");
    }
    if (stream) {
        c->stream = stream;
        c->own_stream = false;
    } else {
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    }
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
        ctx_destroy(c);
        return hip_fail(e, "stream/event creation");
    }
    const uint64_t n = c->n;
    const size_t buf_bytes = lay.elems * elem_size(c->prec);
    struct Alloc {
        void **ptr;
        size_t bytes;
    } allocs[] = {
        {&c->buf[0], buf_bytes},
        {&c->buf[1], buf_bytes},
        {reinterpret_cast<void **>(&c->rho), n * 8},
        {reinterpret_cast<void **>(&c->ux), n * 8},
        {reinterpret_cast<void **>(&c->uy), n * 8},
        {reinterpret_cast<void **>(&c->uz), n * 8},
        {reinterpret_cast<void **>(&c->phi), n * 8},
        {reinterpret_cast<void **>(&c->fy), n * 8},
        {&c->lag, n * side_size(c->prec)},
        {&c->wall_rho, 2ull * p->lx * p->lz * side_size(c->prec)},
        {reinterpret_cast<void **>(&c->partials), max_partials_bytes(p)},
        {reinterpret_cast<void **>(&c->partials2), static_cast<size_t>(lbm::kPartialBlocks) * lbm::NQ * 8},
    };
    for (auto &al : allocs) {
        // one buffer and no wall-rho array in place
        if (c->aa && (al.ptr == &c->buf[1] || al.ptr == &c->wall_rho)) continue;
        if (hipMalloc(al.ptr, al.bytes) != hipSuccess) {
            (void)hipGetLastError();
            ctx_destroy(c);
            return fail(LBM_E_NOMEM, "device allocation of " + std::to_string(al.bytes) + " bytes failed");
        }
    }
    // g starts 27 populations after f (within each row, or after f's planes)
    for (int k = 0; k < 2; ++k) {
        if (!c->buf[k]) continue;
        c->f[k] = c->buf[k];
        c->g[k] = static_cast<char *>(c->buf[k]) + 27ull * c->ps * elem_size(c->prec);
    }
    rc = ensure_diag_cap(c, kDiagRecords);
    if (rc) {
        ctx_destroy(c);
        return rc;
    }
    *out = c;
    return LBM_OK;
}

}  // namespace lbm

extern "C" {

int lbm_init_arrays(lbm_ctx *c) {
    if (int rc = check_ctx(c)) return rc;
    lbm::StepArgs a = make_args(c, nullptr, nullptr, c->f[c->cf], c->g[c->cg]);
    LBM_HIP(lbm::launch_init(a, c->prec, 0, 0, 0.0, c->aa, c->stream));
    c->phase = PHASE_REF;
    c->aa_layout = lbm::AA_LAYOUT_N;
    c->nstep = 0;
    return LBM_OK;
}

int lbm_init_conductive(lbm_ctx *c, uint64_t seed, double amp) {
    if (int rc = check_ctx(c)) return rc;
    lbm::StepArgs a = make_args(c, nullptr, nullptr, c->f[c->cf], c->g[c->cg]);
    LBM_HIP(lbm::launch_init(a, c->prec, 1, seed, amp, c->aa, c->stream));
    c->phase = PHASE_REF;
    c->aa_layout = lbm::AA_LAYOUT_N;
    c->nstep = 0;
    return LBM_OK;
}

// ---- stage interface ------------------------------------------------------

#define LBM_STAGE_PROLOGUE(name)                      \
    if (int rc = check_own(c, name)) return rc;       \
    if (int rc = require_fp64(c, name)) return rc;    \
    if (int rc = ensure_ref(c)) return rc

int lbm_forcing(lbm_ctx *c) {
    LBM_STAGE_PROLOGUE("lbm_forcing");
    LBM_HIP(lbm::launch_forcing(make_args(c, nullptr, nullptr, nullptr, nullptr), c->stream));
    return LBM_OK;
}

int lbm_collision_BGK(lbm_ctx *c) {
    LBM_STAGE_PROLOGUE("lbm_collision_BGK");
    LBM_HIP(lbm::launch_collide_f(make_args(c, nullptr, nullptr, c->f[c->cf], nullptr), c->stream));
    return LBM_OK;
}

int lbm_collision_BGK_scalar(lbm_ctx *c) {
    LBM_STAGE_PROLOGUE("lbm_collision_BGK_scalar");
    LBM_HIP(lbm::launch_collide_g(make_args(c, nullptr, nullptr, nullptr, c->g[c->cg]), c->stream));
    return LBM_OK;
}

int lbm_streaming(lbm_ctx *c) {
    LBM_STAGE_PROLOGUE("lbm_streaming");
    if (c->aa) {
        // in place: each cell moves its populations to the exchanged slots
        // (with the wall rules), then the pairwise exchange completes the pull
        LBM_HIP(lbm::launch_aa_flip(make_args(c, nullptr, nullptr, c->f[0], nullptr), false, c->stream));
        LBM_HIP(lbm::launch_aa_exchange(c->prec, c->f[0], nullptr, c->ps, c->rs, c->n, c->p.lx, c->p.ly, c->p.lz,
                                        c->stream));
        return LBM_OK;
    }
    LBM_HIP(lbm::launch_stream_f(make_args(c, c->f[c->cf], nullptr, c->f[1 - c->cf], nullptr), c->stream));
    c->cf = 1 - c->cf;
    return LBM_OK;
}

int lbm_streaming_scalar(lbm_ctx *c) {
    LBM_STAGE_PROLOGUE("lbm_streaming_scalar");
    if (c->aa) {
        LBM_HIP(lbm::launch_aa_flip(make_args(c, nullptr, nullptr, nullptr, c->g[0]), true, c->stream));
        LBM_HIP(lbm::launch_aa_exchange(c->prec, nullptr, c->g[0], c->ps, c->rs, c->n, c->p.lx, c->p.ly, c->p.lz,
                                        c->stream));
        return LBM_OK;
    }
    LBM_HIP(lbm::launch_stream_g(make_args(c, nullptr, c->g[c->cg], nullptr, c->g[1 - c->cg]), c->stream));
    c->cg = 1 - c->cg;
    return LBM_OK;
}

int lbm_macrovar(lbm_ctx *c) {
    LBM_STAGE_PROLOGUE("lbm_macrovar");
    LBM_HIP(lbm::launch_macrovar(make_args(c, c->f[c->cf], c->g[c->cg], nullptr, nullptr), c->stream));
    c->nstep += 1;
    return LBM_OK;
}

// The reference's execution structure, src/main.cu:27-48: forcing, both
// collisions, sync; both streamings, sync; macrovar, sync -- per step.
int lbm_step_staged(lbm_ctx *c, int64_t nsteps) {
    LBM_STAGE_PROLOGUE("lbm_step_staged");
    if (nsteps < 0) return fail(LBM_E_INVALID, "nsteps < 0");
    for (int64_t k = 0; k < nsteps; ++k) {
        if (int rc = lbm_forcing(c)) return rc;
        if (int rc = lbm_collision_BGK(c)) return rc;
        if (int rc = lbm_collision_BGK_scalar(c)) return rc;
        LBM_HIP(hipStreamSynchronize(c->stream));
        if (int rc = lbm_streaming(c)) return rc;
        if (int rc = lbm_streaming_scalar(c)) return rc;
        LBM_HIP(hipStreamSynchronize(c->stream));
        if (int rc = lbm_macrovar(c)) return rc;
        LBM_HIP(hipStreamSynchronize(c->stream));
    }
    return LBM_OK;
}

// ---- fused path -----------------------------------------------------------

int lbm_step(lbm_ctx *c, int64_t nsteps) {
    if (int rc = check_own(c, "lbm_step")) return rc;
    if (nsteps < 0) return fail(LBM_E_INVALID, "nsteps < 0");
    if (nsteps == 0) return LBM_OK;
    if (c->phase == PHASE_EMPTY) return fail(LBM_E_STATE, "context has no state: call lbm_init_* first");
    int64_t fused = nsteps;
    if (c->phase == PHASE_REF) {
        if (int rc = run_pass(c, lbm::MODE_PROLOGUE, false)) return rc;
        c->phase = PHASE_FUSED;
        fused -= 1;
    }
    // long runs: replay captured sequences of passes (fewer launches)
    if (fused >= 2 * lbm::kGraphPasses)
        for (; fused >= lbm::kGraphPasses; fused -= lbm::kGraphPasses)
            if (int rc = run_fused_graph(c)) return rc;
    for (int64_t k = 0; k < fused; ++k)
        if (int rc = run_pass(c, lbm::MODE_FUSED, false)) return rc;
    c->nstep += nsteps;
    return LBM_OK;
}

int lbm_step_timed(lbm_ctx *c, int64_t nsteps, double *elapsed_ms) {
    if (int rc = check_own(c, "lbm_step_timed")) return rc;
    if (!elapsed_ms) return fail(LBM_E_INVALID, "elapsed_ms is NULL");
    LBM_HIP(hipEventRecord(c->ev0, c->stream));
    if (int rc = lbm_step(c, nsteps)) return rc;
    LBM_HIP(hipEventRecord(c->ev1, c->stream));
    LBM_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    LBM_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *elapsed_ms = ms;
    return LBM_OK;
}

int lbm_run_diag(lbm_ctx *c, int64_t nsteps, int64_t every, lbm_diag *out, int64_t cap, int64_t *nout) {
    if (int rc = check_own(c, "lbm_run_diag")) return rc;
    if (nsteps < 0 || every <= 0 || cap < 0 || (cap > 0 && !out))
        return fail(LBM_E_INVALID, "bad nsteps/every/cap/out");
    if (c->phase == PHASE_EMPTY) return fail(LBM_E_STATE, "context has no state: call lbm_init_* first");
    const int64_t start = c->nstep;  // diagnostics for completed steps s in (start, start+nsteps]
    const int64_t end = start + nsteps;
    if (int rc = ensure_diag_cap(c, cap > 0 ? cap : 1)) return rc;
    int64_t nrec = 0;
    if (nsteps > 0) {
        int64_t todo = nsteps;
        // In FUSED phase the step `start` is collided and its streaming half
        // is pending; from REF the prologue collides step start+1.
        int64_t pending = start;
        if (c->phase == PHASE_REF) {
            if (int rc = run_pass(c, lbm::MODE_PROLOGUE, false)) return rc;
            c->phase = PHASE_FUSED;
            todo -= 1;
            pending = start + 1;
        }
        for (int64_t k = 0; k < todo; ++k) {
            const int64_t s = pending + k;
            const bool red = s > start && s % every == 0 && nrec < cap;
            if (int rc = run_pass(c, lbm::MODE_FUSED, red)) return rc;
            if (red) {
                LBM_HIP(lbm::launch_diag_finish(c->partials, nblocks(c), diag_args(c, s), c->diag_dev + nrec, c->partials2, c->stream));
                ++nrec;
            }
        }
        c->nstep = end;
        const bool red = end % every == 0 && nrec < cap;
        if (int rc = finalize(c, red, nrec)) return rc;
        if (red) ++nrec;
    } else {
        if (int rc = ensure_ref(c)) return rc;
    }
    if (nrec > 0) LBM_HIP(hipMemcpyAsync(out, c->diag_dev, sizeof(lbm_diag) * nrec, hipMemcpyDeviceToHost, c->stream));
    LBM_HIP(hipStreamSynchronize(c->stream));
    if (nout) *nout = nrec;
    return LBM_OK;
}

int lbm_finalize(lbm_ctx *c) {
    if (int rc = check_ctx(c)) return rc;
    return ensure_ref(c);
}

int lbm_reduce(lbm_ctx *c, lbm_diag *out) {
    if (int rc = check_ctx(c)) return rc;
    if (!out) return fail(LBM_E_INVALID, "out is NULL");
    if (int rc = ensure_ref(c)) return rc;
    LBM_HIP(lbm::launch_reduce_fields(make_args(c, nullptr, nullptr, nullptr, nullptr), c->bx, c->stream));
    LBM_HIP(lbm::launch_diag_finish(c->partials, nblocks(c), diag_args(c, c->nstep), c->diag_dev, c->partials2, c->stream));
    LBM_HIP(hipMemcpyAsync(out, c->diag_dev, sizeof(lbm_diag), hipMemcpyDeviceToHost, c->stream));
    LBM_HIP(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_profile(lbm_ctx *c, double *out) {
    if (int rc = check_ctx(c)) return rc;
    if (!out) return fail(LBM_E_INVALID, "out is NULL");
    if (int rc = ensure_ref(c)) return rc;
    // the per-plane results go to the diagnostics partials buffer (at least
    // 16 * LY doubles), which is free between stream-ordered reductions
    double *dev = c->partials;
    LBM_HIP(lbm::launch_profile(make_args(c, nullptr, nullptr, nullptr, nullptr), dev, c->stream));
    LBM_HIP(hipMemcpyAsync(out, dev, sizeof(double) * 5 * c->p.ly, hipMemcpyDeviceToHost, c->stream));
    LBM_HIP(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_step_count(const lbm_ctx *c, int64_t *out) {
    if (int rc = check_ctx(c)) return rc;
    if (!out) return fail(LBM_E_INVALID, "out is NULL");
    *out = c->nstep;
    return LBM_OK;
}

int lbm_synchronize(lbm_ctx *c) {
    if (int rc = check_ctx(c)) return rc;
    LBM_HIP(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

// ---- tuning ---------------------------------------------------------------

int lbm_set_tuning(lbm_ctx *c, int32_t block_x, int32_t flags) {
    if (int rc = check_ctx(c)) return rc;
    if (!valid_block(c->p.lx, block_x)) return fail(LBM_E_INVALID, "block_x must be 64, 128, 256 or 512 and fit LX");
    if (!valid_flags(c->aa, flags))
        return fail(LBM_E_INVALID, c->aa ? "flags must be 0, 3, 4 or 7 (in-place pattern)" : "flags must be in 0..7");
    // Partial sums depend on the block width only within one reducing call,
    // and every variant computes the same values: switching is safe at any time.
    c->bx = block_x;
    c->flags = flags;
    return LBM_OK;
}

int lbm_get_tuning(const lbm_ctx *c, int32_t *block_x, int32_t *flags) {
    if (int rc = check_ctx(c)) return rc;
    if (block_x) *block_x = c->bx;
    if (flags) *flags = c->flags;
    return LBM_OK;
}

int lbm_autotune(lbm_ctx *c, int32_t steps, int64_t *steps_taken, double *best_ms) {
    if (int rc = check_own(c, "lbm_autotune")) return rc;
    if (steps < 1) return fail(LBM_E_INVALID, "steps must be >= 1");
    if (c->phase == PHASE_EMPTY) return fail(LBM_E_STATE, "context has no state: call lbm_init_* first");
    int64_t taken = 0;
    if (c->phase == PHASE_REF) {
        if (int rc = lbm_step(c, 1)) return rc;
        taken += 1;
    }
    const int blocks[] = {64, 128, 256, 512};
    int best_bx = c->bx, best_fl = c->flags;
    double best = 1e300;
    for (int bx : blocks) {
        if (!valid_block(c->p.lx, bx)) continue;
        for (int fl = 0; fl < 8; ++fl) {
            if (!valid_flags(c->aa, fl)) continue;
            c->bx = bx;
            c->flags = fl;
            // one untimed launch per candidate (code object load, caches)
            if (int rc = lbm_step(c, 1)) return rc;
            double ms = 0.0;
            if (int rc = lbm_step_timed(c, steps, &ms)) return rc;
            taken += 1 + steps;
            if (ms < best) {
                best = ms;
                best_bx = bx;
                best_fl = fl;
            }
        }
    }
    c->bx = best_bx;
    c->flags = best_fl;
    if (steps_taken) *steps_taken = taken;
    if (best_ms) *best_ms = best / steps;
    return LBM_OK;
}

int lbm_pattern_time(lbm_ctx *c, int32_t reps, double *ms_per_pass) {
    if (int rc = check_own(c, "lbm_pattern_time")) return rc;
    if (!ms_per_pass || reps < 1) return fail(LBM_E_INVALID, "need reps >= 1 and an output array of two");
    const int lx = c->p.lx, ly = c->p.ly, lz = c->p.lz;
    if (c->ps != static_cast<uint64_t>(lx) || c->rs != 54ull * static_cast<uint64_t>(lx))
        return fail(LBM_E_UNSUPPORTED, "pattern timing needs the row-blocked single-cavity layout");
    const int elem = c->prec == LBM_PREC_FP64 ? 8 : 2;
    const int bx = c->bx <= 128 ? 128 : (c->bx <= 256 ? 256 : 512);
    const int fl = (c->flags & 4) ? 7 : 3;  // bit 2: the XCD-aware block order
    if (lbm::step_grid_blocks(lx, ly, lz, bx) > 0x7fffffffLL) return fail(LBM_E_INVALID, "grid too large");
    c->phase = PHASE_EMPTY;  // the copies overwrite the populations
    for (int par = 0; par < 2; ++par) {
        const int pat = c->aa ? (par == 0 ? LBM_ACCESS_INPLACE_EVEN : LBM_ACCESS_INPLACE_ODD) : LBM_ACCESS_PULL;
        const void *src = c->aa ? c->buf[0] : c->buf[par];
        void *dst = c->aa ? c->buf[0] : c->buf[1 - par];
        for (int r = 0; r < 2; ++r) LBM_HIP(lbm::launch_pattern(pat, elem, src, dst, lx, ly, lz, bx, fl, c->stream));
        LBM_HIP(hipEventRecord(c->ev0, c->stream));
        for (int r = 0; r < reps; ++r) LBM_HIP(lbm::launch_pattern(pat, elem, src, dst, lx, ly, lz, bx, fl, c->stream));
        LBM_HIP(hipEventRecord(c->ev1, c->stream));
        LBM_HIP(hipEventSynchronize(c->ev1));
        float ms = 0.0f;
        LBM_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        ms_per_pass[par] = static_cast<double>(ms) / reps;
    }
    return LBM_OK;
}

int lbm_copy_bandwidth(uint64_t bytes, int32_t reps, double *gbps) {
    if (!gbps || reps < 1 || bytes < 4096) return fail(LBM_E_INVALID, "bad arguments");
    bytes &= ~static_cast<uint64_t>(4095);
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LBM_OK;
    float ms = 0.0f;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        rc = fail(LBM_E_NOMEM, "copy buffers");
        goto done;
    }
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = fail(LBM_E_HIP, "stream/event creation");
        goto done;
    }
    if (hipMemsetAsync(a, 0, bytes, st) != hipSuccess || lbm::launch_copy(a, b, bytes, st) != hipSuccess) {
        rc = fail(LBM_E_HIP, "copy warm-up");
        goto done;
    }
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r) (void)lbm::launch_copy(r % 2 ? b : a, r % 2 ? a : b, bytes, st);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = fail(LBM_E_HIP, "copy timing");
        goto done;
    }
    *gbps = 2.0 * static_cast<double>(bytes) * reps / (ms * 1e-3) / 1e9;
done:
    if (st) (void)hipStreamSynchronize(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

// ---- host copies ----------------------------------------------------------

int lbm_get_field(lbm_ctx *c, int32_t field, double *host) {
    if (int rc = check_ctx(c)) return rc;
    if (!host) return fail(LBM_E_INVALID, "host is NULL");
    if (field < LBM_FIELD_RHO || field > LBM_FIELD_FZ) return fail(LBM_E_INVALID, "unknown field");
    if (int rc = ensure_ref(c)) return rc;
    if (field == LBM_FIELD_FX || field == LBM_FIELD_FZ) {
        // The reference writes 0.0 to F_x and F_z (src/macrovar.cu:44,46).
        LBM_HIP(hipStreamSynchronize(c->stream));
        std::memset(host, 0, sizeof(double) * c->n);
        return LBM_OK;
    }
    LBM_HIP(hipMemcpyAsync(host, field_ptr(c, field), sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
    LBM_HIP(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_set_field(lbm_ctx *c, int32_t field, const double *host) {
    if (int rc = check_ctx(c)) return rc;
    if (!host) return fail(LBM_E_INVALID, "host is NULL");
    if (field < LBM_FIELD_RHO || field > LBM_FIELD_FZ) return fail(LBM_E_INVALID, "unknown field");
    if (int rc = ensure_ref(c)) return rc;
    if (field == LBM_FIELD_FX || field == LBM_FIELD_FZ) {
        for (uint64_t i = 0; i < c->n; ++i)
            if (host[i] != 0.0) return fail(LBM_E_UNSUPPORTED, "F_x and F_z are identically zero in this model");
        return LBM_OK;
    }
    LBM_HIP(hipMemcpyAsync(field_ptr(c, field), host, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
    LBM_HIP(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_get_pops(lbm_ctx *c, int32_t which, double *host) {
    if (int rc = check_ctx(c)) return rc;
    if (!host || (which != LBM_POP_F && which != LBM_POP_G)) return fail(LBM_E_INVALID, "bad which/host");
    if (int rc = ensure_ref(c)) return rc;
    // populations leave in the reference's order [ip][z][y][x], one plane at
    // a time through a staging buffer (the codec also converts 16-bit storage)
    void *src = which == LBM_POP_F ? c->f[c->cf] : c->g[c->cg];
    if (!c->plane && hipMalloc(&c->plane, sizeof(double) * c->n) != hipSuccess)
        return fail(LBM_E_NOMEM, "plane staging buffer");
    const size_t es = elem_size(c->prec);
    for (int ip = 0; ip < 27; ++ip) {
        const void *pl = static_cast<const char *>(src) + static_cast<size_t>(ip) * c->ps * es;
        LBM_HIP(lbm::launch_decode_plane(c->prec, which, ip, pl, c->plane, c->n, codec_args(c), c->stream));
        LBM_HIP(hipMemcpyAsync(host + static_cast<size_t>(ip) * c->n, c->plane, sizeof(double) * c->n,
                               hipMemcpyDeviceToHost, c->stream));
    }
    LBM_HIP(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_set_pops(lbm_ctx *c, int32_t which, const double *host) {
    if (int rc = check_ctx(c)) return rc;
    if (!host || (which != LBM_POP_F && which != LBM_POP_G)) return fail(LBM_E_INVALID, "bad which/host");
    if (int rc = ensure_ref(c)) return rc;
    void *dst = which == LBM_POP_F ? c->f[c->cf] : c->g[c->cg];
    if (!c->plane && hipMalloc(&c->plane, sizeof(double) * c->n) != hipSuccess)
        return fail(LBM_E_NOMEM, "plane staging buffer");
    const size_t es = elem_size(c->prec);
    for (int ip = 0; ip < 27; ++ip) {
        void *pd = static_cast<char *>(dst) + static_cast<size_t>(ip) * c->ps * es;
        LBM_HIP(hipMemcpyAsync(c->plane, host + static_cast<size_t>(ip) * c->n, sizeof(double) * c->n,
                               hipMemcpyHostToDevice, c->stream));
        LBM_HIP(lbm::launch_encode_plane(c->prec, which, ip, c->plane, pd, c->n, codec_args(c), c->stream));
    }
    LBM_HIP(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

}  // extern "C"
