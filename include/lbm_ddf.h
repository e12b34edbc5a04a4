/*
 * lbm_ddf.h -- C ABI of the MI355X-native D3Q27 double-distribution
 * Rayleigh-Benard step (library: lbm_ddf_amd/_lib/liblbm_ddf.so).
 *
 * This header is the drop-in boundary for the single-GPU time step of
 * LBM_DDF.  The reference has no library: its interface is the header
 * src/rb3d.h (kernel prototypes :49-62, host functions :47,64-68), global
 * device pointers (src/globals.cu:28-30) and compile-time parameters
 * (src/parameters.h:5-31).  Each entry point below names the reference
 * interface it replaces.
 *
 * Conventions (differences from the reference are deliberate):
 *   - every function returns 0 on success and a negative LBM_E* code on
 *     failure; lbm_last_error() gives the message.  The library never calls
 *     exit() (the reference's CHECK_CUDA_ERROR exits, src/rb3d.h:71-79);
 *   - an opaque context replaces the global pointers and __constant__
 *     symbols; it owns its device memory and one HIP stream;
 *   - parameters are runtime values per context;
 *   - host buffers are caller-owned; device memory is library-owned;
 *   - a context is not thread-safe; distinct contexts are independent.
 *
 * Host copies of populations (lbm_get_pops / lbm_set_pops, checkpoints) use
 * the reference layout [ip][iz][iy][ix], x fastest (src/collision.cu:10,
 * src/streaming.cu:9).  On the device the library chooses per cavity between
 * that layout (planes staggered in HBM) and a row-blocked one (the 54
 * populations of an x-line side by side); offsets are 64-bit (the
 * reference's int offsets overflow at 512^3).
 */
#ifndef LBM_DDF_H
#define LBM_DDF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_E_INVALID (-1)   /* bad argument */
#define LBM_E_HIP (-2)       /* HIP runtime error */
#define LBM_E_NOMEM (-3)     /* device allocation failed */
#define LBM_E_STATE (-4)     /* call not valid in the context's state */
#define LBM_E_UNSUPPORTED (-5)
#define LBM_E_IO (-6)

/* Storage precision of the populations. */
#define LBM_PREC_FP64 0 /* FP64 storage, FP64 arithmetic */
#define LBM_PREC_FP16 1 /* FP16 deviations (scaled), FP32 arithmetic;
                           needs |t_hot|, |t_cold| <= 1 */
#define LBM_PREC_BF16 2 /* BF16 deviations, FP32 arithmetic */

/* Field selectors for lbm_get_field / lbm_set_field
 * (the reference arrays d_rho, d_ux, d_uy, d_uz, d_phi, d_force_real*). */
#define LBM_FIELD_RHO 0
#define LBM_FIELD_UX 1
#define LBM_FIELD_UY 2
#define LBM_FIELD_UZ 3
#define LBM_FIELD_PHI 4
#define LBM_FIELD_FX 5
#define LBM_FIELD_FY 6
#define LBM_FIELD_FZ 7

#define LBM_POP_F 0
#define LBM_POP_G 1

/* Population storage pattern (no reference counterpart; the reference
 * keeps f, g and the streaming targets f_temp, g_temp, src/globals.cu:28-30). */
#define LBM_PATTERN_AB 0 /* two buffers per distribution, pull streaming */
#define LBM_PATTERN_AA 1 /* one buffer per distribution, updated in place by
                            alternating even/odd passes: half the population
                            memory; FP64 results equal A/B bit for bit (16-bit
                            storage rounds differently: close) */

/* Physical and lattice parameters: the macros of src/parameters.h:8-26
 * made runtime.  beta <= 0 selects the reference default 0.1/(5*LY).
 * pattern: LBM_PATTERN_* (0, the default of a zero-initialised tail, is the
 * A/B pattern). */
typedef struct lbm_params {
    int32_t lx, ly, lz;
    int32_t precision; /* LBM_PREC_* */
    double rayleigh;   /* RAYLEIGH */
    double prandtl;    /* PRANDTL */
    double t_hot;      /* T_HOT: wall at iy = 0 */
    double t_cold;     /* T_COLD: wall at iy = LY-1 */
    double gravity;    /* GRAVITY */
    double beta;       /* BETA */
    int32_t pattern;   /* LBM_PATTERN_* */
    int32_t reserved;  /* must be 0 */
} lbm_params;

/* Derived parameters, src/globals.cu:11-23. */
typedef struct lbm_derived {
    double visc, tau, u0, diff, tauc, t0;
} lbm_derived;

/* Diagnostics of one completed step.
 * nu_*: src/output.cu:199-256; *mean: src/output.cu:24-34;
 * kinetic_energy = 0.5*<|u|^2>, mass = sum(rho'), max_speed = max|u| and
 * nonfinite (count of non-finite moments) have no reference counterpart. */
typedef struct lbm_diag {
    int64_t step;
    double nu_bulk, nu_top, nu_bottom;
    double umean, vmean, wmean, tmean;
    double kinetic_energy, mass;
    double max_speed;
    double nonfinite;
} lbm_diag;

typedef struct lbm_ctx lbm_ctx;

/* Message of the last failed call on this thread. */
const char *lbm_last_error(void);

/* src/globals.cu:11-23 (host-side parameter derivation). */
int lbm_derive(const lbm_params *p, lbm_derived *out);

/* Device bytes a context with these parameters allocates at creation:
 * populations, fields, side values, diagnostic partials and records, each
 * allocation rounded up to 4 KiB.  The first host copy of populations
 * (lbm_get_pops, lbm_set_pops, checkpoints) adds a staging plane of LX*LY*LZ
 * doubles. */
int lbm_bytes_required(const lbm_params *p, uint64_t *bytes);

/* initialize(), src/initial.cu:3-67 (allocation) and :86-99 (constants).
 * Allocates on the current HIP device; the state is undefined until an
 * lbm_init_* call. */
int lbm_create(const lbm_params *p, lbm_ctx **out);
/* The frees of src/main.cu:98-117. */
int lbm_destroy(lbm_ctx *ctx);

/* init_arrays, src/initial.cu:208-242: rho'=0, u=0, phi=0.5, f=0, g=w*phi. */
int lbm_init_arrays(lbm_ctx *ctx);
/* Conductive start (no reference counterpart in src/; the profile is
 * Parallel/initial.cu:89): T(iy) = T_hot - (T_hot-T_cold)*iy/(LY-1) plus
 * amp*U(-1,1) from a counter-based generator keyed by (seed, cell index);
 * rho'=0, u=0, f=0, g=w*phi. */
int lbm_init_conductive(lbm_ctx *ctx, uint64_t seed, double amp);

/* Stage interface (FP64 contexts only), one call per reference kernel
 * launch of src/main.cu:27-45.  Each acts on the reference state.
 *   lbm_forcing              forcing,              src/macrovar.cu:36-47
 *   lbm_collision_BGK        collision_BGK,        src/collision.cu:3-50
 *   lbm_collision_BGK_scalar collision_BGK_scalar, src/collision.cu:52-113
 *   lbm_streaming            streaming + copy-back, src/streaming.cu:3-125, src/main.cu:39
 *   lbm_streaming_scalar     streaming_scalar + copy-back, src/streaming.cu:127-251, src/main.cu:41
 *   lbm_macrovar             macrovar,             src/macrovar.cu:3-34 */
int lbm_forcing(lbm_ctx *ctx);
int lbm_collision_BGK(lbm_ctx *ctx);
int lbm_collision_BGK_scalar(lbm_ctx *ctx);
int lbm_streaming(lbm_ctx *ctx);
int lbm_streaming_scalar(lbm_ctx *ctx);
int lbm_macrovar(lbm_ctx *ctx);
/* nsteps iterations of src/main.cu:27-48 exactly as the reference issues
 * them: the six stage launches with a host synchronization after the
 * collisions, after the streamings and after macrovar (FP64 only). */
int lbm_step_staged(lbm_ctx *ctx, int64_t nsteps);

/* nsteps iterations of the loop body of src/main.cu:27-48, each one fused
 * collide-and-stream launch for f and g.  Asynchronous on the context's
 * stream.  The last step's streaming half stays pending until a call that
 * needs the reference state (lbm_finalize, lbm_get_*, lbm_reduce, the stage
 * calls) completes it. */
int lbm_step(lbm_ctx *ctx, int64_t nsteps);
/* lbm_step, plus the elapsed device time of the nsteps fused launches
 * (HIP events on the context's stream); waits for completion. */
int lbm_step_timed(lbm_ctx *ctx, int64_t nsteps, double *elapsed_ms);
/* lbm_step with the diagnostics of every completed step s with
 * s % every == 0 folded into the fused launch (output_nu / diag_flow,
 * src/main.cu:64-73).  Writes up to cap records; *nout receives the count.
 * Ends with the reference state complete (as lbm_finalize). */
int lbm_run_diag(lbm_ctx *ctx, int64_t nsteps, int64_t every, lbm_diag *out, int64_t cap,
                 int64_t *nout);
/* Complete a pending streaming half-step and the moments (macrovar). */
int lbm_finalize(lbm_ctx *ctx);
/* Diagnostics of the current state (output_nu, diag_flow; src/output.cu). */
int lbm_reduce(lbm_ctx *ctx, lbm_diag *out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, serving loops that must not touch the
 * allocator, harnesses that forbid it): the diagnostics of lbm_reduce over
 * moment fields the caller holds in device memory (rho', u_x, u_y, u_z, phi:
 * LX*LY*LZ doubles each, x fastest), written as one lbm_diag to `out` in
 * device memory with out->step = step, asynchronously on `stream` (a
 * hipStream_t; NULL: the null stream).  Only p's geometry, walls, Ra, Pr,
 * gravity and beta are used.  Nothing is allocated: the partial sums live in
 * a fixed device array of the library (8.1 MiB on each GPU it runs on, held
 * until the process ends), which bounds the partial records
 * ceil(LX/B)*LY*LZ, B = 64, 128 or 256 for LX <= 64, <= 128 or larger, to
 * LBM_NOALLOC_MAX_RECORDS (LY*LZ <= 65536 for LX <= 256); larger shapes
 * return LBM_E_INVALID.  The record equals lbm_reduce's for the same fields
 * bit for bit.  Calls that overlap on one GPU share the array and race; the
 * first call on a device resolves the array's address (make it before any
 * capture). */
#define LBM_NOALLOC_MAX_RECORDS 65536
int lbm_reduce_fields_noalloc(const lbm_params *p, const double *rho, const double *ux, const double *uy,
                              const double *uz, const double *phi, int64_t step, lbm_diag *out, void *stream);
/* Per-y-plane <ux>,<uy>,<uz>,<phi>,phi_rms (output_profile,
 * src/output.cu:282-377); out holds 5*LY doubles. */
int lbm_profile(lbm_ctx *ctx, double *out);
/* Number of completed reference steps. */
int lbm_step_count(const lbm_ctx *ctx, int64_t *out);
/* Wait for all work on the context's stream. */
int lbm_synchronize(lbm_ctx *ctx);

/* Launch-shape tuning of the fused pass (no reference counterpart; the
 * reference fixes an 8x8x4 block, src/parameters.h:29-31).  block_x: threads
 * per block along x (64, 128, 256 or 512); flags: bit 0 non-temporal stores,
 * bit 1 non-temporal loads of unshifted populations, bit 2 XCD-aware block
 * order.  Every setting computes the same values bit for bit. */
int lbm_set_tuning(lbm_ctx *ctx, int32_t block_x, int32_t flags);
int lbm_get_tuning(const lbm_ctx *ctx, int32_t *block_x, int32_t *flags);
/* Times every valid (block_x, flags) on the live state for `steps` fused
 * steps each and keeps the fastest.  Advances the simulation by the steps it
 * runs (*steps_taken); results are unaffected by the choice. */
int lbm_autotune(lbm_ctx *ctx, int32_t steps, int64_t *steps_taken, double *best_ms_per_step);
/* Device-to-device copy bandwidth (read + write bytes per second / 1e9) of a
 * plain copy kernel over two buffers of `bytes`: the achievable-bandwidth
 * reference for the roofline fraction. */
int lbm_copy_bandwidth(uint64_t bytes, int32_t reps, double *gbps);

/* Access-pattern copies (no reference counterpart): copy kernels that move
 * the 54 populations of every cell (elem_bytes 8: FP64 storage; 2: 16-bit
 * storage) in the row-blocked layout of a single cavity -- stream s (f_i at
 * s = i, g_i at s = 27 + i) of cell (ix, iy, iz) at element
 * ((iz*ly + iy)*54 + s)*lx + ix -- along the addresses of one step pass, with
 * no arithmetic: the achievable-bandwidth reference of that pass.  Buffers
 * hold 54*lx*ly*lz elements; the in-place patterns take dst == src, the
 * others dst != src.  block_x: 128, 256 or 512; flags: 3 or 7
 * (lbm_set_tuning's bits: non-temporal stores and non-temporal unshifted
 * loads always, bit 2 the XCD-aware block order).  lbm_pattern_copy enqueues
 * one pass on `stream` (a hipStream_t; NULL: the default stream).
 * lbm_pattern_bandwidth allocates its own buffers, runs `reps` passes after
 * two untimed ones and returns the mean time of a pass in milliseconds. */
#define LBM_ACCESS_COPY 0         /* own slots -> own slots of a second buffer */
#define LBM_ACCESS_PULL 1         /* the A/B pass: gather from x - c_i, wall slots bounced */
#define LBM_ACCESS_PUSH 2         /* own slots -> scatter to x + c_j of a second buffer */
#define LBM_ACCESS_INPLACE_EVEN 3 /* the in-place even pass: own slots, one buffer */
#define LBM_ACCESS_INPLACE_ODD 4  /* the in-place odd pass: neighbours' slots, one buffer */
int lbm_pattern_copy(int32_t pattern, int32_t elem_bytes, int32_t lx, int32_t ly, int32_t lz, int32_t block_x,
                     int32_t flags, const void *src, void *dst, void *stream);
int lbm_pattern_bandwidth(int32_t pattern, int32_t elem_bytes, int32_t lx, int32_t ly, int32_t lz, int32_t block_x,
                          int32_t flags, int32_t reps, double *ms_per_pass);
/* The access-pattern copy of the context's own fused pass, timed on the
 * context's own population buffers, stream and launch shape (block_x outside
 * 128/256/512 becomes the nearest of them; flags 7 when the context's bit 2
 * is set, else 3), so that the fused kernel and its pattern copy read and
 * write the same physical pages.  ms_per_pass[0]: A/B, the pull gather from
 * buffer 0 into buffer 1; in place, the even pass.  ms_per_pass[1]: A/B, the
 * pull back from buffer 1 into buffer 0; in place, the odd pass.  Each is the
 * mean over `reps` passes after two untimed ones.  The copies overwrite the
 * populations: afterwards the context has no state (lbm_init_* before
 * stepping again).  Single-cavity contexts only (LBM_E_STATE for an ensemble
 * member). */
int lbm_pattern_time(lbm_ctx *ctx, int32_t reps, double *ms_per_pass);

/* ---- Checkpoints in the reference's binary formats ------------------------
 * lbm_output_flow: output_flow (src/output.cu:67-146): rho', u_x, u_y, u_z,
 *   phi, each LX*LY*LZ doubles, x fastest, native byte order.
 * lbm_output_fg:   output_fg (src/output.cu:379-419): f then g, 27*LX*LY*LZ
 *   doubles each, [ip][iz][iy][ix].
 * lbm_restart:     the CONTINUE_STEP branch of initialize()
 *   (src/initial.cu:110-205): reads both files, recomputes F from phi
 *   (:202) and sets the step count; the run then continues bit for bit.
 * Files always hold FP64 values (16-bit contexts convert). */
int lbm_output_flow(lbm_ctx *ctx, const char *path);
int lbm_output_fg(lbm_ctx *ctx, const char *path);
int lbm_restart(lbm_ctx *ctx, const char *flow_path, const char *fg_path, int64_t step);

/* ---- Ensembles: many cavities, one launch per step ------------------------
 * No reference counterpart (the reference runs one cavity per process).
 * Cavities of different size, Ra, Pr and walls advance together: one
 * persistent launch per step over all resident cavities, step sequences
 * replayed from a captured hipGraph.  Each cavity's result after k of its own
 * steps equals a standalone lbm_step run of k steps bit for bit.  Cavities
 * join (lbm_ensemble_add: conductive start) and leave (lbm_ensemble_remove)
 * between calls.  A member context may be read through the single-cavity
 * calls (lbm_get_field, lbm_reduce, ...) and re-initialised, but not stepped
 * or destroyed on its own: lbm_step, lbm_step_timed, lbm_run_diag,
 * lbm_autotune, lbm_step_staged, the stage calls and lbm_destroy return
 * LBM_E_STATE for a member. */
typedef struct lbm_ensemble lbm_ensemble;
int lbm_ensemble_create(int32_t precision, lbm_ensemble **out);
int lbm_ensemble_destroy(lbm_ensemble *ens);
/* Allocates a cavity (params->precision must match) with the conductive
 * start of lbm_init_conductive; *slot receives its index.  params->pattern
 * must match the resident members' (the first cavity of an empty ensemble
 * sets it): in-place members need half the population memory, so more of
 * them fit a budget; each member equals its standalone run of the same
 * pattern bit for bit.  A member has fewer than 2^31 cells (LBM_E_INVALID
 * otherwise; lbm_serve rejects such a request at admission). */
int lbm_ensemble_add(lbm_ensemble *ens, const lbm_params *params, uint64_t seed, double amp, int32_t *slot);
int lbm_ensemble_remove(lbm_ensemble *ens, int32_t slot);
int lbm_ensemble_count(const lbm_ensemble *ens, int32_t *slots, int32_t *resident, uint64_t *cells);
/* The member context in a slot (owned by the ensemble). */
int lbm_ensemble_cavity(lbm_ensemble *ens, int32_t slot, lbm_ctx **ctx);
/* nsteps launches; a cavity still in the reference state first gets its
 * prologue (its own step count then advances by nsteps + 1). */
int lbm_ensemble_step(lbm_ensemble *ens, int64_t nsteps);
int lbm_ensemble_step_timed(lbm_ensemble *ens, int64_t nsteps, double *elapsed_ms);
/* nsteps launches, the last with every cavity's diagnostics folded in;
 * out[slot] receives them (step = the cavity's completed step).  cap >= slots. */
int lbm_ensemble_run_diag(lbm_ensemble *ens, int64_t nsteps, lbm_diag *out, int32_t cap);
/* block_x (64, 128, 256) threads per block, blocks_per_cu of the persistent
 * grid (0: no persistent grid, one block per tile), graph_pairs step pairs
 * per graph replay (0: no graph), flags 0, 3
 * (non-temporal hints), 4 (XCD-contiguous tile order) or 7 (both),
 * cells_per_thread (1, 2, 4): a tile is block_x * cells_per_thread
 * consecutive cells of one cavity.  Results do not depend on the choice. */
int lbm_ensemble_set_tuning(lbm_ensemble *ens, int32_t block_x, int32_t blocks_per_cu, int32_t graph_pairs,
                            int32_t flags, int32_t cells_per_thread);
int lbm_ensemble_get_tuning(const lbm_ensemble *ens, int32_t *block_x, int32_t *blocks_per_cu, int32_t *graph_pairs,
                            int32_t *flags, int32_t *cells_per_thread);
/* Times each (block_x, blocks_per_cu, flags) shape at one cell per thread for
 * `steps` steps on the live ensemble, then 2 and 4 cells per thread for the
 * four fastest shapes, re-times the four fastest candidates (twice, over
 * 3*steps steps each) and keeps the fastest of those; with cache_path, a
 * result stored for the same precision, storage pattern and cavity-size multiset is reused
 * without timing. */
int lbm_ensemble_autotune(lbm_ensemble *ens, int32_t steps, const char *cache_path, double *best_ms_per_step);

/* ---- Serving: continuous batching of arriving solves ----------------------
 * No reference counterpart.  Each request is one cavity (conductive start);
 * lbm_serve runs the requests in real time on one GPU: between step chunks
 * it admits every arrived request that fits the HBM budget (first fit in
 * arrival order), advances all resident cavities together (ensemble, one
 * launch per step), checks each request's stopping rule and retires finished
 * ones.  Stopping rule: every `window` of its own steps a request samples
 * Nu_bulk; its running Nu is the mean of the samples; it is converged when the
 * running Nu moves by less than nu_tol between checks (>= 3 samples), else it
 * stops at max_steps.  A request's result does not depend on the mix. */
#define LBM_REQ_CONVERGED 0  /* stopping rule met */
#define LBM_REQ_BUDGET 1     /* max_steps reached */
#define LBM_REQ_REJECTED 2   /* alone larger than the HBM budget */
#define LBM_REQ_DIVERGED 3   /* non-finite moment or max|u| above the sound speed
                                1/sqrt(3): retired, others continue */
#define LBM_REQ_UNFINISHED 4 /* admitted, time limit reached first */
#define LBM_REQ_PENDING 5    /* never admitted before the time limit */

typedef struct lbm_request {
    int32_t lx, ly, lz;
    int32_t reserved;
    double rayleigh, prandtl;
    double t_hot, t_cold;
    uint64_t seed;
    double amp;        /* perturbation amplitude of the conductive start */
    double arrival_s;  /* seconds after the start of lbm_serve */
    int64_t max_steps;
    double nu_tol;
    int64_t window;    /* multiple of lbm_serve_options.chunk_steps */
} lbm_request;

typedef struct lbm_result {
    int32_t status; /* LBM_REQ_* */
    int32_t reserved;
    double arrival_s, admit_s, finish_s;
    int64_t steps;
    double nu_bulk, nu_top, nu_bottom, kinetic_energy; /* at the last check */
    double nu_running;
} lbm_result;

typedef struct lbm_serve_options {
    int32_t precision;
    int32_t max_resident; /* 0: no limit besides memory */
    int64_t chunk_steps;  /* steps between admissions/checks */
    uint64_t hbm_budget;  /* bytes for resident cavities */
    double time_limit_s;
    int32_t block_x, blocks_per_cu, flags; /* ensemble launch shape; block_x 0: default */
    int32_t cells_per_thread;              /* with block_x > 0: 1, 2 or 4 (0: 1) */
    uint64_t max_resident_cells; /* 0: no cap; else admit while resident cells stay below it
                                    (enough cells to saturate HBM, so admitted requests finish) */
    double progress_s;           /* > 0: the loop prints a progress line to stderr at this interval */
    int32_t pattern;             /* LBM_PATTERN_AB (0) or LBM_PATTERN_AA: in place, half the
                                    population bytes per request (FP64: same results) */
    int32_t reserved;            /* must be 0 */
} lbm_serve_options;

typedef struct lbm_serve_stats {
    double wall_s, busy_s;
    double cell_updates, mlups;
    int32_t completed, rejected, diverged, unfinished, pending, peak_resident;
    uint64_t peak_bytes;
    double p50_s, p99_s; /* time to result (finish - arrival) of completed requests */
    /* Every request that arrived before the end of the run counts below:
     * requests still resident or queued at the end with finish = the end of
     * the run (right-censored: the percentiles are lower bounds of the true
     * ones), diverged and rejected ones with their retirement time. */
    int32_t arrived;
    int32_t deferred;            /* admissions put back because the device allocation
                                    failed although the budget had room */
    double p50_all_s, p99_all_s; /* time to result over the arrived requests */
    /* Means over completed requests: time queued (admit - arrival) and time
     * resident (finish - admit). */
    double mean_queue_s, mean_service_s;
} lbm_serve_stats;

int lbm_serve(const lbm_request *reqs, int32_t n, const lbm_serve_options *opt, lbm_result *results,
              lbm_serve_stats *stats);

/* Host copies of the reference state (the D2H copies of src/output.cu).
 * Fields: LX*LY*LZ doubles; populations: 27*LX*LY*LZ doubles. */
int lbm_get_field(lbm_ctx *ctx, int32_t field, double *host);
int lbm_set_field(lbm_ctx *ctx, int32_t field, const double *host);
int lbm_get_pops(lbm_ctx *ctx, int32_t which, double *host);
int lbm_set_pops(lbm_ctx *ctx, int32_t which, const double *host);

#ifdef __cplusplus
}
#endif
#endif
