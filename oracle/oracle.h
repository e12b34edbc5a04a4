/*
 * oracle.h -- CPU FP64 restatement of LBM_DDF's single-GPU D3Q27
 * double-distribution Rayleigh-Benard step (reference tree `src/`).
 *
 * TEST INFRASTRUCTURE ONLY.  This library gives the expected results of the
 * parity tests and the CPU baseline.  Only tests/, __graft_entry__.smoke()
 * and bench.py's CPU baseline legs may load it; the product library
 * (lbm_ddf_amd) never links, loads or calls it.
 *
 * It is a plain C transcription written from reading the reference .cu files
 * as text (the reference is CUDA for sm_80 and cannot be built in this image).
 * Every function names the reference file:line it follows.  Arithmetic is
 * kept in the reference's operation order and the library is compiled with
 * -ffp-contract=off, so that a GPU kernel that keeps the same order can match
 * it bit for bit.
 *
 * Differences from the reference, all deliberate:
 *   - physical/lattice parameters are runtime values (struct orc_params), not
 *     macros (src/parameters.h:8-26); they are derived exactly as
 *     src/globals.cu:11-23 derives them;
 *   - 64-bit cell offsets (the reference's `ip*LXYZ` int product overflows at
 *     512^3, src/parameters.h:11 with src/collision.cu:29);
 *   - an extra initial state (orc_init_conductive): the linear conductive
 *     temperature profile of Parallel/initial.cu:89 with a seeded uniform
 *     perturbation; the reference's own fresh state is orc_init_arrays.
 *   - reductions that the reference does not have (kinetic energy, mass).
 *
 * Pinning status (see tests/test_oracle_*.py): every function is pinned by a
 * closed form, an invariant, or an independent implementation; none is
 * "parity unpinned" except where noted in the function comment.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NPOP 27

/* Physical and lattice parameters of one cavity.
 * Inputs: lx..beta.  Derived by orc_derive: visc..t0. */
typedef struct orc_params {
    int lx, ly, lz;
    double rayl, prand;
    double t_hot, t_cold;
    double grav0, beta;
    /* derived */
    double visc, tau, u0, diff, tauc, t0;
} orc_params;

/* All per-cell arrays of the reference state (src/globals.cu:28-30).
 * f, g, f_tmp, g_tmp hold 27*N doubles laid out [ip][iz][iy][ix];
 * the others hold N doubles laid out [iz][iy][ix]. */
typedef struct orc_fields {
    double *f, *g, *f_tmp, *g_tmp;
    double *rho, *ux, *uy, *uz, *phi;
    double *fx, *fy, *fz;
} orc_fields;

/* D3Q27 tables, src/initial.cu:71-83. */
void orc_tables(int *cix, int *ciy, int *ciz, double *tp);

/* src/globals.cu:11-23 (beta default src/parameters.h:26 when beta <= 0). */
void orc_derive(orc_params *p);

/* src/initial.cu:208-242: rho'=0, u=0, phi=0.5, F_y, f=0, g=w*phi. */
void orc_init_arrays(const orc_params *p, const orc_fields *s);

/* Conductive start: T(iy) = T_hot - (T_hot-T_cold)*iy/(LY-1)
 * (Parallel/initial.cu:89) plus amp*U(-1,1) from splitmix64(seed, idx);
 * rho'=0, u=0, f=0, g=w*phi, F_y from phi (forcing rule, src/macrovar.cu:44-46). */
void orc_init_conductive(const orc_params *p, const orc_fields *s,
                         uint64_t seed, double amp);

/* The perturbation draw used by orc_init_conductive, for one cell. */
double orc_perturbation(uint64_t seed, uint64_t idx, double amp);

/* src/macrovar.cu:36-47 */
void orc_forcing(const orc_params *p, const orc_fields *s);
/* src/collision.cu:3-32 (the discarded stress tensor :19-20,34-49 is dead) */
void orc_collision_bgk(const orc_params *p, const orc_fields *s);
/* src/collision.cu:52-81 */
void orc_collision_bgk_scalar(const orc_params *p, const orc_fields *s);
/* src/streaming.cu:3-125, then the copy-back of src/main.cu:39 */
void orc_streaming(const orc_params *p, const orc_fields *s);
/* src/streaming.cu:127-251, then the copy-back of src/main.cu:41 */
void orc_streaming_scalar(const orc_params *p, const orc_fields *s);
/* src/macrovar.cu:3-34 */
void orc_macrovar(const orc_params *p, const orc_fields *s);

/* nsteps iterations of the loop body of src/main.cu:27-48. */
void orc_run(const orc_params *p, const orc_fields *s, long nsteps);

/* src/output.cu:199-256: out = {Nu_bulk, Nu_top, Nu_bottom}. */
void orc_output_nu(const orc_params *p, const orc_fields *s, double out[3]);
/* src/output.cu:24-34: out = {umean, vmean, wmean, tmean}. */
void orc_diag_flow(const orc_params *p, const orc_fields *s, double out[4]);
/* src/output.cu:336-358: out[iy*5 + k] = {<ux>,<uy>,<uz>,<phi>,phi_rms}. */
void orc_output_profile(const orc_params *p, const orc_fields *s, double *out);
/* No reference counterpart: out = {KE = 0.5*mean(|u|^2), mass = sum(rho')}. */
void orc_energy_mass(const orc_params *p, const orc_fields *s, double out[2]);

#ifdef __cplusplus
}
#endif
#endif
