/*
 * oracle.c -- CPU FP64 restatement of LBM_DDF src/ (see oracle.h).
 * TEST INFRASTRUCTURE ONLY: the expected results of the parity tests and the
 * CPU baseline.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off -fPIC -shared).
 * Cells are processed in parallel with OpenMP where each cell's result
 * depends only on the previous stage's arrays, so the result does not depend
 * on the thread count.  Reductions run serially in the reference's order.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* src/initial.cu:71-73 */
static const int CIX[ORC_NPOP] = {0, 1, 0, -1, 0, 0, 0, 1, -1, -1, 1, 1, 0, -1, 0, 1, 0, -1, 0, 1, -1, -1, 1, 1, -1, -1, 1};
static const int CIY[ORC_NPOP] = {0, 0, 1, 0, -1, 0, 0, 1, 1, -1, -1, 0, 1, 0, -1, 0, 1, 0, -1, 1, 1, -1, -1, 1, 1, -1, -1};
static const int CIZ[ORC_NPOP] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1};

/* src/initial.cu:75-83 */
static double TP[ORC_NPOP];
static int tables_ready = 0;

static void init_tables(void) {
    if (tables_ready) return;
    for (int i = 0; i < ORC_NPOP; i++) {
        if (i == 0) TP[i] = 8.0 / 27.0;
        else if (i <= 6) TP[i] = 2.0 / 27.0;
        else if (i <= 18) TP[i] = 1.0 / 54.0;
        else TP[i] = 1.0 / 216.0;
    }
    tables_ready = 1;
}

void orc_tables(int *cix, int *ciy, int *ciz, double *tp) {
    init_tables();
    for (int i = 0; i < ORC_NPOP; i++) {
        cix[i] = CIX[i];
        ciy[i] = CIY[i];
        ciz[i] = CIZ[i];
        tp[i] = TP[i];
    }
}

/* src/globals.cu:11-23; BETA default from src/parameters.h:26. */
void orc_derive(orc_params *p) {
    init_tables();
    if (!(p->beta > 0.0)) p->beta = 0.1 / ((double)p->ly * 5.0);
    p->visc = (double)p->ly * sqrt(p->grav0 * p->beta * (p->t_hot - p->t_cold) * (double)p->ly * p->prand / p->rayl);
    p->tau = 3.0 * p->visc + 0.5;
    p->u0 = sqrt(p->grav0 * p->beta * (p->t_hot - p->t_cold) * (double)p->ly);
    p->diff = p->visc / p->prand;
    p->tauc = 3.0 * p->diff + 0.5;
    p->t0 = 0.5 * (p->t_cold + p->t_hot);
}

static inline size_t ncell(const orc_params *p) {
    return (size_t)p->lx * (size_t)p->ly * (size_t)p->lz;
}

/* Counter-based draw: splitmix64 finaliser of seed + (idx+1)*golden gamma,
 * top 53 bits as u in [0,1), returned as amp*(2u-1).  The scaling steps are
 * exact in binary64 except the final multiply by amp. */
double orc_perturbation(uint64_t seed, uint64_t idx, double amp) {
    uint64_t z = seed + (idx + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    double u = (double)(z >> 11) * 0x1.0p-53;
    return amp * (2.0 * u - 1.0);
}

/* Shared tail of both initial states: src/initial.cu:218-241 given phi. */
static void init_common(const orc_params *p, const orc_fields *s, size_t idx, double phi) {
    const size_t n = ncell(p);
    s->rho[idx] = 0.0;
    s->ux[idx] = 0.0;
    s->uy[idx] = 0.0;
    s->uz[idx] = 0.0;
    s->phi[idx] = phi;
    s->fx[idx] = 0.0;
    s->fy[idx] = p->grav0 * p->beta * (s->phi[idx] - p->t0);
    s->fz[idx] = 0.0;
    double usq = 0.0;
    for (int ip = 0; ip < ORC_NPOP; ip++) {
        double feq = TP[ip] * s->rho[idx] * (1.0 - 1.5 * usq);
        s->f[(size_t)ip * n + idx] = feq - 0.5 * s->fy[idx] * CIY[ip] / (1.0 / 3.0) * feq;
        double geq = TP[ip] * s->phi[idx] * (1.0 - 1.5 * usq);
        s->g[(size_t)ip * n + idx] = geq;
    }
}

/* src/initial.cu:208-242 */
void orc_init_arrays(const orc_params *p, const orc_fields *s) {
    init_tables();
    const int LX = p->lx, LY = p->ly, LZ = p->lz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < LZ; iz++)
        for (int iy = 0; iy < LY; iy++)
            for (int ix = 0; ix < LX; ix++) {
                size_t idx = ((size_t)iz * LY + iy) * LX + ix;
                init_common(p, s, idx, 0.5);
            }
}

/* Linear profile of Parallel/initial.cu:89 plus the seeded perturbation,
 * then the src/ init semantics (rho'=0, f=0, g=w*phi). */
void orc_init_conductive(const orc_params *p, const orc_fields *s, uint64_t seed, double amp) {
    init_tables();
    const int LX = p->lx, LY = p->ly, LZ = p->lz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = 0; iz < LZ; iz++)
        for (int iy = 0; iy < LY; iy++)
            for (int ix = 0; ix < LX; ix++) {
                size_t idx = ((size_t)iz * LY + iy) * LX + ix;
                double t = p->t_hot - (p->t_hot - p->t_cold) * iy / (LY - 1.0);
                double phi = t + orc_perturbation(seed, (uint64_t)idx, amp);
                init_common(p, s, idx, phi);
            }
}

/* src/macrovar.cu:36-47 */
void orc_forcing(const orc_params *p, const orc_fields *s) {
    const size_t n = ncell(p);
#pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < n; idx++) {
        s->fx[idx] = 0.0;
        s->fy[idx] = p->grav0 * p->beta * (s->phi[idx] - p->t0);
        s->fz[idx] = 0.0;
    }
}

/* Cells are the innermost loop in the stage functions below: the
 * reference computes each cell independently, so iterating direction-major
 * changes no element's arithmetic and lets the compiler vectorise. */

/* src/collision.cu:3-32.  The stress tensor of :19-20,34-49 is computed and
 * discarded by the reference; it has no effect and is not restated. */
void orc_collision_bgk(const orc_params *p, const orc_fields *s) {
    init_tables();
    const size_t n = ncell(p);
    const double tau = p->tau;
    const double *restrict fy = s->fy, *restrict rho = s->rho;
    const double *restrict ux = s->ux, *restrict uy = s->uy, *restrict uz = s->uz;
#pragma omp parallel
    for (int ip = 0; ip < ORC_NPOP; ip++) {
        double *restrict f = s->f + (size_t)ip * n;
        const double cx = CIX[ip], cy = CIY[ip], cz = CIZ[ip], tp = TP[ip];
#pragma omp for schedule(static) nowait
        for (size_t idx = 0; idx < n; idx++) {
            double fy9 = fy[idx];
            double rho9 = rho[idx];
            double u9 = ux[idx];
            double v9 = uy[idx];
            double w9 = uz[idx];
            double RT = 1.0 / 3.0;
            double eu = (cx * u9 + cy * v9 + cz * w9) / RT;
            double uv = (u9 * u9 + v9 * v9 + w9 * w9) / RT;
            double feq = tp * (rho9 + eu + 0.5 * (eu * eu - uv));
            double fchange = -(f[idx] - feq) / tau;
            double feq_rho0 = tp * (1.0 + eu + 0.5 * (eu * eu - uv));
            fchange += (1.0 - 0.5 / tau) * fy9 * (cy - v9) / RT * feq_rho0;
            f[idx] += fchange;
        }
    }
}

/* src/collision.cu:52-81 (the source-term variant at :83-112 is commented
 * out in the reference and is not restated). */
void orc_collision_bgk_scalar(const orc_params *p, const orc_fields *s) {
    init_tables();
    const size_t n = ncell(p);
    const double tauc = p->tauc;
    const double *restrict phi = s->phi;
    const double *restrict ux = s->ux, *restrict uy = s->uy, *restrict uz = s->uz;
#pragma omp parallel
    for (int ip = 0; ip < ORC_NPOP; ip++) {
        double *restrict g = s->g + (size_t)ip * n;
        const double cx = CIX[ip], cy = CIY[ip], cz = CIZ[ip], tp = TP[ip];
#pragma omp for schedule(static) nowait
        for (size_t idx = 0; idx < n; idx++) {
            double phi9 = phi[idx];
            double u9 = ux[idx];
            double v9 = uy[idx];
            double w9 = uz[idx];
            double RT = 1.0 / 3.0;
            double g9 = g[idx];
            double eu = (cx * u9 + cy * v9 + cz * w9) / RT;
            double uv = (u9 * u9 + v9 * v9 + w9 * w9) / RT;
            double geq = tp * phi9 * (1.0 + eu + 0.5 * (eu * eu - uv));
            double gchange = -(g9 - geq) / tauc;
            g[idx] += gchange;
        }
    }
}

/* Wall blocks of src/streaming.cu, in the order the reference applies them:
 * x- (:59), x+ (:70), y- (:81), y+ (:92), z- (:103), z+ (:114).
 * Each pair is {destination direction, source direction}. */
static const int WALL_XLO[9][2] = {{1, 3}, {7, 9}, {10, 8}, {11, 17}, {19, 25}, {22, 24}, {15, 13}, {23, 21}, {26, 20}};
static const int WALL_XHI[9][2] = {{3, 1}, {9, 7}, {8, 10}, {17, 11}, {25, 19}, {24, 22}, {13, 15}, {21, 23}, {20, 26}};
static const int WALL_YLO[9][2] = {{2, 4}, {7, 9}, {8, 10}, {12, 18}, {19, 25}, {20, 26}, {16, 14}, {23, 21}, {24, 22}};
static const int WALL_YHI[9][2] = {{4, 2}, {9, 7}, {10, 8}, {18, 12}, {25, 19}, {26, 20}, {14, 16}, {21, 23}, {22, 24}};
static const int WALL_ZLO[9][2] = {{5, 6}, {11, 17}, {12, 18}, {13, 15}, {14, 16}, {19, 25}, {20, 26}, {21, 23}, {22, 24}};
static const int WALL_ZHI[9][2] = {{6, 5}, {17, 11}, {18, 12}, {15, 13}, {16, 14}, {25, 19}, {26, 20}, {23, 21}, {24, 22}};

/* Pull part shared by both streaming kernels, src/streaming.cu:20-56 and
 * :144-180: tmp_i(x) = src_i(x - c_i) where x - c_i lies in the domain.
 * Cells whose source lies outside are left for the wall blocks. */
static void pull_all(const orc_params *p, const double *src, double *tmp) {
    const int LX = p->lx, LY = p->ly, LZ = p->lz;
    const size_t n = ncell(p);
#pragma omp parallel for collapse(2) schedule(static)
    for (int ip = 0; ip < ORC_NPOP; ip++)
        for (int iz = 0; iz < LZ; iz++) {
            const int sz = iz - CIZ[ip];
            if (sz < 0 || sz >= LZ) continue;
            for (int iy = 0; iy < LY; iy++) {
                const int sy = iy - CIY[ip];
                if (sy < 0 || sy >= LY) continue;
                const int x0 = CIX[ip] > 0 ? 1 : 0;
                const int x1 = CIX[ip] < 0 ? LX - 1 : LX;
                const size_t drow = (size_t)ip * n + ((size_t)iz * LY + iy) * LX;
                const size_t srow = (size_t)ip * n + ((size_t)sz * LY + sy) * LX;
                for (int ix = x0; ix < x1; ix++) tmp[drow + ix] = src[srow + ix - CIX[ip]];
            }
        }
}

static inline void bounce_back(const double *src, double *tmp, size_t n, size_t idx, const int w[9][2]) {
    for (int k = 0; k < 9; k++) tmp[(size_t)w[k][0] * n + idx] = src[(size_t)w[k][1] * n + idx];
}

/* Isothermal anti-bounce-back of src/streaming.cu:205-228:
 * tmp_i = -src_opp + 2*w_i*T_w*(1 + rho'). */
static inline void anti_bounce_back(const double *src, double *tmp, size_t n, size_t idx,
                                    const int w[9][2], double t_wall, double density) {
    for (int k = 0; k < 9; k++) {
        int d = w[k][0], o = w[k][1];
        tmp[(size_t)d * n + idx] = -src[(size_t)o * n + idx] + 2.0 * TP[d] * t_wall * (1.0 + density);
    }
}

/* The six wall blocks over their planes, in the reference order.  A cell on
 * an edge or corner is visited by several blocks; the later block wins for
 * the directions both write, as in the reference. */
static void walls(const orc_params *p, const double *src, double *tmp, const double *rho, int scalar) {
    const int LX = p->lx, LY = p->ly, LZ = p->lz;
    const size_t n = ncell(p);
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < LZ; iz++)
        for (int iy = 0; iy < LY; iy++) {
            size_t row = ((size_t)iz * LY + iy) * LX;
            bounce_back(src, tmp, n, row, WALL_XLO);
            bounce_back(src, tmp, n, row + LX - 1, WALL_XHI);
        }
#pragma omp parallel for schedule(static)
    for (int iz = 0; iz < LZ; iz++)
        for (int ix = 0; ix < LX; ix++) {
            size_t lo = ((size_t)iz * LY + 0) * LX + ix;
            size_t hi = ((size_t)iz * LY + (LY - 1)) * LX + ix;
            if (scalar) {
                anti_bounce_back(src, tmp, n, lo, WALL_YLO, p->t_hot, rho[lo]);
                anti_bounce_back(src, tmp, n, hi, WALL_YHI, p->t_cold, rho[hi]);
            } else {
                bounce_back(src, tmp, n, lo, WALL_YLO);
                bounce_back(src, tmp, n, hi, WALL_YHI);
            }
        }
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < LY; iy++)
        for (int ix = 0; ix < LX; ix++) {
            size_t lo = ((size_t)0 * LY + iy) * LX + ix;
            size_t hi = ((size_t)(LZ - 1) * LY + iy) * LX + ix;
            bounce_back(src, tmp, n, lo, WALL_ZLO);
            bounce_back(src, tmp, n, hi, WALL_ZHI);
        }
}

/* src/streaming.cu:3-125 + copy-back src/main.cu:39.  The x+ block runs
 * after the x- block on the same cell when LX = 1, as in the reference. */
void orc_streaming(const orc_params *p, const orc_fields *s) {
    const size_t n = ncell(p);
    pull_all(p, s->f, s->f_tmp);
    walls(p, s->f, s->f_tmp, s->rho, 0);
    memcpy(s->f, s->f_tmp, (size_t)ORC_NPOP * n * sizeof(double));
}

/* src/streaming.cu:127-251 + copy-back src/main.cu:41.  Wall order x, y, z:
 * at y/z edges the adiabatic z rule overwrites the isothermal y rule. */
void orc_streaming_scalar(const orc_params *p, const orc_fields *s) {
    init_tables();
    const size_t n = ncell(p);
    pull_all(p, s->g, s->g_tmp);
    walls(p, s->g, s->g_tmp, s->rho, 1);
    memcpy(s->g, s->g_tmp, (size_t)ORC_NPOP * n * sizeof(double));
}

/* src/macrovar.cu:3-34: moments summed in direction order 0..26. */
void orc_macrovar(const orc_params *p, const orc_fields *s) {
    const size_t n = ncell(p);
    double *restrict rho = s->rho, *restrict ux = s->ux, *restrict uy = s->uy, *restrict uz = s->uz;
    double *restrict phi = s->phi;
#pragma omp parallel
    {
#pragma omp for schedule(static)
        for (size_t idx = 0; idx < n; idx++) {
            rho[idx] = 0.0;
            ux[idx] = 0.0;
            uy[idx] = 0.0;
            uz[idx] = 0.0;
            phi[idx] = 0.0;
        }
        for (int ip = 0; ip < ORC_NPOP; ip++) {
            const double *restrict f = s->f + (size_t)ip * n;
            const double *restrict g = s->g + (size_t)ip * n;
            const double cx = CIX[ip], cy = CIY[ip], cz = CIZ[ip];
#pragma omp for schedule(static)
            for (size_t idx = 0; idx < n; idx++) {
                double f_val = f[idx];
                double g_val = g[idx];
                rho[idx] += f_val;
                ux[idx] += cx * f_val;
                uy[idx] += cy * f_val;
                uz[idx] += cz * f_val;
                phi[idx] += g_val;
            }
        }
#pragma omp for schedule(static)
        for (size_t idx = 0; idx < n; idx++) {
            ux[idx] = ux[idx] + 0.5 * s->fx[idx];
            uy[idx] = uy[idx] + 0.5 * s->fy[idx];
            uz[idx] = uz[idx] + 0.5 * s->fz[idx];
        }
    }
}

/* Loop body of src/main.cu:27-48 (timing, output and syncs omitted). */
void orc_run(const orc_params *p, const orc_fields *s, long nsteps) {
    for (long k = 0; k < nsteps; k++) {
        orc_forcing(p, s);
        orc_collision_bgk(p, s);
        orc_collision_bgk_scalar(p, s);
        orc_streaming(p, s);
        orc_streaming_scalar(p, s);
        orc_macrovar(p, s);
    }
}

/* src/output.cu:199-256 (requires LY >= 3). */
void orc_output_nu(const orc_params *p, const orc_fields *s, double out[3]) {
    const int LX = p->lx, LY = p->ly, LZ = p->lz;
    const size_t n = ncell(p);
    const double *phi = s->phi;
    double *uyT = (double *)malloc(n * sizeof(double));
    double *Ty = (double *)malloc(n * sizeof(double));
    for (size_t i = 0; i < n; i++) uyT[i] = s->uy[i] * phi[i];
    for (int iz = 0; iz < LZ; iz++)
        for (int ix = 0; ix < LX; ix++)
            for (int iy = 0; iy < LY; iy++) {
                size_t idx = ((size_t)iz * LY + iy) * LX + ix;
                if (iy == 0) {
                    size_t p1 = ((size_t)iz * LY + (iy + 1)) * LX + ix;
                    size_t p2 = ((size_t)iz * LY + (iy + 2)) * LX + ix;
                    Ty[idx] = (-3.0 * phi[idx] + 4.0 * phi[p1] - phi[p2]) / 2.0;
                } else if (iy == LY - 1) {
                    size_t m1 = ((size_t)iz * LY + (iy - 1)) * LX + ix;
                    size_t m2 = ((size_t)iz * LY + (iy - 2)) * LX + ix;
                    Ty[idx] = (3.0 * phi[idx] - 4.0 * phi[m1] + phi[m2]) / 2.0;
                } else {
                    size_t p1 = ((size_t)iz * LY + (iy + 1)) * LX + ix;
                    size_t m1 = ((size_t)iz * LY + (iy - 1)) * LX + ix;
                    Ty[idx] = (phi[p1] - phi[m1]) / 2.0;
                }
            }
    double uyT_avg = 0.0;
    for (size_t i = 0; i < n; i++) uyT_avg += uyT[i];
    uyT_avg /= (double)n;
    double nu_bulk = 1.0 + uyT_avg * LY / (p->diff * (p->t_hot - p->t_cold));
    double top = 0.0, bottom = 0.0;
    long ctop = 0, cbot = 0;
    for (int ix = 0; ix < LX; ix++)
        for (int iz = 0; iz < LZ; iz++) {
            size_t it = ((size_t)iz * LY + (LY - 1)) * LX + ix;
            size_t ib = (size_t)iz * LY * LX + ix;
            top += Ty[it];
            bottom += Ty[ib];
            ctop++;
            cbot++;
        }
    top /= ctop;
    bottom /= cbot;
    out[0] = nu_bulk;
    out[1] = -top * LY / (p->t_hot - p->t_cold);
    out[2] = -bottom * LY / (p->t_hot - p->t_cold);
    free(uyT);
    free(Ty);
}

/* src/output.cu:24-34 */
void orc_diag_flow(const orc_params *p, const orc_fields *s, double out[4]) {
    const size_t n = ncell(p);
    double um = 0.0, vm = 0.0, wm = 0.0, tm = 0.0;
    for (size_t i = 0; i < n; i++) {
        um += s->ux[i];
        vm += s->uy[i];
        wm += s->uz[i];
        tm += s->phi[i];
    }
    out[0] = um / (double)n;
    out[1] = vm / (double)n;
    out[2] = wm / (double)n;
    out[3] = tm / (double)n;
}

/* src/output.cu:336-358 */
void orc_output_profile(const orc_params *p, const orc_fields *s, double *out) {
    const int LX = p->lx, LY = p->ly, LZ = p->lz;
    for (int iy = 0; iy < LY; iy++) {
        double su = 0.0, sv = 0.0, sw = 0.0, sp = 0.0, sq = 0.0;
        int count = 0;
        for (int ix = 0; ix < LX; ix++)
            for (int iz = 0; iz < LZ; iz++) {
                size_t idx = ((size_t)iz * LY + iy) * LX + ix;
                su += s->ux[idx];
                sv += s->uy[idx];
                sw += s->uz[idx];
                sp += s->phi[idx];
                sq += s->phi[idx] * s->phi[idx];
                count++;
            }
        double ap = sp / count;
        out[iy * 5 + 0] = su / count;
        out[iy * 5 + 1] = sv / count;
        out[iy * 5 + 2] = sw / count;
        out[iy * 5 + 3] = ap;
        out[iy * 5 + 4] = sqrt(sq / count - ap * ap);
    }
}

/* No reference counterpart (the reference has no energy or mass output). */
void orc_energy_mass(const orc_params *p, const orc_fields *s, double out[2]) {
    const size_t n = ncell(p);
    double ke = 0.0, mass = 0.0;
    for (size_t i = 0; i < n; i++) {
        ke += s->ux[i] * s->ux[i] + s->uy[i] * s->uy[i] + s->uz[i] * s->uz[i];
        mass += s->rho[i];
    }
    out[0] = 0.5 * ke / (double)n;
    out[1] = mass;
}
