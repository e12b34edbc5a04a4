/*
 * selftest.c -- a small driver of the oracle for sanitizer builds (test
 * infrastructure only; see `make -C oracle sanitize`).
 *
 * Runs every oracle entry point on small odd shapes (including LX = 1 and
 * LZ = 1, where the wall blocks overlap most) with the stage sequence of
 * src/main.cu:27-45 and the diagnostics of src/output.cu, under
 * AddressSanitizer and UndefinedBehaviorSanitizer.  Exits 0 when every run
 * completes with finite diagnostics and conserved mass.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "oracle.h"

static int run_case(int lx, int ly, int lz, double ra) {
    orc_params p = {0};
    p.lx = lx;
    p.ly = ly;
    p.lz = lz;
    p.rayl = ra;
    p.prand = 0.71;
    p.t_hot = 0.5;
    p.t_cold = -0.5;
    p.grav0 = 1.0;
    p.beta = 0.0;
    orc_derive(&p);
    const size_t n = (size_t)lx * ly * lz;
    orc_fields s;
    double **pops[] = {&s.f, &s.g, &s.f_tmp, &s.g_tmp};
    double **flds[] = {&s.rho, &s.ux, &s.uy, &s.uz, &s.phi, &s.fx, &s.fy, &s.fz};
    for (int k = 0; k < 4; ++k) *pops[k] = calloc(ORC_NPOP * n, sizeof(double));
    for (int k = 0; k < 8; ++k) *flds[k] = calloc(n, sizeof(double));
    double *prof = calloc(5 * (size_t)ly, sizeof(double));
    orc_init_conductive(&p, &s, 7, 1e-2);
    double em0[2], em[2], nu[3], dg[4];
    orc_energy_mass(&p, &s, em0);
    for (int step = 0; step < 5; ++step) {
        orc_forcing(&p, &s);
        orc_collision_bgk(&p, &s);
        orc_collision_bgk_scalar(&p, &s);
        orc_streaming(&p, &s);
        orc_streaming_scalar(&p, &s);
        orc_macrovar(&p, &s);
    }
    orc_run(&p, &s, 5);
    orc_output_nu(&p, &s, nu);
    orc_diag_flow(&p, &s, dg);
    orc_output_profile(&p, &s, prof);
    orc_energy_mass(&p, &s, em);
    int ok = isfinite(nu[0]) && isfinite(nu[1]) && isfinite(nu[2]) && isfinite(dg[3]) && isfinite(prof[4]) &&
             fabs(em[1] - em0[1]) < 1e-10;
    orc_init_arrays(&p, &s);
    printf("%dx%dx%d Ra=%g: Nu %.6f %.6f %.6f mass %.3e -> %s\n", lx, ly, lz, ra, nu[0], nu[1], nu[2], em[1],
           ok ? "ok" : "FAIL");
    for (int k = 0; k < 4; ++k) free(*pops[k]);
    for (int k = 0; k < 8; ++k) free(*flds[k]);
    free(prof);
    return ok ? 0 : 1;
}

int main(void) {
    int bad = 0;
    bad += run_case(5, 7, 9, 1e6);
    bad += run_case(1, 4, 3, 1e5);
    bad += run_case(3, 3, 1, 1e5);
    bad += run_case(6, 5, 4, 1e7);
    return bad ? 1 : 0;
}
