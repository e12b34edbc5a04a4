// rb3d_main.cpp -- host driver with the reference's run structure
// (src/main.cu:3-122) on top of the C ABI.
//
// The reference fixes everything at compile time (src/parameters.h); here the
// same knobs are command-line options with the reference's defaults.  Output
// files keep the reference's names and formats in the directory
// Ra%.1ePr%.2f (src/output.cu:5):
//   statistics.dat   every NDIAG steps   (diag_flow,      src/output.cu:3-65)
//   %09d.dat         every NFLOWOUT/NEND (output_flow,    src/output.cu:67-146)
//   Nu_<Ra>.txt      every NNUOUT steps  (output_nu,      src/output.cu:148-280)
//   Profile_<Ra>.txt every NNUOUT steps  (output_profile, src/output.cu:282-377)
//   fg%09d.dat       at NEND             (output_fg,      src/output.cu:379-419)
// and --continue N restarts from %09d.dat + fg%09d.dat (src/initial.cu:110-205).
//
// Differences: the time loop calls the fused lbm_step between output events
// instead of five kernels per step; errors print and return a non-zero exit
// status; --init conductive selects the seeded conductive start; --pattern aa
// keeps one population buffer per distribution (in place).
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_ddf.h"

namespace {

struct Options {
    int lx = 400, ly = 400, lz = 400;
    long nend = 1000, ndiag = 10000, nflowout = 10000, nnuout = 10000, cont = 0;
    double ra = 1.0e7, pr = 0.71, t_hot = 1.0, t_cold = 0.0, gravity = 1.0, beta = 0.0;
    std::string precision = "fp64", init = "reference", outdir = ".", pattern = "ab";
    unsigned long long seed = 1;
    double amp = 1.0e-3;
};

#define CHECK(call)                                                                           \
    do {                                                                                      \
        int rc_ = (call);                                                                     \
        if (rc_ != 0) {                                                                       \
            std::fprintf(stderr, "lbm error %d in %s: %s\n", rc_, #call, lbm_last_error());   \
            return 1;                                                                         \
        }                                                                                     \
    } while (0)

void usage() {
    std::fprintf(stderr,
                 "rb3d [--lx N --ly N --lz N] [--nend N] [--ndiag N] [--nflowout N] [--nnuout N]\n"
                 "     [--ra X] [--pr X] [--thot X] [--tcold X] [--gravity X] [--beta X]\n"
                 "     [--precision fp64|fp16|bf16] [--pattern ab|aa] [--init reference|conductive]\n"
                 "     [--seed N] [--amp X]\n"
                 "     [--continue STEP] [--outdir DIR]\n");
}

bool parse(int argc, char **argv, Options &o) {
    for (int i = 1; i < argc; ++i) {
        std::string k = argv[i];
        if (k == "-h" || k == "--help") return false;
        if (i + 1 >= argc) return false;
        const char *v = argv[++i];
        if (k == "--lx") o.lx = std::atoi(v);
        else if (k == "--ly") o.ly = std::atoi(v);
        else if (k == "--lz") o.lz = std::atoi(v);
        else if (k == "--nend") o.nend = std::atol(v);
        else if (k == "--ndiag") o.ndiag = std::atol(v);
        else if (k == "--nflowout") o.nflowout = std::atol(v);
        else if (k == "--nnuout") o.nnuout = std::atol(v);
        else if (k == "--continue") o.cont = std::atol(v);
        else if (k == "--ra") o.ra = std::atof(v);
        else if (k == "--pr") o.pr = std::atof(v);
        else if (k == "--thot") o.t_hot = std::atof(v);
        else if (k == "--tcold") o.t_cold = std::atof(v);
        else if (k == "--gravity") o.gravity = std::atof(v);
        else if (k == "--beta") o.beta = std::atof(v);
        else if (k == "--precision") o.precision = v;
        else if (k == "--init") o.init = v;
        else if (k == "--pattern") o.pattern = v;
        else if (k == "--seed") o.seed = std::strtoull(v, nullptr, 10);
        else if (k == "--amp") o.amp = std::atof(v);
        else if (k == "--outdir") o.outdir = v;
        else return false;
    }
    return o.ndiag > 0 && o.nflowout > 0 && o.nnuout > 0 && o.nend >= o.cont &&
           (o.pattern == "ab" || o.pattern == "aa");
}

// File suffix of output_nu / output_profile (src/output.cu:164-178).
std::string ra_tag(double ra) {
    const double known[] = {1e5, 5e5, 1e6, 5e6, 1e7, 1e8};
    const char *names[] = {"1e50", "5e50", "1e60", "5e60", "1e70", "1e80"};
    for (int i = 0; i < 6; ++i)
        if (ra == known[i]) return names[i];
    char b[64];
    std::snprintf(b, sizeof b, "%.2e", ra);
    return b;
}

}  // namespace

int main(int argc, char **argv) {
    Options o;
    if (!parse(argc, argv, o)) {
        usage();
        return 2;
    }
    int prec = o.precision == "fp16" ? LBM_PREC_FP16 : (o.precision == "bf16" ? LBM_PREC_BF16 : LBM_PREC_FP64);
    lbm_params p{o.lx, o.ly, o.lz, prec, o.ra, o.pr, o.t_hot, o.t_cold, o.gravity, o.beta,
                 o.pattern == "aa" ? LBM_PATTERN_AA : LBM_PATTERN_AB, 0};
    char dirname[512];
    std::snprintf(dirname, sizeof dirname, "%s/Ra%.1ePr%.2f", o.outdir.c_str(), o.ra, o.pr);
    mkdir(dirname, 0700);

    const auto t_start = std::chrono::steady_clock::now();
    lbm_ctx *ctx = nullptr;
    CHECK(lbm_create(&p, &ctx));
    if (o.cont > 0) {
        char flow[600], fg[600];
        std::snprintf(flow, sizeof flow, "%s/%09ld.dat", dirname, o.cont);
        std::snprintf(fg, sizeof fg, "%s/fg%09ld.dat", dirname, o.cont);
        CHECK(lbm_restart(ctx, flow, fg, o.cont));
    } else if (o.init == "conductive") {
        CHECK(lbm_init_conductive(ctx, o.seed, o.amp));
    } else {
        CHECK(lbm_init_arrays(ctx));
    }

    const size_t n = static_cast<size_t>(o.lx) * o.ly * o.lz;
    std::vector<double> ux, uy, uz, phi;
    double gpu_s = 0.0, host_s = 0.0;
    long istep = o.cont;
    while (istep < o.nend) {
        // next step with an output event (src/main.cu:64-76)
        long next = o.nend;
        for (long cadence : {o.ndiag, o.nflowout, o.nnuout}) next = std::min(next, (istep / cadence + 1) * cadence);
        auto t0 = std::chrono::steady_clock::now();
        CHECK(lbm_step(ctx, next - istep));
        CHECK(lbm_synchronize(ctx));
        gpu_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        istep = next;
        t0 = std::chrono::steady_clock::now();
        if (istep % o.ndiag == 0) {
            lbm_diag d;
            CHECK(lbm_reduce(ctx, &d));
            std::printf("Step %-7ld umean = %14.6e, vmean = %14.6e, wmean = %14.6e, tmean = %14.6e\n", istep, d.umean,
                        d.vmean, d.wmean, d.tmean);
            std::fflush(stdout);
            char fn[600];
            std::snprintf(fn, sizeof fn, "%s/statistics.dat", dirname);
            FILE *f = std::fopen(fn, "a");
            if (!f) return 1;
            std::fprintf(f, "%-7ld %14.6e %14.6e %14.6e %14.6e\n", istep, d.umean, d.vmean, d.wmean, d.tmean);
            std::fclose(f);
        }
        if (istep % o.nflowout == 0 || istep == o.nend) {
            char fn[600];
            std::snprintf(fn, sizeof fn, "%s/%09ld.dat", dirname, istep);
            CHECK(lbm_output_flow(ctx, fn));
        }
        if (istep % o.nnuout == 0) {
            lbm_diag d;
            CHECK(lbm_reduce(ctx, &d));
            char fn[600];
            std::snprintf(fn, sizeof fn, "%s/Nu_%s.txt", dirname, ra_tag(o.ra).c_str());
            FILE *f = std::fopen(fn, "a");
            if (!f) return 1;
            std::fprintf(f, "%15.6E%15.6E%15.6E%15.6E%15.6E\n", static_cast<double>(istep), 0.0, d.nu_bulk, d.nu_top,
                         d.nu_bottom);
            std::fclose(f);
            std::vector<double> prof(5 * static_cast<size_t>(o.ly));
            CHECK(lbm_profile(ctx, prof.data()));
            std::snprintf(fn, sizeof fn, "%s/Profile_%s.txt", dirname, ra_tag(o.ra).c_str());
            f = std::fopen(fn, "a");
            if (!f) return 1;
            std::fprintf(f, "%ld\n", istep);
            for (int iy = 0; iy < o.ly; ++iy)
                std::fprintf(f, "%14.6e %14.6e %14.6e %14.6e %14.6e\n", prof[5 * iy], prof[5 * iy + 1], prof[5 * iy + 2],
                             prof[5 * iy + 3], prof[5 * iy + 4]);
            std::fclose(f);
        }
        if (istep == o.nend) {
            char fn[600];
            std::snprintf(fn, sizeof fn, "%s/fg%09ld.dat", dirname, istep);
            CHECK(lbm_output_fg(ctx, fn));
        }
        host_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    (void)n;
    std::printf("Simulation completed.\n");
    std::printf("Rayleigh number: %.2e\n", o.ra);
    std::printf("Prandtl number: %.2f\n", o.pr);
    std::printf("Total iterations: %ld\n", o.nend);
    std::printf("CPU computation time: %.2f seconds\n", host_s);
    std::printf("GPU computation time: %.2f seconds\n", gpu_s);
    std::printf("Total computation time: %.2f seconds\n", total);
    const double cells = static_cast<double>(o.lx) * o.ly * o.lz;
    if (gpu_s > 0) std::printf("MLUPS: %.1f\n", cells * (o.nend - o.cont) / gpu_s / 1e6);
    CHECK(lbm_destroy(ctx));
    return 0;
}
