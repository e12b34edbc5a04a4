// server.hip -- lbm_serve: the continuous-batching request loop of
// serve_loop.hpp over an ensemble (one persistent launch per step for all
// resident cavities, diagnostics folded into each chunk's last launch; A/B or
// in-place members) and the steady clock.
#include <chrono>
#include <thread>
#include <vector>

#include "div_magic.hpp"
#include "internal.hpp"
#include "serve_loop.hpp"

namespace {

// Cavities of an ensemble, with the request's walls and the loop's precision.
struct EnsembleBackend {
    lbm_ensemble *ens;
    int32_t precision;
    int32_t pattern;

    lbm_params params(const lbm_request &q) const {
        lbm_params p{};
        p.lx = q.lx;
        p.ly = q.ly;
        p.lz = q.lz;
        p.precision = precision;
        p.rayleigh = q.rayleigh;
        p.prandtl = q.prandtl;
        p.t_hot = q.t_hot;
        p.t_cold = q.t_cold;
        p.gravity = 1.0;
        p.beta = 0.0;
        p.pattern = pattern;
        return p;
    }
    int need(const lbm_request &q, uint64_t *bytes) const {
        const lbm_params p = params(q);
        // an ensemble member has fewer than 2^31 cells: larger requests are
        // rejected at admission rather than failing the run at lbm_ensemble_add
        if (q.lx > 0 && q.ly > 0 && q.lz > 0 && static_cast<uint64_t>(q.lx) * q.ly * q.lz >= lbm::kDivLimit)
            return LBM_E_INVALID;
        return lbm_bytes_required(&p, bytes);
    }
    int add(const lbm_request &q, int32_t *slot) {
        const lbm_params p = params(q);
        return lbm_ensemble_add(ens, &p, q.seed, q.amp, slot);
    }
    void remove(int32_t slot) { (void)lbm_ensemble_remove(ens, slot); }
    int steps(int32_t slot, int64_t *n) {
        lbm_ctx *c = nullptr;
        if (int rc = lbm_ensemble_cavity(ens, slot, &c)) return rc;
        return lbm_step_count(c, n);
    }
    int chunk(int64_t nsteps, std::vector<lbm_diag> &diag) {
        int32_t slots = 0;
        if (int rc = lbm_ensemble_count(ens, &slots, nullptr, nullptr)) return rc;
        diag.assign(static_cast<size_t>(std::max(1, slots)), lbm_diag{});
        return lbm_ensemble_run_diag(ens, nsteps, diag.data(), static_cast<int32_t>(diag.size()));
    }
};

struct SteadyClock {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    void sleep_until(double t) const {
        const double wait = t - now();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
    }
};

}  // namespace

extern "C" {

int lbm_serve(const lbm_request *reqs, int32_t n, const lbm_serve_options *opt, lbm_result *results,
              lbm_serve_stats *stats) {
    if (const char *err = lbm::serve_args_error(reqs, n, opt, results)) return lbm::ctx_fail(LBM_E_INVALID, err);
    lbm_ensemble *ens = nullptr;
    if (int rc = lbm_ensemble_create(opt->precision, &ens)) return rc;
    if (opt->block_x > 0)
        if (int rc = lbm_ensemble_set_tuning(ens, opt->block_x, opt->blocks_per_cu, 8, opt->flags,
                                             opt->cells_per_thread > 0 ? opt->cells_per_thread : 1)) {
            lbm_ensemble_destroy(ens);
            return rc;
        }
    EnsembleBackend be{ens, opt->precision, opt->pattern};
    SteadyClock clk;
    const int rc = lbm::serve_loop(be, clk, reqs, n, opt, results, stats);
    lbm_ensemble_destroy(ens);
    return rc;
}

}  // extern "C"
