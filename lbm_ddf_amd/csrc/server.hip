// server.hip -- continuous-batching request loop over an ensemble.
//
// Requests (one Rayleigh-Benard cavity each) arrive at given times.  Between
// step chunks the loop admits every arrived request that fits the HBM budget
// (first fit in arrival order), advances all resident cavities together by
// one chunk (one ensemble launch per step, the last with the diagnostics
// folded in), checks each cavity's stopping rule, and retires finished ones,
// freeing their memory for the queue.  A cavity's trajectory and its check
// points depend only on its own step count, so its result is the same alone
// and in any mix.
//
// Stopping rule: after every `window` steps of its own the cavity samples
// Nu_bulk (src/output.cu:228-233); the running Nu is the mean of its samples;
// the request is converged when the running Nu changes by less than nu_tol
// between two checks (at least three samples), or stops at its step budget.
// A cavity with a non-finite moment or a speed above the lattice sound speed
// is retired as diverged; the others continue.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

// Lattice sound speed of D3Q27 (c_s^2 = 1/3).
const double kSoundSpeed = std::sqrt(1.0 / 3.0);

struct Live {
    int32_t req = -1;
    int32_t slot = -1;
    uint64_t bytes = 0;
    uint64_t cells = 0;
    int64_t next_check = 0;
    int64_t samples = 0;
    double nu_sum = 0.0;
    double run_nu = NAN;
};

double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return NAN;
    std::sort(v.begin(), v.end());
    const double pos = p * (v.size() - 1);
    const size_t i = static_cast<size_t>(pos);
    const double frac = pos - i;
    return i + 1 < v.size() ? v[i] * (1 - frac) + v[i + 1] * frac : v[i];
}

}  // namespace

extern "C" {

int lbm_serve(const lbm_request *reqs, int32_t n, const lbm_serve_options *opt, lbm_result *results,
              lbm_serve_stats *stats) {
    if (!reqs || n < 0 || !opt || !results) return lbm::ctx_fail(LBM_E_INVALID, "NULL argument");
    if (opt->chunk_steps < 1) return lbm::ctx_fail(LBM_E_INVALID, "chunk_steps must be >= 1");
    for (int32_t i = 0; i < n; ++i)
        if (reqs[i].window < opt->chunk_steps || reqs[i].window % opt->chunk_steps != 0 || reqs[i].max_steps < 1)
            return lbm::ctx_fail(LBM_E_INVALID, "each window must be a positive multiple of chunk_steps");
    lbm_ensemble *ens = nullptr;
    if (int rc = lbm_ensemble_create(opt->precision, &ens)) return rc;
    if (opt->block_x > 0)
        if (int rc = lbm_ensemble_set_tuning(ens, opt->block_x, opt->blocks_per_cu, 8, opt->flags)) {
            lbm_ensemble_destroy(ens);
            return rc;
        }

    std::vector<int32_t> order(n);
    for (int32_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    for (int32_t i = 0; i < n; ++i) {
        results[i] = lbm_result{};
        results[i].status = LBM_REQ_PENDING;
        results[i].arrival_s = reqs[i].arrival_s;
        results[i].admit_s = results[i].finish_s = NAN;
    }

    std::vector<int32_t> waiting;  // arrived, not admitted (arrival order)
    size_t next_arrival = 0;
    std::vector<Live> live;
    std::vector<lbm_diag> diag;
    uint64_t bytes_used = 0, peak_bytes = 0, cells_used = 0;
    int32_t peak_resident = 0;
    double cell_updates = 0.0, busy = 0.0;
    int rc = LBM_OK;
    const Clock::time_point t0 = Clock::now();

    auto finish = [&](Live &l, int32_t status, const lbm_diag *d, int64_t steps) {
        lbm_result &r = results[l.req];
        r.status = status;
        r.finish_s = seconds_since(t0);
        r.steps = steps;
        if (d) {
            r.nu_bulk = d->nu_bulk;
            r.nu_top = d->nu_top;
            r.nu_bottom = d->nu_bottom;
            r.kinetic_energy = d->kinetic_energy;
        }
        r.nu_running = l.run_nu;
        (void)lbm_ensemble_remove(ens, l.slot);
        bytes_used -= l.bytes;
        cells_used -= l.cells;
        l.req = -1;
    };

    while (true) {
        double now = seconds_since(t0);
        while (next_arrival < order.size() && reqs[order[next_arrival]].arrival_s <= now)
            waiting.push_back(order[next_arrival++]);
        if (now >= opt->time_limit_s) break;

        // ---- admission: first fit in arrival order -------------------------
        for (size_t w = 0; w < waiting.size();) {
            const int32_t id = waiting[w];
            const lbm_request &q = reqs[id];
            lbm_params p{q.lx, q.ly, q.lz, opt->precision, q.rayleigh, q.prandtl, q.t_hot, q.t_cold, 1.0, 0.0};
            uint64_t need = 0;
            if (lbm_bytes_required(&p, &need) != LBM_OK || need > opt->hbm_budget) {
                results[id].status = LBM_REQ_REJECTED;
                results[id].finish_s = now;
                waiting.erase(waiting.begin() + w);
                continue;
            }
            const uint64_t cells = static_cast<uint64_t>(q.lx) * q.ly * q.lz;
            const bool room = bytes_used + need <= opt->hbm_budget &&
                              (opt->max_resident <= 0 || static_cast<int32_t>(live.size()) < opt->max_resident) &&
                              (opt->max_resident_cells == 0 || live.empty() ||
                               cells_used + cells <= opt->max_resident_cells);
            if (!room) {
                ++w;
                continue;
            }
            Live l;
            l.req = id;
            l.bytes = need;
            l.cells = cells;
            if ((rc = lbm_ensemble_add(ens, &p, q.seed, q.amp, &l.slot)) != LBM_OK) goto out;
            l.next_check = q.window;
            bytes_used += need;
            cells_used += cells;
            results[id].admit_s = seconds_since(t0);
            results[id].status = LBM_REQ_UNFINISHED;
            live.push_back(l);
            waiting.erase(waiting.begin() + w);
        }
        peak_bytes = std::max(peak_bytes, bytes_used);
        peak_resident = std::max(peak_resident, static_cast<int32_t>(live.size()));

        if (live.empty()) {
            if (next_arrival >= order.size() && waiting.empty()) break;
            if (next_arrival < order.size()) {
                const double wait = std::min(reqs[order[next_arrival]].arrival_s, opt->time_limit_s) - seconds_since(t0);
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            }
            continue;
        }

        // ---- one chunk over all resident cavities --------------------------
        {
            const double tb = seconds_since(t0);
            int32_t slots = 0;
            if ((rc = lbm_ensemble_count(ens, &slots, nullptr, nullptr)) != LBM_OK) goto out;
            diag.assign(std::max(1, slots), lbm_diag{});
            std::vector<int64_t> before(live.size());
            for (size_t i = 0; i < live.size(); ++i) {
                lbm_ctx *c = nullptr;
                lbm_ensemble_cavity(ens, live[i].slot, &c);
                before[i] = c->nstep;  // a joining cavity also gets its prologue step
            }
            if ((rc = lbm_ensemble_run_diag(ens, opt->chunk_steps, diag.data(), static_cast<int32_t>(diag.size()))) != LBM_OK)
                goto out;
            busy += seconds_since(t0) - tb;
            for (size_t i = 0; i < live.size(); ++i) {
                Live &l = live[i];
                lbm_ctx *c = nullptr;
                lbm_ensemble_cavity(ens, l.slot, &c);
                const lbm_diag &d = diag[l.slot];
                const int64_t done = d.step;  // the cavity's completed step at the chunk end
                cell_updates += static_cast<double>(l.cells) * static_cast<double>(c->nstep - before[i]);
                const lbm_request &q = reqs[l.req];
                // health: a non-finite moment, or a speed beyond the lattice
                // sound speed 1/sqrt(3), where the BGK model has left its range
                if (d.nonfinite != 0 || !(d.max_speed <= kSoundSpeed)) {
                    finish(l, LBM_REQ_DIVERGED, &d, done);
                    continue;
                }
                if (done >= l.next_check) {
                    const double prev = l.run_nu;
                    l.nu_sum += d.nu_bulk;
                    l.samples += 1;
                    l.run_nu = l.nu_sum / l.samples;
                    l.next_check += q.window;
                    if (l.samples >= 3 && std::fabs(l.run_nu - prev) < q.nu_tol) {
                        finish(l, LBM_REQ_CONVERGED, &d, done);
                        continue;
                    }
                }
                if (done >= q.max_steps) finish(l, LBM_REQ_BUDGET, &d, done);
            }
            live.erase(std::remove_if(live.begin(), live.end(), [](const Live &l) { return l.req < 0; }), live.end());
        }
    }

    // ---- end of run: whatever is resident or queued did not finish ---------
    for (Live &l : live) {
        lbm_ctx *c = nullptr;
        lbm_ensemble_cavity(ens, l.slot, &c);
        finish(l, LBM_REQ_UNFINISHED, nullptr, c ? c->nstep : 0);
    }
    live.clear();

out:
    {
        const double wall = seconds_since(t0);
        if (stats) {
            *stats = lbm_serve_stats{};
            std::vector<double> ttr;
            for (int32_t i = 0; i < n; ++i) {
                const lbm_result &r = results[i];
                switch (r.status) {
                    case LBM_REQ_CONVERGED:
                    case LBM_REQ_BUDGET:
                        stats->completed++;
                        ttr.push_back(r.finish_s - r.arrival_s);
                        break;
                    case LBM_REQ_REJECTED: stats->rejected++; break;
                    case LBM_REQ_DIVERGED: stats->diverged++; break;
                    case LBM_REQ_UNFINISHED: stats->unfinished++; break;
                    default: stats->pending++; break;
                }
            }
            stats->wall_s = wall;
            stats->busy_s = busy;
            stats->cell_updates = cell_updates;
            stats->mlups = wall > 0 ? cell_updates / wall / 1e6 : 0.0;
            stats->p50_s = percentile(ttr, 0.50);
            stats->p99_s = percentile(ttr, 0.99);
            stats->peak_resident = peak_resident;
            stats->peak_bytes = peak_bytes;
        }
    }
    lbm_ensemble_destroy(ens);
    return rc;
}

}  // extern "C"
