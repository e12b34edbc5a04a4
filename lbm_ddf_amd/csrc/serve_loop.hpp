// serve_loop.hpp -- the scheduling policy of lbm_serve, independent of the
// GPU: admission, chunked advance, stopping rules, retirement and statistics.
//
// Requests (one Rayleigh-Benard cavity each) arrive at given times.  Between
// step chunks the loop admits every arrived request that fits the HBM budget
// (first fit in arrival order: a request that does not fit does not block
// later ones), advances all resident cavities together by one chunk, checks
// each cavity's stopping rule, and retires finished ones, freeing their
// memory for the queue.  A cavity's trajectory and its check points depend
// only on its own step count, so its result is the same alone and in any mix.
//
// Stopping rule: after every `window` steps of its own the cavity samples
// Nu_bulk (src/output.cu:228-233); the running Nu is the mean of its samples;
// the request is converged when the running Nu changes by less than nu_tol
// between two checks (at least three samples), or stops at its step budget.
// A cavity with a non-finite moment or a speed above the lattice sound speed
// is retired as diverged; the others continue.
//
// The template is instantiated by lbm_serve (server.hip) with the ensemble
// backend and a steady clock, and by the CPU tests with a synthetic backend
// and a virtual clock.  Backend interface:
//   int need(const lbm_request &q, uint64_t *bytes)   device bytes of a cavity
//   int add(const lbm_request &q, int32_t *slot)       allocate, conductive start;
//       LBM_E_NOMEM when the device has less free memory than the budget
//       assumed: the request goes back to the queue while other cavities are
//       resident (their retirement frees memory), and is rejected when nothing
//       is resident (it cannot run at all); the loop keeps serving
//   void remove(int32_t slot)
//   int steps(int32_t slot, int64_t *n)                completed steps so far
//   int chunk(int64_t nsteps, std::vector<lbm_diag> &diag)
//       advance every resident cavity by nsteps launches; diag[slot] holds the
//       diagnostics at the end (diag[slot].step = the completed step)
// Clock interface: double now() (seconds since the start), void sleep_until(double t).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "lbm_ddf.h"

namespace lbm {

// Lattice sound speed of D3Q27 (c_s^2 = 1/3).
inline double sound_speed() { return std::sqrt(1.0 / 3.0); }

inline double percentile(std::vector<double> v, double p) {
    if (v.empty()) return NAN;
    std::sort(v.begin(), v.end());
    const double pos = p * static_cast<double>(v.size() - 1);
    const size_t i = static_cast<size_t>(pos);
    const double frac = pos - static_cast<double>(i);
    return i + 1 < v.size() ? v[i] * (1 - frac) + v[i + 1] * frac : v[i];
}

// Argument checks of lbm_serve; 0 or a message.
inline const char *serve_args_error(const lbm_request *reqs, int32_t n, const lbm_serve_options *opt,
                                    const lbm_result *results) {
    if (!reqs || n < 0 || !opt || !results) return "NULL argument";
    if (opt->chunk_steps < 1) return "chunk_steps must be >= 1";
    if (opt->pattern != LBM_PATTERN_AB && opt->pattern != LBM_PATTERN_AA) return "unknown pattern";
    if (opt->reserved != 0) return "reserved must be 0";
    for (int32_t i = 0; i < n; ++i)
        if (reqs[i].window < opt->chunk_steps || reqs[i].window % opt->chunk_steps != 0 || reqs[i].max_steps < 1)
            return "each window must be a positive multiple of chunk_steps";
    return nullptr;
}

template <class Backend, class Clock>
int serve_loop(Backend &be, Clock &clk, const lbm_request *reqs, int32_t n, const lbm_serve_options *opt,
               lbm_result *results, lbm_serve_stats *stats) {
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

    std::vector<int32_t> order(n);
    for (int32_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    for (int32_t i = 0; i < n; ++i) {
        results[i] = lbm_result{};
        results[i].status = LBM_REQ_PENDING;
        results[i].arrival_s = reqs[i].arrival_s;
        results[i].admit_s = results[i].finish_s = NAN;
        results[i].nu_running = NAN;
    }

    std::vector<int32_t> waiting;  // arrived, not admitted (arrival order)
    size_t next_arrival = 0;
    std::vector<Live> live;
    std::vector<lbm_diag> diag;
    std::vector<int64_t> before;
    uint64_t bytes_used = 0, peak_bytes = 0, cells_used = 0;
    int32_t peak_resident = 0, retired = 0, deferred = 0;
    double cell_updates = 0.0, busy = 0.0, next_report = opt->progress_s;
    int rc = LBM_OK;

    auto finish = [&](Live &l, int32_t status, const lbm_diag *d, int64_t steps) {
        lbm_result &r = results[l.req];
        r.status = status;
        r.finish_s = clk.now();
        r.steps = steps;
        if (d) {
            r.nu_bulk = d->nu_bulk;
            r.nu_top = d->nu_top;
            r.nu_bottom = d->nu_bottom;
            r.kinetic_energy = d->kinetic_energy;
        }
        r.nu_running = l.run_nu;
        ++retired;
        be.remove(l.slot);
        bytes_used -= l.bytes;
        cells_used -= l.cells;
        l.req = -1;
    };

    while (rc == LBM_OK) {
        const double now = clk.now();
        while (next_arrival < order.size() && reqs[order[next_arrival]].arrival_s <= now)
            waiting.push_back(order[next_arrival++]);
        if (now >= opt->time_limit_s) break;

        // ---- admission: first fit in arrival order -------------------------
        for (size_t w = 0; w < waiting.size();) {
            const int32_t id = waiting[w];
            const lbm_request &q = reqs[id];
            uint64_t need = 0;
            if (be.need(q, &need) != LBM_OK || need > opt->hbm_budget) {
                results[id].status = LBM_REQ_REJECTED;
                results[id].finish_s = now;
                waiting.erase(waiting.begin() + static_cast<std::ptrdiff_t>(w));
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
            const int arc = be.add(q, &l.slot);
            if (arc == LBM_E_NOMEM) {
                if (live.empty()) {
                    results[id].status = LBM_REQ_REJECTED;
                    results[id].finish_s = now;
                    waiting.erase(waiting.begin() + static_cast<std::ptrdiff_t>(w));
                    continue;
                }
                ++deferred;
                break;  // admit again after the next chunk's retirements
            }
            if (arc != LBM_OK) {
                rc = arc;
                break;
            }
            l.next_check = q.window;
            bytes_used += need;
            cells_used += cells;
            results[id].admit_s = clk.now();
            results[id].status = LBM_REQ_UNFINISHED;
            live.push_back(l);
            waiting.erase(waiting.begin() + static_cast<std::ptrdiff_t>(w));
        }
        if (rc != LBM_OK) break;
        peak_bytes = std::max(peak_bytes, bytes_used);
        peak_resident = std::max(peak_resident, static_cast<int32_t>(live.size()));

        if (live.empty()) {
            if (next_arrival >= order.size() && waiting.empty()) break;
            if (next_arrival < order.size())
                clk.sleep_until(std::min(reqs[order[next_arrival]].arrival_s, opt->time_limit_s));
            else
                break;  // queued requests that fit nothing while nothing is resident: cannot progress
            continue;
        }

        // ---- one chunk over all resident cavities --------------------------
        const double tb = clk.now();
        before.resize(live.size());
        for (size_t i = 0; i < live.size(); ++i)
            if ((rc = be.steps(live[i].slot, &before[i])) != LBM_OK) break;
        if (rc != LBM_OK) break;
        if ((rc = be.chunk(opt->chunk_steps, diag)) != LBM_OK) break;
        busy += clk.now() - tb;
        for (size_t i = 0; i < live.size(); ++i) {
            Live &l = live[i];
            const lbm_diag &d = diag[static_cast<size_t>(l.slot)];
            const int64_t done = d.step;  // the cavity's completed step at the chunk end
            // a joining cavity also gets its prologue step
            cell_updates += static_cast<double>(l.cells) * static_cast<double>(done - before[i]);
            const lbm_request &q = reqs[l.req];
            // health: a non-finite moment, or a speed beyond the lattice
            // sound speed, where the BGK model has left its range
            if (d.nonfinite != 0 || !(d.max_speed <= sound_speed())) {
                finish(l, LBM_REQ_DIVERGED, &d, done);
                continue;
            }
            if (done >= l.next_check) {
                const double prev = l.run_nu;
                l.nu_sum += d.nu_bulk;
                l.samples += 1;
                l.run_nu = l.nu_sum / static_cast<double>(l.samples);
                l.next_check += q.window;
                if (l.samples >= 3 && std::fabs(l.run_nu - prev) < q.nu_tol) {
                    finish(l, LBM_REQ_CONVERGED, &d, done);
                    continue;
                }
            }
            if (done >= q.max_steps) finish(l, LBM_REQ_BUDGET, &d, done);
        }
        live.erase(std::remove_if(live.begin(), live.end(), [](const Live &l) { return l.req < 0; }), live.end());
        if (opt->progress_s > 0 && clk.now() >= next_report) {
            const double t = clk.now();
            std::fprintf(stderr, "[serve] %.0f s: %zu resident, %zu queued, %zu not yet arrived, %d retired, %.0f MLUPS\n",
                         t, live.size(), waiting.size(), order.size() - next_arrival, retired,
                         t > 0 ? cell_updates / t / 1e6 : 0.0);
            std::fflush(stderr);
            next_report = t + opt->progress_s;
        }
    }

    // ---- end of run: whatever is resident did not finish -------------------
    for (Live &l : live) {
        int64_t s = 0;
        (void)be.steps(l.slot, &s);
        finish(l, LBM_REQ_UNFINISHED, nullptr, s);
    }
    live.clear();

    if (stats) {
        *stats = lbm_serve_stats{};
        const double wall = clk.now();
        std::vector<double> ttr, ttr_all;
        double queue = 0.0, service = 0.0;
        for (int32_t i = 0; i < n; ++i) {
            const lbm_result &r = results[i];
            switch (r.status) {
                case LBM_REQ_CONVERGED:
                case LBM_REQ_BUDGET:
                    stats->completed++;
                    ttr.push_back(r.finish_s - r.arrival_s);
                    queue += r.admit_s - r.arrival_s;
                    service += r.finish_s - r.admit_s;
                    break;
                case LBM_REQ_REJECTED: stats->rejected++; break;
                case LBM_REQ_DIVERGED: stats->diverged++; break;
                case LBM_REQ_UNFINISHED: stats->unfinished++; break;
                default: stats->pending++; break;
            }
            if (r.arrival_s > wall) continue;  // never arrived
            stats->arrived++;
            const bool open = r.status == LBM_REQ_UNFINISHED || r.status == LBM_REQ_PENDING;
            ttr_all.push_back((open ? wall : r.finish_s) - r.arrival_s);
        }
        stats->wall_s = wall;
        stats->busy_s = busy;
        stats->cell_updates = cell_updates;
        stats->mlups = wall > 0 ? cell_updates / wall / 1e6 : 0.0;
        stats->p50_s = percentile(ttr, 0.50);
        stats->p99_s = percentile(ttr, 0.99);
        stats->p50_all_s = percentile(ttr_all, 0.50);
        stats->p99_all_s = percentile(ttr_all, 0.99);
        stats->mean_queue_s = stats->completed ? queue / stats->completed : NAN;
        stats->mean_service_s = stats->completed ? service / stats->completed : NAN;
        stats->deferred = deferred;
        stats->peak_resident = peak_resident;
        stats->peak_bytes = peak_bytes;
    }
    return rc;
}

}  // namespace lbm
