// serve_loop_cpu.cpp -- the scheduling policy of lbm_serve (serve_loop.hpp)
// on the CPU, with a synthetic backend and a virtual clock (test double; the
// product instantiation is server.hip).
//
// Synthetic cavity: its Nu_bulk after s completed steps is a function of
// (seed, s) only; a chunk of k launches advances each resident cavity by k
// steps (k + 1 for a cavity's first chunk, which includes the prologue, as
// the ensemble does) and costs cells * k / rate seconds of virtual time.
// Seeds >= 1000 go non-finite at step seed - 1000; seeds >= 2000 exceed the
// sound speed at step seed - 2000; seeds >= 3000 have a constant Nu.
//
// Built and run by tests/test_serve_loop_cpu.py; exits non-zero on the first
// failed check, printing it.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "serve_loop.hpp"

namespace {

struct VirtualClock {
    double t = 0.0;
    double now() const { return t; }
    void sleep_until(double x) {
        if (x > t) t = x;
    }
};

struct FakeBackend {
    VirtualClock *clk = nullptr;
    double rate = 1.0e9;  // cell updates per virtual second
    uint64_t bytes_per_cell = 100;
    struct Cav {
        uint64_t seed, cells;
        int64_t steps = 0;
    };
    std::map<int32_t, Cav> cavs;
    int add_calls = 0;
    // device memory actually free for cavities (the budget may claim more):
    // add() fails with LBM_E_NOMEM beyond it
    uint64_t device_bytes = ~0ull;
    uint64_t allocated = 0;

    static double nu_of(uint64_t seed, int64_t s) {
        if (seed >= 3000) return 4.0;
        // decays towards 2 + seed/10 with an oscillation: converges for any tol
        return 2.0 + 0.1 * static_cast<double>(seed % 100) + std::exp(-static_cast<double>(s) / 400.0) *
                                                                  std::cos(static_cast<double>(s) / 50.0);
    }
    int need(const lbm_request &q, uint64_t *bytes) const {
        *bytes = static_cast<uint64_t>(q.lx) * q.ly * q.lz * bytes_per_cell;
        return 0;
    }
    int add(const lbm_request &q, int32_t *slot) {
        ++add_calls;
        uint64_t b = 0;
        need(q, &b);
        if (allocated + b > device_bytes) return LBM_E_NOMEM;
        int32_t s = 0;
        while (cavs.count(s)) ++s;
        cavs[s] = Cav{q.seed, static_cast<uint64_t>(q.lx) * q.ly * q.lz};
        allocated += b;
        *slot = s;
        return 0;
    }
    void remove(int32_t slot) {
        allocated -= cavs.at(slot).cells * bytes_per_cell;
        cavs.erase(slot);
    }
    int steps(int32_t slot, int64_t *n) {
        *n = cavs.at(slot).steps;
        return 0;
    }
    int chunk(int64_t k, std::vector<lbm_diag> &diag) {
        int32_t top = 0;
        for (auto &kv : cavs) top = std::max(top, kv.first + 1);
        diag.assign(static_cast<size_t>(std::max(1, top)), lbm_diag{});
        double work = 0.0;
        for (auto &kv : cavs) {
            Cav &c = kv.second;
            const int64_t adv = c.steps == 0 ? k + 1 : k;
            c.steps += adv;
            work += static_cast<double>(c.cells) * static_cast<double>(adv);
            lbm_diag &d = diag[static_cast<size_t>(kv.first)];
            d.step = c.steps;
            d.nu_bulk = nu_of(c.seed, c.steps);
            d.nu_top = d.nu_bottom = d.nu_bulk;
            d.max_speed = 0.05;
            if (c.seed >= 2000 && c.seed < 3000 && c.steps >= static_cast<int64_t>(c.seed - 2000)) d.max_speed = 0.7;
            else if (c.seed >= 1000 && c.seed < 2000 && c.steps >= static_cast<int64_t>(c.seed - 1000)) d.nonfinite = 1;
        }
        clk->t += work / rate;
        return 0;
    }
};

int failures = 0;
#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);        \
            ++failures;                                                          \
        }                                                                        \
    } while (0)

lbm_request req(int side, uint64_t seed, double arrival, int64_t max_steps = 4000, double tol = 1e-12,
                int64_t window = 200) {
    lbm_request q{};
    q.lx = q.ly = q.lz = side;
    q.rayleigh = 1e5;
    q.prandtl = 0.71;
    q.t_hot = 0.5;
    q.t_cold = -0.5;
    q.seed = seed;
    q.amp = 1e-3;
    q.arrival_s = arrival;
    q.max_steps = max_steps;
    q.nu_tol = tol;
    q.window = window;
    return q;
}

lbm_serve_options options() {
    lbm_serve_options o{};
    o.precision = LBM_PREC_FP16;
    o.chunk_steps = 100;
    o.hbm_budget = 1ull << 40;
    o.time_limit_s = 1e9;
    return o;
}

struct Run {
    std::vector<lbm_result> res;
    lbm_serve_stats st{};
    int rc = 0;
};

Run serve(const std::vector<lbm_request> &reqs, const lbm_serve_options &o, uint64_t device_bytes = ~0ull) {
    VirtualClock clk;
    FakeBackend be;
    be.clk = &clk;
    be.device_bytes = device_bytes;
    Run r;
    r.res.resize(reqs.size());
    r.rc = lbm::serve_loop(be, clk, reqs.data(), static_cast<int32_t>(reqs.size()), &o, r.res.data(), &r.st);
    if (!be.cavs.empty()) {
        std::printf("FAILED: %zu cavities left resident\n", be.cavs.size());
        ++failures;
    }
    return r;
}

bool same_result(const lbm_result &a, const lbm_result &b) {
    return a.status == b.status && a.steps == b.steps && a.nu_bulk == b.nu_bulk &&
           ((std::isnan(a.nu_running) && std::isnan(b.nu_running)) || a.nu_running == b.nu_running);
}

void test_budget_status_and_alone_vs_mix() {
    const auto o = options();
    const lbm_request a = req(20, 7, 0.0, 1000);
    Run alone = serve({a}, o);
    CHECK(alone.rc == 0);
    CHECK(alone.res[0].status == LBM_REQ_BUDGET);
    CHECK(alone.res[0].steps == 1001);  // chunks of 100 + the prologue step: first check point >= 1000
    Run mix = serve({req(30, 1, 0.0, 3000), req(10, 2, 0.0, 500), a, req(16, 3, 0.0, 2000)}, o);
    CHECK(same_result(alone.res[0], mix.res[2]));
    CHECK(mix.st.completed == 4 && mix.st.peak_resident == 4);
}

void test_convergence_rule() {
    auto o = options();
    // tolerance that the decaying oscillation meets after a few windows
    Run r = serve({req(12, 5, 0.0, 1000000, 1e-3, 200)}, o);
    CHECK(r.res[0].status == LBM_REQ_CONVERGED);
    CHECK(r.res[0].steps % 100 == 1);  // check points fall on chunk ends (prologue + k*100)
    CHECK(r.res[0].steps >= 3 * 200);
    // the running Nu is the mean of the samples taken every window
    double sum = 0.0;
    int64_t nsamp = 0;
    for (int64_t s = 201; s <= r.res[0].steps; s += 200) {  // first chunk end at or past each window
        sum += FakeBackend::nu_of(5, s);
        ++nsamp;
    }
    CHECK(std::fabs(r.res[0].nu_running - sum / static_cast<double>(nsamp)) < 1e-12);
    // converges exactly when two consecutive running means differ by less than tol
    double prev = NAN, run = 0.0;
    sum = 0.0;
    nsamp = 0;
    int64_t expect = -1;
    for (int64_t s = 201; s < 1000000; s += 200) {
        sum += FakeBackend::nu_of(5, s);
        ++nsamp;
        run = sum / static_cast<double>(nsamp);
        if (nsamp >= 3 && std::fabs(run - prev) < 1e-3) {
            expect = s;
            break;
        }
        prev = run;
    }
    CHECK(r.res[0].steps == expect);
    // a constant Nu converges at the third sample, not before
    Run c = serve({req(12, 3000, 0.0, 1000000, 1e-3, 200)}, o);
    CHECK(c.res[0].status == LBM_REQ_CONVERGED && c.res[0].steps == 601 && c.res[0].nu_running == 4.0);
}

void test_hbm_budget_first_fit_and_rejection() {
    auto o = options();
    const uint64_t per = 16ull * 16 * 16 * 100;
    o.hbm_budget = 2 * per + per / 2;  // two 16^3 cavities at a time
    std::vector<lbm_request> reqs = {req(16, 1, 0.0, 500), req(16, 2, 0.0, 800), req(16, 3, 0.0, 500),
                                     req(40, 4, 0.0, 500),  // larger than the whole budget
                                     req(8, 5, 0.0, 500)};  // arrives last, fits beside two 16^3
    Run r = serve(reqs, o);
    CHECK(r.res[3].status == LBM_REQ_REJECTED);
    CHECK(r.st.rejected == 1 && r.st.completed == 4);
    CHECK(r.st.peak_bytes <= o.hbm_budget);
    // first fit: the small one is admitted at once although request 2 waits
    CHECK(r.res[4].admit_s == 0.0);
    CHECK(r.res[2].admit_s > 0.0);
    CHECK(r.res[2].admit_s >= r.res[0].finish_s);  // after the first retirement freed memory
}

void test_resident_caps() {
    auto o = options();
    o.max_resident = 2;
    Run r = serve({req(10, 1, 0.0, 300), req(10, 2, 0.0, 300), req(10, 3, 0.0, 300)}, o);
    CHECK(r.st.peak_resident == 2 && r.st.completed == 3);
    auto o2 = options();
    o2.max_resident_cells = 2 * 10 * 10 * 10;
    Run r2 = serve({req(10, 1, 0.0, 300), req(10, 2, 0.0, 300), req(10, 3, 0.0, 300), req(30, 4, 0.0, 300)}, o2);
    CHECK(r2.st.peak_resident == 2 && r2.st.completed == 4);  // a lone oversize cavity is still admitted
}

void test_divergence_retires_only_the_sick() {
    auto o = options();
    const lbm_request healthy = req(12, 9, 0.0, 900);
    Run alone = serve({healthy}, o);
    Run r = serve({req(12, 1000 + 250, 0.0, 5000), healthy, req(12, 2000 + 420, 0.0, 5000)}, o);
    CHECK(r.res[0].status == LBM_REQ_DIVERGED && r.res[0].steps == 301);  // first chunk end past step 250
    CHECK(r.res[2].status == LBM_REQ_DIVERGED && r.res[2].steps == 501);
    CHECK(same_result(alone.res[0], r.res[1]));
    CHECK(r.st.diverged == 2 && r.st.completed == 1);
}

void test_time_limit_and_arrivals() {
    auto o = options();
    o.time_limit_s = 0.5;
    // 24^3 cells * 100 steps / 1e9 per s = 1.4 ms per chunk: 10^9 steps never finish
    Run r = serve({req(24, 1, 0.0, 1000000000), req(24, 2, 30.0, 100)}, o);
    CHECK(r.res[0].status == LBM_REQ_UNFINISHED && r.res[0].steps > 0);
    CHECK(r.res[1].status == LBM_REQ_PENDING);
    CHECK(r.st.unfinished == 1 && r.st.pending == 1);
    CHECK(r.st.wall_s >= 0.5);
    // an idle loop waits for the next arrival
    auto o2 = options();
    Run r2 = serve({req(8, 1, 2.5, 200)}, o2);
    CHECK(r2.res[0].admit_s == 2.5 && r2.res[0].status == LBM_REQ_BUDGET);
}

void test_time_to_result_statistics() {
    auto o = options();
    std::vector<lbm_request> reqs;
    for (int i = 0; i < 20; ++i) reqs.push_back(req(8 + i, static_cast<uint64_t>(i), 0.01 * i, 300));
    Run r = serve(reqs, o);
    std::vector<double> ttr;
    for (const auto &x : r.res) {
        CHECK(x.finish_s >= x.admit_s && x.admit_s >= x.arrival_s);
        ttr.push_back(x.finish_s - x.arrival_s);
    }
    CHECK(r.st.completed == 20);
    CHECK(std::fabs(r.st.p50_s - lbm::percentile(ttr, 0.5)) < 1e-15);
    CHECK(std::fabs(r.st.p99_s - lbm::percentile(ttr, 0.99)) < 1e-15);
    CHECK(r.st.p99_s >= r.st.p50_s);
    double cells = 0.0;
    for (const auto &q : reqs) cells += static_cast<double>(q.lx) * q.ly * q.lz * 301.0;
    CHECK(std::fabs(r.st.cell_updates - cells) < 1e-6 * cells);
}

void test_allocation_failure_defers_or_rejects() {
    // The budget admits three 16^3 cavities, the device holds two: the third
    // admission fails with LBM_E_NOMEM and the request waits for a
    // retirement instead of ending the run.  A request that the device cannot
    // hold even alone is rejected, and the others are served.
    auto o = options();
    const uint64_t per = 16ull * 16 * 16 * 100;
    o.hbm_budget = 3 * per;
    Run r = serve({req(16, 1, 0.0, 500), req(16, 2, 0.0, 900), req(16, 3, 0.0, 500)}, o, 2 * per);
    CHECK(r.rc == 0);
    CHECK(r.st.completed == 3 && r.st.rejected == 0);
    CHECK(r.st.deferred >= 1);
    CHECK(r.res[2].admit_s >= r.res[0].finish_s);  // after the first retirement freed memory
    CHECK(r.st.peak_resident == 2);
    Run alone = serve({req(16, 3, 0.0, 500)}, o);
    CHECK(same_result(alone.res[0], r.res[2]));
    Run big = serve({req(20, 1, 0.0, 300), req(10, 2, 0.0, 300)}, o, per);
    CHECK(big.rc == 0);
    CHECK(big.res[0].status == LBM_REQ_REJECTED && big.res[1].status == LBM_REQ_BUDGET);
}

void test_censored_latency_and_queueing() {
    // Three requests arrive at once; one GPU slot (max_resident 1); the time
    // limit ends the run while the third is queued.  The completed-only
    // percentiles see the first request alone; the all-arrived ones count the
    // resident one and the queued one with the end of the run as their finish.
    auto o = options();
    o.max_resident = 1;
    // 20^3 * 100 / 1e9 = 0.8 ms per chunk; 1001 steps: 8 ms
    const double chunk = 20.0 * 20 * 20 * 100 / 1.0e9;
    o.time_limit_s = 14.5 * chunk;
    Run r = serve({req(20, 1, 0.0, 1000), req(20, 2, 0.0, 100000), req(20, 3, 0.0, 1000),
                   req(20, 4, 1.0e6, 100)}, o);  // the last never arrives
    CHECK(r.res[0].status == LBM_REQ_BUDGET);
    CHECK(r.res[1].status == LBM_REQ_UNFINISHED && r.res[2].status == LBM_REQ_PENDING);
    CHECK(r.res[3].status == LBM_REQ_PENDING);
    CHECK(r.st.completed == 1 && r.st.arrived == 3);
    const double t0 = r.res[0].finish_s - r.res[0].arrival_s;
    CHECK(r.st.p50_s == t0 && r.st.p99_s == t0);
    // all arrived: {t0, wall, wall}
    CHECK(std::fabs(r.st.p50_all_s - r.st.wall_s) < 1e-15);
    CHECK(r.st.p50_all_s > r.st.p50_s);
    CHECK(std::fabs(r.st.p99_all_s - r.st.wall_s) < 1e-15);
    CHECK(r.st.mean_queue_s == 0.0);
    CHECK(std::fabs(r.st.mean_service_s - t0) < 1e-15);
    // queueing shows when requests wait: with arrivals at once and one slot,
    // the second completed request queued for the first one's whole run
    auto o2 = options();
    o2.max_resident = 1;
    Run q = serve({req(10, 1, 0.0, 300), req(10, 2, 0.0, 300)}, o2);
    CHECK(q.st.completed == 2);
    const double s0 = q.res[0].finish_s - q.res[0].admit_s;
    CHECK(std::fabs(q.st.mean_queue_s - 0.5 * (q.res[1].admit_s - q.res[1].arrival_s)) < 1e-15);
    CHECK(q.res[1].admit_s - q.res[1].arrival_s >= s0);
}

void test_argument_checks() {
    auto o = options();
    lbm_request q = req(8, 1, 0.0, 100, 1e-3, 150);  // window not a multiple of the chunk
    lbm_result res{};
    CHECK(lbm::serve_args_error(&q, 1, &o, &res) != nullptr);
    q.window = 200;
    CHECK(lbm::serve_args_error(&q, 1, &o, &res) == nullptr);
    o.chunk_steps = 0;
    CHECK(lbm::serve_args_error(&q, 1, &o, &res) != nullptr);
}

}  // namespace

int main() {
    test_budget_status_and_alone_vs_mix();
    test_convergence_rule();
    test_hbm_budget_first_fit_and_rejection();
    test_resident_caps();
    test_divergence_retires_only_the_sick();
    test_time_limit_and_arrivals();
    test_time_to_result_statistics();
    test_allocation_failure_defers_or_rejects();
    test_censored_latency_and_queueing();
    test_argument_checks();
    if (failures) {
        std::printf("%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("all checks passed\n");
    return 0;
}
