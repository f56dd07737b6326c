// Host-only stand-in for the engine's C ABI under the serving loop
// (inflation-easy_amd/csrc/ie_serve.cpp), for ThreadSanitizer / AddressSanitizer
// builds of that loop on the CPU.  A "run" holds host memory sized from its
// lattice, its steps advance a scale factor, its compute calls sleep; the stub
// checks, and aborts with a message when one is broken:
//   - lifecycle calls (ie_create, ie_destroy, ie_batch_create / _launch /
//     _wait / _destroy) come from one thread, the pump's;
//   - a batch is launched only on the runs it was built from (alive, same
//     generation: a run destroyed and another created at its address is caught);
//   - no compute call runs on a run that a launched batch is stepping, and no run
//     is destroyed while a compute call is on it;
//   - the device footprint of the live runs never exceeds the budget the driver
//     below gives the server.
// Failures are injected per request: seed < 0 fails the initial conditions,
// V0 < 0 raises the kernel table-range flag during stepping, Nend < 0 fails the
// deltaN phase.  main() drives a serving stream of the benchmark's mix and
// checks that every request comes back with the expected status and that the
// footprint returns to zero.
#include <assert.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "ie_api.h"

#define CHECK(c, msg)                                                         \
    do {                                                                      \
        if (!(c)) {                                                           \
            fprintf(stderr, "serve_stub: invariant broken: %s (%s:%d)\n", msg, __FILE__, __LINE__); \
            abort();                                                          \
        }                                                                     \
    } while (0)

struct ie_run {
    uint64_t gen = 0;
    ie_config cfg{};
    std::vector<unsigned char> mem;      // stands for the device state
    ie_background bg{};
    std::atomic<int> busy{0};            // compute calls in progress
    std::atomic<int> stepping{0};        // a launched batch is stepping this run
    size_t footprint = 0;
};

struct ie_batch {
    std::vector<std::pair<ie_run *, uint64_t>> runs;
    int launched_steps = 0;
};

namespace {

std::mutex g_mu;
std::map<ie_run *, uint64_t> g_live;     // live runs and their generation
uint64_t g_gen = 0;
size_t g_outstanding = 0, g_budget = 0, g_peak = 0;
std::atomic<bool> g_have_pump{false};
std::thread::id g_pump;

void lifecycle_thread()
{
    if (!g_have_pump.load()) {
        g_pump = std::this_thread::get_id();
        g_have_pump.store(true);
    }
    CHECK(std::this_thread::get_id() == g_pump, "lifecycle call off the pump thread");
}

void nap_us(int us) { std::this_thread::sleep_for(std::chrono::microseconds(us)); }

struct Busy {
    ie_run *r;
    explicit Busy(ie_run *run) : r(run)
    {
        CHECK(r != nullptr, "null run");
        {
            std::lock_guard<std::mutex> g(g_mu);
            auto it = g_live.find(r);
            CHECK(it != g_live.end() && it->second == r->gen, "compute call on a destroyed run");
        }
        CHECK(r->stepping.load() == 0, "compute call on a run a batch is stepping");
        r->busy.fetch_add(1);
    }
    ~Busy() { r->busy.fetch_sub(1); }
};

size_t stub_footprint(const ie_config *c)
{
    const size_t gs = (size_t)c->N * c->N * c->N;
    size_t b = gs * (c->precision == IE_SPLIT ? 40 : 56);
    if (c->calculate_sigw) b += gs * 72;
    return b + 4096;
}

}  // namespace

extern "C" {

size_t ie_footprint_bytes(const ie_config *c) { return c && c->N > 0 ? stub_footprint(c) : 0; }

ie_status ie_create(const ie_config *cfg, ie_run **out)
{
    lifecycle_thread();
    ie_run *r = new ie_run();
    r->cfg = *cfg;
    r->footprint = stub_footprint(cfg);
    r->mem.assign(std::min<size_t>(r->footprint / 4096, 1 << 20), 0);
    r->bg.a = 1.0;
    r->bg.astep = 1.0;
    {
        std::lock_guard<std::mutex> g(g_mu);
        r->gen = ++g_gen;
        g_live[r] = r->gen;
        g_outstanding += r->footprint;
        g_peak = std::max(g_peak, g_outstanding);
        CHECK(g_outstanding <= g_budget, "live runs exceed the server's budget");
    }
    *out = r;
    return IE_OK;
}

ie_status ie_destroy(ie_run *r)
{
    lifecycle_thread();
    CHECK(r->busy.load() == 0, "run destroyed during a compute call");
    CHECK(r->stepping.load() == 0, "run destroyed while a batch steps it");
    {
        std::lock_guard<std::mutex> g(g_mu);
        auto it = g_live.find(r);
        CHECK(it != g_live.end(), "double destroy");
        g_live.erase(it);
        g_outstanding -= r->footprint;
    }
    delete r;
    return IE_OK;
}

ie_status ie_get_config(const ie_run *r, ie_config *out)
{
    *out = r->cfg;
    return IE_OK;
}

ie_status ie_initialize(ie_run *r, const ie_ic_params *ic, ie_background *out)
{
    Busy b(r);
    nap_us(50 + r->cfg.N * 4);
    for (size_t q = 0; q < r->mem.size(); q += 4096) r->mem[q] = (unsigned char)ic->seed;
    if (ic->seed < 0) return IE_ERR_TABLE_RANGE;
    r->bg.t = 0.0;
    r->bg.a = 1.0;
    r->bg.ad = 0.6;
    if (out) *out = r->bg;
    return IE_OK;
}

ie_status ie_evolve_fields(ie_run *r, double d)
{
    Busy b(r);
    r->bg.t += d;
    r->bg.a += d * r->bg.ad;
    return IE_OK;
}

ie_status ie_get_background(ie_run *r, ie_background *out)
{
    Busy b(r);
    *out = r->bg;
    return IE_OK;
}

ie_status ie_set_background(ie_run *r, const ie_background *in)
{
    Busy b(r);
    r->bg = *in;
    return IE_OK;
}

ie_status ie_deltaN_init(ie_run *r)
{
    Busy b(r);
    nap_us(30);
    return IE_OK;
}

ie_status ie_evolve_fieldsN(ie_run *r, double)
{
    Busy b(r);
    return IE_OK;
}

ie_status ie_get_phiref(ie_run *r, double *out)
{
    Busy b(r);
    *out = 0.1;
    return IE_OK;
}

ie_status ie_deltaN_steps(ie_run *r, double, int64_t n)
{
    Busy b(r);
    nap_us((int)std::min<int64_t>(n / 10, 2000) + r->cfg.N);
    return r->cfg.Nend < 0 ? IE_ERR_TABLE_RANGE : IE_OK;
}

ie_status ie_deltaN_zeta(ie_run *r, double, float *nmean)
{
    Busy b(r);
    *nmean = 0.0f;
    return IE_OK;
}

ie_status ie_deltaN_zero_unfinished(ie_run *r, double)
{
    Busy b(r);
    return IE_OK;
}

ie_status ie_deltaN_center(ie_run *r)
{
    Busy b(r);
    return IE_OK;
}

int ie_spectrum_numbins(int N) { return (int)(1.73205 * (N / 2)) + 1; }

static void fill_spectrum(ie_run *r, double scale, double *p, int *counts, double *vals)
{
    const int nb = ie_spectrum_numbins(r->cfg.N);
    for (int q = 0; q < nb; ++q) {
        p[q] = q;
        counts[q] = 1 + q;
        vals[q] = scale * (1.0 + q) * (r->cfg.precision == IE_SPLIT ? 1.0 + 1e-7 : 1.0);
    }
}

ie_status ie_spectrum(ie_run *r, int, double, double *p, int *counts, double *vals)
{
    Busy b(r);
    nap_us(20 + r->cfg.N);
    fill_spectrum(r, 1.0, p, counts, vals);
    return IE_OK;
}

ie_status ie_gw_spectrum(ie_run *r, int hdot, double *p, int *counts, double *vals)
{
    Busy b(r);
    if (!r->cfg.calculate_sigw) return IE_ERR_STATE;
    nap_us(40 + r->cfg.N);
    fill_spectrum(r, hdot ? 2.0 : 0.5, p, counts, vals);
    return IE_OK;
}

ie_status ie_modes(int N, double, double, double *out)
{
    const int nb = (int)(sqrt(3.0) * (N / 2)) + 1;
    for (int q = 0; q < nb; ++q) out[q] = 0.5 + q;
    return IE_OK;
}

ie_status ie_batch_create(ie_run *const *runs, int nruns, ie_batch **out)
{
    lifecycle_thread();
    ie_batch *b = new ie_batch();
    std::lock_guard<std::mutex> g(g_mu);
    for (int q = 0; q < nruns; ++q) {
        auto it = g_live.find(runs[q]);
        CHECK(it != g_live.end(), "batch built on a destroyed run");
        b->runs.emplace_back(runs[q], it->second);
    }
    *out = b;
    return IE_OK;
}

ie_status ie_batch_launch(ie_batch *b, int nsteps, int)
{
    lifecycle_thread();
    {
        std::lock_guard<std::mutex> g(g_mu);
        for (auto &m : b->runs) {
            auto it = g_live.find(m.first);
            CHECK(it != g_live.end() && it->second == m.second,
                  "batch launched on a run it was not built from (destroyed, or another run at its address)");
        }
    }
    for (auto &m : b->runs) {
        CHECK(m.first->busy.load() == 0, "batch launched on a run with a compute call in progress");
        m.first->stepping.fetch_add(1);
    }
    b->launched_steps = nsteps;
    return IE_OK;
}

ie_status ie_batch_wait(ie_batch *b)
{
    lifecycle_thread();
    int raised = 0;
    for (auto &m : b->runs) {
        ie_run *r = m.first;
        ie_background &g = r->bg;
        for (int s = 0; s < b->launched_steps && !g.done; ++s) {
            g.a *= 1.0 + 0.6 * r->cfg.dt;
            g.steps += 1;
            if (r->cfg.V0 < 0 && g.steps > 20) g.error = IE_ERR_TABLE_RANGE;
            g.done = !(g.a <= r->cfg.af);
        }
        if (g.error) raised = 1;
        nap_us(b->launched_steps / 20 + 1);
    }
    for (auto &m : b->runs) m.first->stepping.fetch_sub(1);
    b->launched_steps = 0;
    return raised ? IE_ERR_TABLE_RANGE : IE_OK;
}

ie_status ie_batch_destroy(ie_batch *b)
{
    lifecycle_thread();
    delete b;
    return IE_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ driver

int main(int argc, char **argv)
{
    const int nreq = argc > 1 ? atoi(argv[1]) : 120;
    ie_config c256{};
    c256.N = 256;
    c256.precision = IE_SPLIT;
    c256.calculate_sigw = 1;
    const double big = (double)stub_footprint(&c256);
    g_budget = (size_t)(3.0 * big);   // a few large runs at a time: admission queues
    ie_server *s = nullptr;
    CHECK(ie_server_create(g_budget, 20, 4, &s) == IE_OK, "server create");
    std::mt19937_64 rng(20261203);
    std::discrete_distribution<int> size_of({0.4, 0.3, 0.2, 0.1});
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<double> tx(32), tv(32), tdv(32);
    for (int q = 0; q < 32; ++q) {
        tx[q] = 3.5 - 0.03 * q;
        tv[q] = 1.0;
        tdv[q] = 0.1;
    }
    std::map<int64_t, ie_status> expect;
    std::map<int64_t, int> sigw_of;
    int submitted = 0, got = 0, pumps = 0;
    while (got < nreq) {
        // a few arrivals between rounds (Poisson-like), then one round
        const int arrivals = submitted < nreq ? (int)(u(rng) * 4.0) : 0;
        for (int a = 0; a < arrivals && submitted < nreq; ++a) {
            const int n = submitted++;
            ie_request q{};
            q.cfg.N = 32 << size_of(rng);
            q.cfg.precision = IE_SPLIT;
            q.cfg.numerical_potential = u(rng) < 0.5;
            q.cfg.calculate_sigw = u(rng) < 0.2;
            q.cfg.integrator = IE_LEAPFROG;
            q.cfg.dt = 0.001;
            q.cfg.af = 1.0 + 0.05 * (1 + n % 3);
            q.cfg.V0 = 3e-9;
            q.cfg.rescale_B = 1.0;
            q.cfg.L = 10.0;
            q.cfg.dN = 1e-4;
            q.cfg.Nend = 0.01;
            if (q.cfg.numerical_potential) {
                q.cfg.tab_n = 32;
                q.cfg.tab_x = tx.data();
                q.cfg.tab_v = tv.data();
                q.cfg.tab_dv = tdv.data();
            }
            q.ic.seed = 1000 + n;
            q.deltaN = 1;
            q.fp64_check = n % 20 == 0;
            ie_status want = IE_OK;
            if (n % 37 == 5) q.ic.seed = -1, want = IE_ERR_TABLE_RANGE;     // initial conditions fail
            if (n % 41 == 7) q.cfg.V0 = -1.0, want = IE_ERR_TABLE_RANGE;    // kernel error while stepping
            if (n % 43 == 9) q.cfg.Nend = -1.0, want = IE_ERR_TABLE_RANGE;  // deltaN phase fails
            int64_t id = 0;
            CHECK(ie_server_submit(s, &q, &id) == IE_OK, "submit");
            expect[id] = want;
            sigw_of[id] = q.cfg.calculate_sigw;
        }
        int active = 0, pending = 0;
        CHECK(ie_server_pump(s, &active, &pending) == IE_OK, "pump failed (one failing run must not stall the server)");
        ++pumps;
        ie_result r;
        while (ie_server_pop_result(s, &r) == IE_OK) {
            auto it = expect.find(r.id);
            CHECK(it != expect.end(), "unknown result id");
            if (r.status != it->second) {
                fprintf(stderr, "request %lld: status %d, expected %d\n", (long long)r.id, r.status, it->second);
                abort();
            }
            if (r.status == IE_OK) {
                CHECK(r.nbins > 0, "P_zeta missing");
                CHECK((r.gw_nbins > 0) == (sigw_of[r.id] != 0), "Omega_GW present exactly for SIGW runs");
                CHECK(r.done_s >= r.admit_s && r.admit_s >= r.submit_s, "clock order");
            }
            expect.erase(it);
            ++got;
        }
        CHECK(pumps < 200000, "serving loop does not finish");
    }
    ie_server_destroy(s);
    CHECK(expect.empty(), "missing results");
    CHECK(g_outstanding == 0, "footprint not returned");
    CHECK(g_live.empty(), "runs left alive");
    printf("serve_stub ok: %d requests, %d rounds, peak footprint %.2f of budget\n", nreq, pumps,
           (double)g_peak / (double)g_budget);
    return 0;
}
