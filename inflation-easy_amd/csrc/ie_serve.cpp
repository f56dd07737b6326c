// ie_serve.cpp -- request loop with HBM-footprint admission and continuous
// batching, built on the public C ABI (include/ie_api.h).
//
// A request is a whole run: initial conditions, the inflation loop up to af,
// then (optionally) the deltaN map and P_zeta, as the reference program does
// for one lattice per process (src/main.cpp:103-146).  Here runs of any size
// share the GPU:
//   - admission: FIFO; a request is admitted when its device footprint fits the
//     budget still free; its initial conditions are generated on a worker thread;
//   - stepping: the inflating runs of one storage kind and lattice size form one
//     ie_batch, rebuilt only when its membership changes; each round advances a
//     batch of N^3 lattices by quantum * max(1, 256 / N) steps (one launch per
//     kernel per step, replayed from a graph), so that small runs, whose steps
//     cost little, do not wait for the rounds of large ones;
//   - finishing: a run whose loop test ends leaves the batch and is completed on
//     a worker thread (final leapfrog sync, deltaN phase, zeta spectrum), on its
//     own stream, while the remaining runs keep stepping; its memory is released
//     and the next requests can be admitted.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <future>
#include <mutex>
#include <string>
#include <vector>

#include "ie_api.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    int64_t id = 0;
    ie_request req{};
    std::vector<double> tx, tv, tdv;
    ie_run *run = nullptr, *twin = nullptr;
    size_t footprint = 0;
    double submit = 0, admit = 0;
    std::future<ie_status> init;      // initial conditions (worker)
    std::future<void> finish;         // deltaN + spectrum (worker)
    bool ready = false, inflating = false, finishing = false;
    ie_result result{};
};

double now_s(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

// Initial conditions and the first leapfrog half drift (src/evolution.cpp:935).
ie_status init_run(ie_run *run, const ie_request &q)
{
    ie_background b;
    ie_status st = ie_initialize(run, &q.ic, &b);
    if (st != IE_OK) return st;
    if (q.cfg.integrator != IE_LEAPFROG) return IE_ERR_ARG;
    return ie_evolve_fields(run, 0.5 * q.cfg.dt);
}

// After the loop: final synchronisation (src/evolution.cpp:993-995), then the
// deltaN phase and saveN's zeta spectrum (src/evolution.cpp:1291-1375,
// src/output.cpp:1522-1560).  Writes P_zeta into out[nb].
ie_status finish_run(ie_run *run, const ie_config &c, int deltaN, double *out, int *nb, int64_t *steps, double *a_end)
{
    ie_background b;
    ie_status st = ie_get_background(run, &b);
    if (st != IE_OK) return st;
    *steps = b.steps;
    st = ie_evolve_fields(run, -0.5 * c.dt * pow(b.astep, c.rescale_s - 1.0));
    if (st != IE_OK) return st;
    *a_end = b.a;
    *nb = 0;
    if (!deltaN) return IE_OK;
    if ((st = ie_deltaN_init(run)) != IE_OK) return st;
    if ((st = ie_get_background(run, &b)) != IE_OK) return st;
    b.phiref = 0.0;
    if ((st = ie_set_background(run, &b)) != IE_OK) return st;
    if ((st = ie_evolve_fieldsN(run, 0.5 * c.dN)) != IE_OK) return st;
    double phiref = c.phiref_manual_value;
    if (!c.use_phiref_manual && (st = ie_get_phiref(run, &phiref)) != IE_OK) return st;
    if ((st = ie_get_background(run, &b)) != IE_OK) return st;
    b.phiref = phiref;
    if ((st = ie_set_background(run, &b)) != IE_OK) return st;
    double Ne = 0.0;
    int64_t n = 0;
    while (Ne <= c.Nend) {
        Ne += c.dN;
        ++n;
    }
    if ((st = ie_deltaN_steps(run, c.dN, n)) != IE_OK) return st;
    if ((st = ie_evolve_fieldsN(run, -0.5 * c.dN)) != IE_OK) return st;
    const double cutoff = c.Nend - c.dN;
    float nmean;
    if ((st = ie_deltaN_zeta(run, cutoff, &nmean)) != IE_OK) return st;
    if ((st = ie_deltaN_zero_unfinished(run, cutoff)) != IE_OK) return st;
    if ((st = ie_deltaN_center(run)) != IE_OK) return st;
    const int bins = ie_spectrum_numbins(c.N);
    std::vector<double> p(bins), v(bins);
    std::vector<int> cnt(bins);
    const double norm1 = pow(c.L / c.rescale_B, 3) / pow((double)c.N, 6);
    if ((st = ie_spectrum(run, IE_FIELD_DELTAN, norm1, p.data(), cnt.data(), v.data())) != IE_OK) return st;
    // P_zeta(k) = k^3/(2 pi^2) P_dN on the modes.dat k (notebooks/plot.ipynb cell 9)
    const int nm = (int)(sqrt(3.0) * (c.N / 2)) + 1;
    std::vector<double> k(nm), pz(nm);
    if ((st = ie_modes(c.N, c.L, c.rescale_B, k.data())) != IE_OK) return st;
    *nb = std::min(std::min(bins, nm), 512);
    if ((st = ie_pzeta_deltaN(*nb, k.data(), v.data(), pz.data())) != IE_OK) return st;
    for (int q = 0; q < *nb; ++q) out[q] = pz[q];
    return IE_OK;
}

}  // namespace

struct ie_server {
    size_t budget = 0, used = 0;
    int quantum = 100, threads = 2;
    int64_t next_id = 1;
    Clock::time_point t0 = Clock::now();
    std::deque<Job *> pending;
    std::vector<Job *> live;              // admitted, not yet reported
    std::deque<ie_result> results;
    std::mutex mu;
    struct Group {
        int kind, N;                  // storage (IE_FP64 / IE_SPLIT), lattice size
        ie_batch *batch;
        std::vector<ie_run *> members;
    };
    std::vector<Group> groups;
};

extern "C" {

ie_status ie_server_create(size_t budget, int quantum, int threads, ie_server **out)
{
    if (!out || budget == 0 || quantum < 1) return IE_ERR_ARG;
    ie_server *s = new ie_server();
    s->budget = budget;
    s->quantum = quantum;
    s->threads = threads < 1 ? 1 : threads;
    *out = s;
    return IE_OK;
}

double ie_server_clock(const ie_server *s) { return s ? now_s(s->t0) : 0.0; }

ie_status ie_server_submit(ie_server *s, const ie_request *req, int64_t *id)
{
    if (!s || !req || !id) return IE_ERR_ARG;
    Job *j = new Job();
    j->id = s->next_id++;
    j->req = *req;
    if (req->cfg.numerical_potential) {
        const int n = req->cfg.tab_n;
        if (n < 2 || !req->cfg.tab_x || !req->cfg.tab_v || !req->cfg.tab_dv) {
            delete j;
            return IE_ERR_ARG;
        }
        j->tx.assign(req->cfg.tab_x, req->cfg.tab_x + n);
        j->tv.assign(req->cfg.tab_v, req->cfg.tab_v + n);
        j->tdv.assign(req->cfg.tab_dv, req->cfg.tab_dv + n);
        j->req.cfg.tab_x = j->tx.data();
        j->req.cfg.tab_v = j->tv.data();
        j->req.cfg.tab_dv = j->tdv.data();
    }
    j->footprint = ie_footprint_bytes(&j->req.cfg);
    if (req->fp64_check) {
        ie_config c64 = j->req.cfg;
        c64.precision = IE_FP64;
        j->footprint += ie_footprint_bytes(&c64);
    }
    if (j->footprint == 0 || j->footprint > s->budget) {
        delete j;
        return IE_ERR_OOM;
    }
    j->submit = now_s(s->t0);
    j->result.id = j->id;
    j->result.N = req->cfg.N;
    j->result.submit_s = j->submit;
    j->result.fp64_max_rel_err = -1.0;
    s->pending.push_back(j);
    *id = j->id;
    return IE_OK;
}

static void drop_batches(ie_server *s)
{
    for (auto &g : s->groups)
        if (g.batch) ie_batch_destroy(g.batch);
    s->groups.clear();
}

static void fail_job(ie_server *s, Job *j, ie_status st)
{
    j->result.status = st;
    j->result.done_s = now_s(s->t0);
}

ie_status ie_server_pump(ie_server *s, int *active, int *pending)
{
    if (!s) return IE_ERR_ARG;
    // 1. admission by footprint (FIFO: a large request is not overtaken)
    for (;;) {
        size_t used;
        {
            std::lock_guard<std::mutex> g(s->mu);
            used = s->used;
        }
        if (s->pending.empty() || used + s->pending.front()->footprint > s->budget) break;
        Job *j = s->pending.front();
        s->pending.pop_front();
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->used += j->footprint;
        }
        j->admit = now_s(s->t0);
        j->result.admit_s = j->admit;
        ie_status st = ie_create(&j->req.cfg, &j->run);
        if (st == IE_OK && j->req.fp64_check) {
            ie_config c64 = j->req.cfg;
            c64.precision = IE_FP64;
            st = ie_create(&c64, &j->twin);
        }
        s->live.push_back(j);
        if (st != IE_OK) {
            fail_job(s, j, st);
            j->finishing = true;
            continue;
        }
        j->init = std::async(std::launch::async, [j]() {
            ie_status e = init_run(j->run, j->req);
            if (e == IE_OK && j->twin) e = init_run(j->twin, j->req);
            return e;
        });
    }
    // 2. runs whose initial conditions are ready join the batches
    for (Job *j : s->live) {
        if (j->ready || j->finishing || !j->init.valid()) continue;
        if (j->init.wait_for(std::chrono::seconds(0)) != std::future_status::ready) continue;
        const ie_status st = j->init.get();
        j->ready = true;
        if (st != IE_OK) {
            fail_job(s, j, st);
            j->finishing = true;
            continue;
        }
        j->inflating = true;
    }
    // 3. one round of steps for every inflating run: one batch per storage kind
    //    and lattice size; batches whose members are unchanged are kept
    std::vector<ie_server::Group> want;
    auto add = [&want](int kind, int N, ie_run *r) {
        for (auto &g : want)
            if (g.kind == kind && g.N == N) {
                g.members.push_back(r);
                return;
            }
        want.push_back(ie_server::Group{kind, N, nullptr, {r}});
    };
    for (Job *j : s->live) {
        if (!j->inflating) continue;
        add(j->req.cfg.precision == IE_SPLIT ? IE_SPLIT : IE_FP64, j->req.cfg.N, j->run);
        if (j->twin) add(IE_FP64, j->req.cfg.N, j->twin);
    }
    for (auto &w : want)
        for (auto &g : s->groups)
            if (g.batch && g.kind == w.kind && g.N == w.N && g.members == w.members) {
                w.batch = g.batch;
                g.batch = nullptr;
            }
    drop_batches(s);
    s->groups = std::move(want);
    for (auto &g : s->groups)
        if (!g.batch && ie_batch_create(g.members.data(), (int)g.members.size(), &g.batch) != IE_OK) {
            drop_batches(s);
            return IE_ERR_HIP;
        }
    // all batches enqueued first (one stream each), so that they run together
    ie_status st = IE_OK;
    for (auto &g : s->groups) {
        const int q = s->quantum * std::max(1, 256 / g.N);
        if (st == IE_OK) st = ie_batch_launch(g.batch, q, s->quantum);
    }
    for (auto &g : s->groups) {
        const ie_status w = ie_batch_wait(g.batch);
        if (st == IE_OK) st = w;
    }
    if (st != IE_OK) return st;
    // 4. runs past af leave the batch and are finished on a worker thread
    int finishing_now = 0;
    for (Job *j : s->live)
        if (j->finishing && j->finish.valid()) ++finishing_now;
    for (Job *j : s->live) {
        if (!j->inflating) continue;
        ie_background b, bt;
        if (ie_get_background(j->run, &b) != IE_OK) return IE_ERR_HIP;
        bool done = b.done != 0;
        if (j->twin) {
            if (ie_get_background(j->twin, &bt) != IE_OK) return IE_ERR_HIP;
            done = done && bt.done != 0;
        }
        if (!done || finishing_now >= s->threads) continue;
        j->inflating = false;
        j->finishing = true;
        ++finishing_now;
        j->finish = std::async(std::launch::async, [s, j]() {
            ie_result &r = j->result;
            int nb = 0;
            ie_status st = finish_run(j->run, j->req.cfg, j->req.deltaN, r.pzeta, &nb, &r.steps, &r.a_end);
            r.nbins = nb;
            if (st == IE_OK && j->twin) {
                std::vector<double> pz(512);
                int nb2 = 0;
                int64_t st2;
                double a2;
                st = finish_run(j->twin, j->req.cfg, j->req.deltaN, pz.data(), &nb2, &st2, &a2);
                double e = 0.0;
                for (int q = 1; q < std::min(nb, nb2); ++q)
                    if (pz[q] > 0) e = std::max(e, fabs(r.pzeta[q] - pz[q]) / pz[q]);
                r.fp64_max_rel_err = e;
            }
            r.status = st;
            ie_destroy(j->run);
            j->run = nullptr;
            if (j->twin) ie_destroy(j->twin);
            j->twin = nullptr;
            std::lock_guard<std::mutex> g(s->mu);
            s->used -= j->footprint;
            r.done_s = now_s(s->t0);
        });
    }
    // 5. collect completed results
    std::vector<Job *> keep;
    for (Job *j : s->live) {
        bool complete = false;
        if (j->finishing) {
            if (!j->finish.valid()) complete = true;   // failed at admission / init
            else if (j->finish.wait_for(std::chrono::seconds(0)) == std::future_status::ready) {
                j->finish.get();
                complete = true;
            }
        }
        if (complete) {
            if (j->run) ie_destroy(j->run);
            if (j->twin) ie_destroy(j->twin);
            if (!j->finish.valid() && j->result.status != IE_OK) {
                std::lock_guard<std::mutex> g(s->mu);
                s->used -= j->footprint;
            }
            s->results.push_back(j->result);
            delete j;
        } else {
            keep.push_back(j);
        }
    }
    s->live.swap(keep);
    if (active) *active = (int)s->live.size();
    if (pending) *pending = (int)s->pending.size();
    return IE_OK;
}

ie_status ie_server_pop_result(ie_server *s, ie_result *out)
{
    if (!s || !out) return IE_ERR_ARG;
    if (s->results.empty()) return IE_ERR_STATE;
    *out = s->results.front();
    s->results.pop_front();
    return IE_OK;
}

ie_status ie_server_destroy(ie_server *s)
{
    if (!s) return IE_OK;
    for (Job *j : s->live) {
        if (j->init.valid()) j->init.wait();
        if (j->finish.valid()) j->finish.wait();
    }
    drop_batches(s);
    for (Job *j : s->live) {
        if (j->run) ie_destroy(j->run);
        if (j->twin) ie_destroy(j->twin);
        delete j;
    }
    for (Job *j : s->pending) delete j;
    delete s;
    return IE_OK;
}

}  // extern "C"
