// ie_serve.cpp -- request loop with HBM-footprint admission and continuous
// batching, built on the public C ABI (include/ie_api.h).
//
// A request is a whole run: initial conditions, the inflation loop up to af,
// then the GW spectra at the end of inflation (SIGW runs) and the deltaN map and
// P_zeta, as the reference program does for one lattice per process
// (src/main.cpp:103-146).  Here runs of any size share the GPU:
//   - admission: FIFO; a request is admitted when its device footprint fits the
//     budget still free; its initial conditions are generated by a worker;
//   - stepping: the inflating runs of one storage kind and lattice size form one
//     ie_batch, rebuilt only when its membership changes; each round advances a
//     batch of N^3 lattices by quantum * max(1, 256 / N) steps (one launch per
//     kernel per step, replayed from a graph), so that small runs, whose steps
//     cost little, do not wait for the rounds of large ones; every batch of a
//     round is enqueued on its own stream before any is waited for;
//   - finishing: a run whose loop test ends leaves the batch and is completed by
//     a worker (final leapfrog sync, GW spectra and Omega_GW, deltaN phase, zeta
//     spectrum) on its own stream while the remaining runs keep stepping.
//
// Threads.  The thread that calls ie_server_pump makes every lifecycle call:
// ie_create, ie_destroy, ie_batch_create / ie_batch_destroy and the batches'
// graph capture.  A fixed pool of workers only enqueues work on the streams of
// the runs handed to it and waits for it, so no device allocation, free, stream
// destruction or graph capture runs beside another thread's.  A run's footprint
// is returned to the budget once, on the pump thread, when its runs are
// destroyed.
//
// Batches are matched across rounds by their members' request ids, never by run
// pointers: the address of a destroyed run is reused by later allocations, and a
// batch kept for a new run at an old address would step it with the old run's
// device descriptor (its potential and table).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "ie_api.h"

namespace {

using Clock = std::chrono::steady_clock;

double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

// Fixed pool of worker threads; finishing tasks (which free memory for the
// next admissions) are taken before initial conditions.  The destructor runs
// the tasks still queued, then joins.
class Pool {
  public:
    explicit Pool(int n)
    {
        for (int q = 0; q < n; ++q) threads_.emplace_back([this]() { loop(); });
    }
    ~Pool()
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
    }
    void post(std::function<void()> f, bool urgent)
    {
        {
            std::lock_guard<std::mutex> g(mu_);
            (urgent ? urgent_ : normal_).push_back(std::move(f));
        }
        cv_.notify_one();
    }

  private:
    void loop()
    {
        for (;;) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this]() { return stop_ || !urgent_.empty() || !normal_.empty(); });
                std::deque<std::function<void()>> &q = !urgent_.empty() ? urgent_ : normal_;
                if (q.empty()) return;   // stopping, nothing left
                f = std::move(q.front());
                q.pop_front();
            }
            f();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::function<void()>> urgent_, normal_;
    bool stop_ = false;
    std::vector<std::thread> threads_;
};

// What a run leaves at its end (filled by finish_run).
struct RunOutputs {
    int64_t steps = 0;
    double a_end = 0.0;
    int nb = 0, gw_nb = 0;
    double pzeta[512];
    double omega_gw[512];
};

struct Job {
    enum Stage { QUEUED, INIT, INFLATING, FINISH, COMPLETE };
    int64_t id = 0;
    ie_request req{};
    std::vector<double> tx, tv, tdv;
    ie_run *run = nullptr, *twin = nullptr;
    size_t footprint = 0;
    Stage stage = QUEUED;
    // a worker's task on this job: its status and end time are written before
    // task_done is set (release); the pump reads them after seeing it (acquire)
    std::atomic<int> task_done{0};
    ie_status task_status = IE_OK;
    double task_end_s = 0.0;
    ie_result result{};
};

// Initial conditions and the first leapfrog half drift (src/evolution.cpp:935).
ie_status init_run(ie_run *run, const ie_request &q)
{
    ie_background b;
    ie_status st = ie_initialize(run, &q.ic, &b);
    if (st != IE_OK) return st;
    return ie_evolve_fields(run, 0.5 * q.cfg.dt);
}

// TT tensor spectra at the end of inflation and Omega_GW on the modes.dat k
// (spectraGW / spectraGWdot, src/output.cpp:631-648; notebooks/plot.ipynb cell
// 23), H and a as sf.dat writes them (src/output.cpp:272-293).
ie_status gw_outputs(ie_run *run, const ie_config &c, const ie_background &b, const double *k, RunOutputs *o)
{
    const int bins = ie_spectrum_numbins(c.N);
    std::vector<double> p(bins), vh(bins), vhd(bins);
    std::vector<int> cnt(bins);
    ie_status st = ie_gw_spectrum(run, 0, p.data(), cnt.data(), vh.data());
    if (st != IE_OK) return st;
    if ((st = ie_gw_spectrum(run, 1, p.data(), cnt.data(), vhd.data())) != IE_OK) return st;
    const int kept = std::min((int)floor(0.8 * bins), 512);   // bins >= floor(0.8 numbins) are empty
    const double hubble = b.ad * c.rescale_B * pow(b.a, c.rescale_s - 2.0);
    if ((st = ie_omega_gw(kept, k, vh.data(), vhd.data(), b.a, hubble, o->omega_gw)) != IE_OK) return st;
    o->gw_nb = kept;
    return IE_OK;
}

// After the loop: final synchronisation (src/evolution.cpp:993-995), the GW
// outputs of a SIGW run, then the deltaN phase and saveN's zeta spectrum
// (src/evolution.cpp:1291-1375, src/output.cpp:1522-1560).
ie_status finish_run(ie_run *run, int deltaN, RunOutputs *o)
{
    ie_config c;
    ie_status st = ie_get_config(run, &c);
    if (st != IE_OK) return st;
    ie_background b;
    st = ie_get_background(run, &b);
    if (st != IE_OK) return st;
    o->steps = b.steps;
    if ((st = ie_evolve_fields(run, -0.5 * c.dt * pow(b.astep, c.rescale_s - 1.0))) != IE_OK) return st;
    if ((st = ie_get_background(run, &b)) != IE_OK) return st;
    o->a_end = b.a;
    const int nm = (int)(sqrt(3.0) * (c.N / 2)) + 1;
    std::vector<double> k(nm);
    if ((st = ie_modes(c.N, c.L, c.rescale_B, k.data())) != IE_OK) return st;
    if (c.calculate_sigw && (st = gw_outputs(run, c, b, k.data(), o)) != IE_OK) return st;
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
    o->nb = std::min(std::min(bins, nm), 512);
    return ie_pzeta_deltaN(o->nb, k.data(), v.data(), o->pzeta);
}

// Largest relative difference of a against the reference b over entries 1 .. n-1
// with b > 0.
double max_rel(const double *a, const double *b, int n)
{
    double e = 0.0;
    for (int q = 1; q < n; ++q)
        if (b[q] > 0) e = std::max(e, fabs(a[q] - b[q]) / b[q]);
    return e;
}

}  // namespace

struct ie_server {
    size_t budget = 0, used = 0;   // pump thread only
    int quantum = 100;
    int graph_steps = -1;          // steps per graph launch of the batches (-1: the quantum, 0: direct launches)
    int64_t next_id = 1;
    Clock::time_point t0 = Clock::now();
    std::deque<Job *> pending;
    std::vector<Job *> live;       // admitted, not yet reported
    std::deque<ie_result> results;
    struct Group {
        int kind = 0, N = 0;               // storage (IE_FP64 / IE_SPLIT), lattice size
        ie_batch *batch = nullptr;
        std::vector<ie_run *> members;
        std::vector<int64_t> keys;         // 2 id (+1 for an FP64 twin), in member order
    };
    std::vector<Group> groups;
    Pool *pool = nullptr;
};

namespace {

void drop_batches(ie_server *s)
{
    for (auto &g : s->groups)
        if (g.batch) ie_batch_destroy(g.batch);
    s->groups.clear();
}

void complete(ie_server *s, Job *j, ie_status st)
{
    j->result.status = st;
    j->result.done_s = since(s->t0);
    j->stage = Job::COMPLETE;
}

// 1. admission by footprint (FIFO: a large request is not overtaken)
void admit(ie_server *s)
{
    while (!s->pending.empty() && s->used + s->pending.front()->footprint <= s->budget) {
        Job *j = s->pending.front();
        s->pending.pop_front();
        s->used += j->footprint;
        j->result.admit_s = since(s->t0);
        s->live.push_back(j);
        ie_status st = ie_create(&j->req.cfg, &j->run);
        if (st == IE_OK && j->req.fp64_check) {
            ie_config c64 = j->req.cfg;
            c64.precision = IE_FP64;
            st = ie_create(&c64, &j->twin);
        }
        if (st != IE_OK) {
            complete(s, j, st);
            continue;
        }
        j->stage = Job::INIT;
        s->pool->post(
            [j, s]() {
                ie_status e = init_run(j->run, j->req);
                if (e == IE_OK && j->twin) e = init_run(j->twin, j->req);
                j->task_status = e;
                j->task_end_s = since(s->t0);
                j->task_done.store(1, std::memory_order_release);
            },
            false);
    }
}

// 3. one round of steps for every inflating run: one batch per storage kind and
// lattice size; batches whose members are unchanged are kept.  Returns a HIP
// failure; a run that raised a kernel error is completed with it.
ie_status step_round(ie_server *s)
{
    std::vector<ie_server::Group> want;
    auto add = [&want](int kind, int N, ie_run *r, int64_t key) {
        for (auto &g : want)
            if (g.kind == kind && g.N == N) {
                g.members.push_back(r);
                g.keys.push_back(key);
                return;
            }
        ie_server::Group g;
        g.kind = kind;
        g.N = N;
        g.members.push_back(r);
        g.keys.push_back(key);
        want.push_back(std::move(g));
    };
    for (Job *j : s->live) {
        if (j->stage != Job::INFLATING) continue;
        add(j->req.cfg.precision == IE_SPLIT ? IE_SPLIT : IE_FP64, j->req.cfg.N, j->run, 2 * j->id);
        if (j->twin) add(IE_FP64, j->req.cfg.N, j->twin, 2 * j->id + 1);
    }
    for (auto &w : want)
        for (auto &g : s->groups)
            if (g.batch && g.kind == w.kind && g.N == w.N && g.keys == w.keys) {
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
    ie_status st = IE_OK;
    for (auto &g : s->groups) {
        const int q = s->quantum * std::max(1, 256 / g.N);
        if (st == IE_OK) st = ie_batch_launch(g.batch, q, s->graph_steps < 0 ? s->quantum : s->graph_steps);
    }
    for (auto &g : s->groups) {
        const ie_status w = ie_batch_wait(g.batch);
        if (st == IE_OK && w != IE_ERR_TABLE_RANGE) st = w;
    }
    if (st != IE_OK) return st;
    for (Job *j : s->live) {
        if (j->stage != Job::INFLATING) continue;
        ie_background b;
        int err = 0;
        for (ie_run *r : {j->run, j->twin}) {
            if (!r) continue;
            if (ie_get_background(r, &b) != IE_OK) return IE_ERR_HIP;
            if (b.error) err = b.error;
        }
        if (err) complete(s, j, (ie_status)err);
    }
    return IE_OK;
}

// 4. runs past af leave their batch and are finished by a worker
void start_finishing(ie_server *s)
{
    for (Job *j : s->live) {
        if (j->stage != Job::INFLATING) continue;
        ie_background b;
        bool done = ie_get_background(j->run, &b) == IE_OK && b.done != 0;
        if (done && j->twin) done = ie_get_background(j->twin, &b) == IE_OK && b.done != 0;
        if (!done) continue;
        j->stage = Job::FINISH;
        j->task_done.store(0, std::memory_order_relaxed);
        s->pool->post(
            [j, s]() {
                RunOutputs o, t;
                ie_result &r = j->result;
                ie_status e = finish_run(j->run, j->req.deltaN, &o);
                r.steps = o.steps;
                r.a_end = o.a_end;
                r.nbins = o.nb;
                memcpy(r.pzeta, o.pzeta, sizeof(double) * (size_t)o.nb);
                r.gw_nbins = o.gw_nb;
                memcpy(r.omega_gw, o.omega_gw, sizeof(double) * (size_t)o.gw_nb);
                if (e == IE_OK && j->twin) {
                    e = finish_run(j->twin, j->req.deltaN, &t);
                    r.fp64_max_rel_err = max_rel(o.pzeta, t.pzeta, std::min(o.nb, t.nb));
                    if (o.gw_nb > 0) r.gw_fp64_max_rel_err = max_rel(o.omega_gw, t.omega_gw, std::min(o.gw_nb, t.gw_nb));
                }
                j->task_status = e;
                j->task_end_s = since(s->t0);
                j->task_done.store(1, std::memory_order_release);
            },
            true);
    }
}

}  // namespace

extern "C" {

ie_status ie_run_outputs(ie_run *run, int deltaN, ie_result *out)
{
    if (!run || !out) return IE_ERR_ARG;
    RunOutputs o;
    const ie_status st = finish_run(run, deltaN, &o);
    out->steps = o.steps;
    out->a_end = o.a_end;
    out->nbins = o.nb;
    memcpy(out->pzeta, o.pzeta, sizeof(double) * (size_t)o.nb);
    out->gw_nbins = o.gw_nb;
    memcpy(out->omega_gw, o.omega_gw, sizeof(double) * (size_t)o.gw_nb);
    return st;
}

ie_status ie_server_create(size_t budget, int quantum, int threads, ie_server **out)
{
    if (!out || budget == 0 || quantum < 1) return IE_ERR_ARG;
    ie_server *s = new ie_server();
    s->budget = budget;
    s->quantum = quantum;
    s->pool = new Pool(threads < 1 ? 1 : threads);
    *out = s;
    return IE_OK;
}

double ie_server_clock(const ie_server *s) { return s ? since(s->t0) : 0.0; }

ie_status ie_server_set_graph_steps(ie_server *s, int graph_steps)
{
    if (!s || graph_steps < 0) return IE_ERR_ARG;
    s->graph_steps = graph_steps;
    return IE_OK;
}

ie_status ie_server_submit(ie_server *s, const ie_request *req, int64_t *id)
{
    if (!s || !req || !id) return IE_ERR_ARG;
    if (req->cfg.integrator != IE_LEAPFROG) return IE_ERR_ARG;   // batches step the leapfrog loop
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
    j->result.id = j->id;
    j->result.N = req->cfg.N;
    j->result.submit_s = since(s->t0);
    j->result.fp64_max_rel_err = -1.0;
    j->result.gw_fp64_max_rel_err = -1.0;
    s->pending.push_back(j);
    *id = j->id;
    return IE_OK;
}

ie_status ie_server_pump(ie_server *s, int *active, int *pending)
{
    if (!s) return IE_ERR_ARG;
    admit(s);
    // 2. runs whose initial conditions are ready join the batches
    for (Job *j : s->live)
        if (j->stage == Job::INIT && j->task_done.load(std::memory_order_acquire)) {
            if (j->task_status != IE_OK) complete(s, j, j->task_status);
            else j->stage = Job::INFLATING;
        }
    const ie_status st = step_round(s);
    if (st != IE_OK) return st;
    start_finishing(s);
    // 5. completed runs are destroyed here, their footprint returned, their
    // results queued
    std::vector<Job *> keep;
    for (Job *j : s->live) {
        if (j->stage == Job::FINISH && j->task_done.load(std::memory_order_acquire)) {
            j->result.status = j->task_status;
            j->result.done_s = j->task_end_s;
            j->stage = Job::COMPLETE;
        }
        if (j->stage != Job::COMPLETE) {
            keep.push_back(j);
            continue;
        }
        if (j->run) ie_destroy(j->run);
        if (j->twin) ie_destroy(j->twin);
        s->used -= j->footprint;
        s->results.push_back(j->result);
        delete j;
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
    delete s->pool;   // runs the queued tasks, then joins the workers
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
