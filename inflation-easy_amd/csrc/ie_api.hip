// ie_api.hip -- C ABI of the gfx950 lattice engine (declared in include/ie_api.h).
//
// A run handle owns the device memory of one lattice and one HIP stream.  Every
// entry point validates its arguments, enqueues work on the run's stream and
// returns once its host-visible outputs are complete; errors are returned as
// ie_status with a message kept in the handle (the reference prints and exits,
// e.g. src/potential.cpp:19-27; the host driver maps the status back to that).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "ie_fft.h"
#include "ie_internal.h"

using namespace ie;

struct ie_run {
    ie_config cfg;
    int N = 0, logN = 0;
    size_t gs = 0;
    double dx = 0.0;
    hipStream_t stream = nullptr;
    Fields F{};
    Pot P{};
    DevBg *bg = nullptr;         // device
    DevBg *bg_host = nullptr;    // pinned mirror
    double *partials = nullptr;  // reduction partials
    double *table = nullptr;     // x, V, V' (device)
    double *scratch = nullptr;   // N^3 doubles: split staging, then the deltaN map
    FftPlan *fft = nullptr;
    double *work = nullptr;      // 4096 doubles: small reductions
    float *nmean = nullptr;      // saveN mean (device)
    unsigned long long *hist = nullptr;
    int hist_bins = 0;
    double *dn_f = nullptr, *dn_fd = nullptr;  // FP64 deltaN state (split runs)
    double *dn5[3] = {nullptr, nullptr, nullptr};   // deltaN RK45 trial state (f, fd, N)
    int in_deltaN = 0;          // FP64 state (dn_f / F.f): deltaN and post-inflation phases
    int in_post = 0;            // post-inflation phase
    int pinned = 0;             // kernel shapes set by ie_set_shapes: not tuned
    ie_post_params post{};
    double *post_f2 = nullptr, *post_fd2 = nullptr;   // second Phi, Phi' buffers (split runs)
    RkWork *rk4 = nullptr, *rk45 = nullptr;    // RK stage buffers, made on first use
    std::vector<double> htab_x, htab_v, htab_dv;   // host copy of the table
    std::vector<void *> allocs;
    std::string err;
};

namespace {

thread_local std::string g_create_error;

ie_status fail(ie_run *r, ie_status st, const std::string &msg)
{
    if (r) r->err = msg;
    else g_create_error = msg;
    return st;
}

#define IE_HIP(r, call)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail((r), IE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

ie_status dev_alloc(ie_run *r, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(r, IE_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    r->allocs.push_back(*p);
    return IE_OK;
}

ie_status sync(ie_run *r)
{
    IE_HIP(r, hipStreamSynchronize(r->stream));
    return IE_OK;
}

ie_status pull_bg(ie_run *r)
{
    IE_HIP(r, hipMemcpyAsync(r->bg_host, r->bg, sizeof(DevBg), hipMemcpyDeviceToHost, r->stream));
    return sync(r);
}

ie_status push_bg(ie_run *r)
{
    IE_HIP(r, hipMemcpyAsync(r->bg, r->bg_host, sizeof(DevBg), hipMemcpyHostToDevice, r->stream));
    return sync(r);
}

ie_status check_kernel_error(ie_run *r)
{
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    if (r->bg_host->error) {
        r->bg_host->error = 0;
        push_bg(r);
        return fail(r, IE_ERR_TABLE_RANGE, "field value outside the numerical potential range");
    }
    return IE_OK;
}

// Accept test and step-size proposal of an RK45 trial step (src/evolution.cpp:
// 880-904, 1250-1267): accept when the weighted RMS error is <= 1; grow by
// safety * err^{-1/5} (within [0.1, 5]) after an accepted step, shrink by at most
// 0.5 (at least 0.1) after a rejection, clamped to [max(rk45_min_dt, 1e-16), hmax].
int rk45_propose(const ie_config &c, double h, double hmax, double err_norm, double *h_next)
{
    const double safe = err_norm > 1e-16 ? err_norm : 1e-16;
    double factor = c.rk45_safety * pow(safe, -0.2);
    factor = factor < 0.1 ? 0.1 : (factor > 5.0 ? 5.0 : factor);
    const int acc = err_norm <= 1.0;
    const double hmin = c.rk45_min_dt > 1e-16 ? c.rk45_min_dt : 1e-16;
    const double hhi = hmax > hmin ? hmax : hmin;
    double hn = acc ? h * factor : h * std::max(std::min(factor, 0.5), 0.1);
    hn = hn < hmin ? hmin : (hn > hhi ? hhi : hn);
    *h_next = hn;
    return acc;
}

// Kernel shapes chosen by the tuner, per (N, split storage): sites per lane of the
// kick and of the tensor launches (0: not tuned yet).
struct Shapes {
    int kick_e = 0, tensor_e = 0;   // 0: not tuned
    int march_pf = -1;              // -1: not tuned
};
std::mutex g_shape_mu;
std::map<std::pair<int, int>, Shapes> g_shapes;

Shapes tuned_shapes(int N, int split)
{
    std::lock_guard<std::mutex> lock(g_shape_mu);
    auto it = g_shapes.find({N, split});
    return it == g_shapes.end() ? Shapes() : it->second;
}

// Shapes of a class no run of this process has tuned: the tuner's choices on
// MI355X (16 steps per candidate, params.analytic.txt and params.numerical.txt
// presets).  N = 64: the march tensor kernel with early loads, one site per lane
// (0.017 ms per tensor step against 0.039 for the row kernel); N = 128: the
// march step with the j +- 1 rows from L2, tensor march with early loads (E = 1
// split, 2 FP64); N = 256: the march step from L2 in split storage and for the
// tabulated potential in FP64, through LDS for the hilltop in FP64, tensor march
// E = 4 in split storage (0.335 ms against 0.451 for the row kernel), E = 2 with
// early loads in FP64.  Results do not depend on the shapes.
Shapes default_shapes(int N, int split, int numerical)
{
    Shapes s;
    if (N == 64) {
        s.kick_e = 1;
        s.tensor_e = kTensorMarch | kTensorPf | 1;
    } else if (N == 128) {
        s.kick_e = 2;
        s.march_pf = 1;
        s.tensor_e = kTensorMarch | kTensorPf | (split ? 1 : 2);
    } else if (N == 256) {
        s.kick_e = 4;
        s.march_pf = split || numerical ? 1 : 0;
        s.tensor_e = split ? kTensorMarch | 4 : kTensorMarch | kTensorPf | 2;
    }
    return s;
}

// A class's tuned shapes, else its defaults; the parts the tuner left unset
// take the defaults too.
Shapes class_shapes(int N, int split, int numerical)
{
    const Shapes t = tuned_shapes(N, split), d = default_shapes(N, split, numerical);
    Shapes s;
    s.kick_e = t.kick_e ? t.kick_e : d.kick_e;
    s.tensor_e = t.tensor_e ? t.tensor_e : d.tensor_e;
    s.march_pf = t.march_pf >= 0 ? t.march_pf : d.march_pf;
    return s;
}

void apply_shapes(Fields &F, const Shapes &sh)
{
    F.kick_e = sh.kick_e;
    F.tensor_e = sh.tensor_e;
    F.march_pf = sh.march_pf > 0 ? 1 : 0;
}

void store_shapes(int N, int split, int kick_e, int tensor_e, int march_pf)
{
    std::lock_guard<std::mutex> lock(g_shape_mu);
    Shapes &v = g_shapes[{N, split}];
    if (kick_e > 0) v.kick_e = kick_e;
    if (tensor_e > 0) v.tensor_e = tensor_e;
    if (march_pf >= 0) v.march_pf = march_pf;
}

__global__ void k_loop_begin(DevBg *bg)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) bg->done = !(bg->a <= bg->af) ? 1 : 0;
}

}  // namespace

namespace ie {
void set_unbound_error(const std::string &msg) { g_create_error = msg; }
}  // namespace ie

extern "C" {

const char *ie_version(void) { return "inflation-easy-amd 0.1 (gfx950)"; }

int ie_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char *ie_last_error(const ie_run *run) { return run ? run->err.c_str() : g_create_error.c_str(); }

// Device bytes of a tabulated potential: x, V, V' as given, then the interval
// records and the two index levels (ie_tabindex.h), each part 256-byte aligned.
static size_t align_up(size_t b) { return (b + 255) & ~(size_t)255; }

static size_t table_bytes(const ie_config &c)
{
    if (!c.numerical_potential || c.tab_n < 2 || !c.tab_x) return 0;
    const int n = c.tab_n;
    long subs = 0;
    const int k1 = ie::tab_k1(c.tab_x, n, &subs);
    return align_up(3 * (size_t)n * 8) + align_up((size_t)n * sizeof(ie::TabRec)) +
           align_up((size_t)k1 * sizeof(ie::TabBucket)) + align_up((size_t)subs * sizeof(ie::TabSub));
}

size_t ie_footprint_bytes(const ie_config *c)
{
    if (!c || !is_pow2(c->N)) return 0;
    const size_t gs = (size_t)c->N * c->N * c->N;
    const size_t es = c->precision == IE_SPLIT ? 4 : 8;
    size_t b = 4 * gs * es;                        // f and fd ping-pong
    if (c->calculate_sigw) b += 18 * gs * 4;       // h ping-pong + hd
    b += gs * 8;                                   // scratch (downloads, deltaN)
    if (c->precision == IE_SPLIT) b += 2 * gs * 8; // FP64 state of the deltaN phase
    b += ie::fft_workspace_bytes(c->N);            // spectra
    b += 2048 * 8 * 8 + 3 * sizeof(DevBg) + table_bytes(*c);
    return b;
}

ie_status ie_create(const ie_config *cfg, ie_run **out)
{
    if (!cfg || !out) return fail(nullptr, IE_ERR_ARG, "null argument");
    *out = nullptr;
    const ie_config &c = *cfg;
    if (!is_pow2(c.N) || c.N < 8 || c.N > 1024)
        return fail(nullptr, IE_ERR_ARG, "N must be a power of two in [8, 1024]");
    if (c.precision != IE_FP64 && c.precision != IE_SPLIT) return fail(nullptr, IE_ERR_ARG, "bad precision");
    if (c.numerical_potential && (c.tab_n < 2 || !c.tab_x || !c.tab_v || !c.tab_dv))
        return fail(nullptr, IE_ERR_ARG, "Numerical potential table must contain at least two points.");
    if (!(c.L > 0.0) || !(c.dt > 0.0)) return fail(nullptr, IE_ERR_ARG, "L and dt must be positive");
    const double dx = c.L / (double)c.N;
    if (c.dt > dx / sqrt(3.0)) {
        char buf[160];
        snprintf(buf, sizeof(buf), "Time step too large. Courant condition violated: dt/dx = %f", c.dt / dx);
        return fail(nullptr, IE_ERR_COURANT, buf);
    }
    if (c.numerical_potential)
        for (int q = 1; q < c.tab_n; ++q)
            if (!(c.tab_x[q] < c.tab_x[q - 1]))
                return fail(nullptr, IE_ERR_ARG, "tabulated field grid must be strictly descending");

    ie_run *r = new ie_run();
    r->cfg = c;
    r->cfg.tab_x = r->cfg.tab_v = r->cfg.tab_dv = nullptr;
    if (c.numerical_potential) {
        r->htab_x.assign(c.tab_x, c.tab_x + c.tab_n);
        r->htab_v.assign(c.tab_v, c.tab_v + c.tab_n);
        r->htab_dv.assign(c.tab_dv, c.tab_dv + c.tab_n);
        r->cfg.tab_x = r->htab_x.data();
        r->cfg.tab_v = r->htab_v.data();
        r->cfg.tab_dv = r->htab_dv.data();
    }
    r->N = c.N;
    while ((1 << r->logN) < c.N) ++r->logN;
    r->gs = (size_t)c.N * c.N * c.N;
    r->dx = dx;
    ie_status st = IE_OK;
    auto bail = [&](ie_status s) {
        ie_destroy(r);
        return s;
    };
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) {
        delete r;
        return fail(nullptr, IE_ERR_HIP, "hipStreamCreate failed");
    }
    const size_t es = c.precision == IE_SPLIT ? sizeof(float) : sizeof(double);
    r->F.N = c.N;
    r->F.logN = r->logN;
    r->F.sigw = c.calculate_sigw ? 1 : 0;
    apply_shapes(r->F, class_shapes(c.N, c.precision == IE_SPLIT, c.numerical_potential ? 1 : 0));
    for (int p = 0; p < 2; ++p)
        if ((st = dev_alloc(r, &r->F.f[p], r->gs * es)) != IE_OK) return bail(st);
    for (int p = 0; p < 2; ++p)
        if ((st = dev_alloc(r, &r->F.fd[p], r->gs * es)) != IE_OK) return bail(st);
    if (r->F.sigw) {
        for (int q = 0; q < 6; ++q) {
            for (int p = 0; p < 2; ++p)
                if ((st = dev_alloc(r, (void **)&r->F.h[p][q], r->gs * 4)) != IE_OK) return bail(st);
            if ((st = dev_alloc(r, (void **)&r->F.hd[q], r->gs * 4)) != IE_OK) return bail(st);
        }
    }
    if ((st = dev_alloc(r, (void **)&r->scratch, r->gs * 8)) != IE_OK) return bail(st);
    // split runs: the FP64 state of the deltaN phase, allocated here so that no
    // later phase allocates (ie_footprint_bytes counts it)
    if (c.precision == IE_SPLIT) {
        if ((st = dev_alloc(r, (void **)&r->dn_f, r->gs * 8)) != IE_OK) return bail(st);
        if ((st = dev_alloc(r, (void **)&r->dn_fd, r->gs * 8)) != IE_OK) return bail(st);
    }
    if ((st = dev_alloc(r, (void **)&r->partials, 2048 * 8 * 8)) != IE_OK) return bail(st);
    if ((st = dev_alloc(r, (void **)&r->bg, sizeof(DevBg))) != IE_OK) return bail(st);
    if ((st = dev_alloc(r, (void **)&r->F.rec, 2 * sizeof(DevBg))) != IE_OK) return bail(st);
    if ((st = dev_alloc(r, (void **)&r->work, 4096 * 8)) != IE_OK) return bail(st);
    if ((st = dev_alloc(r, (void **)&r->nmean, 256)) != IE_OK) return bail(st);
    if (hipHostMalloc((void **)&r->bg_host, sizeof(DevBg), hipHostMallocDefault) != hipSuccess)
        return bail(fail(nullptr, IE_ERR_OOM, "hipHostMalloc failed"));
    r->F.dn = r->scratch;
    for (int p = 0; p < 2; ++p) {
        if (hipMemsetAsync(r->F.f[p], 0, r->gs * es, r->stream) != hipSuccess) return bail(IE_ERR_HIP);
    }
    for (int p = 0; p < 2; ++p)
        if (hipMemsetAsync(r->F.fd[p], 0, r->gs * es, r->stream) != hipSuccess) return bail(IE_ERR_HIP);
    if (r->F.sigw)
        for (int q = 0; q < 6; ++q) {
            for (int p = 0; p < 2; ++p)
                if (hipMemsetAsync(r->F.h[p][q], 0, r->gs * 4, r->stream) != hipSuccess) return bail(IE_ERR_HIP);
            if (hipMemsetAsync(r->F.hd[q], 0, r->gs * 4, r->stream) != hipSuccess) return bail(IE_ERR_HIP);
        }
    r->P.numerical = c.numerical_potential ? 1 : 0;
    r->P.V0 = c.V0;
    r->P.ns = c.ns;
    r->P.B2 = c.rescale_B * c.rescale_B;
    r->P.invB2 = 1.0 / r->P.B2;
    r->P.hV = c.V0 / r->P.B2;
    r->P.hV4 = c.V0 * ((1.0 - c.ns) / 4.0) / r->P.B2;
    r->P.hV2 = c.V0 * ((1.0 - c.ns) / 2.0) / r->P.B2;
    r->P.n = 0;
    if (c.numerical_potential) {
        const size_t n = (size_t)c.tab_n;
        // x, V, V' as given; then the interval records and the O(1) index of
        // the lattice kernels
        const ie::TabIndexHost H = ie::tab_index_build(c.tab_x, c.tab_v, c.tab_dv, c.tab_n, r->P.invB2);
        const size_t o_rec = align_up(3 * n * 8), o_l1 = o_rec + align_up(H.rec.size() * sizeof(ie::TabRec));
        const size_t o_l2 = o_l1 + align_up(H.l1.size() * sizeof(ie::TabBucket));
        const size_t total = o_l2 + align_up(H.l2.size() * sizeof(ie::TabSub));
        if (total > table_bytes(c)) return bail(fail(nullptr, IE_ERR_ARG, "table index larger than its footprint"));
        std::vector<char> t(total, 0);
        memcpy(t.data(), c.tab_x, n * 8);
        memcpy(t.data() + n * 8, c.tab_v, n * 8);
        memcpy(t.data() + 2 * n * 8, c.tab_dv, n * 8);
        memcpy(t.data() + o_rec, H.rec.data(), H.rec.size() * sizeof(ie::TabRec));
        memcpy(t.data() + o_l1, H.l1.data(), H.l1.size() * sizeof(ie::TabBucket));
        memcpy(t.data() + o_l2, H.l2.data(), H.l2.size() * sizeof(ie::TabSub));
        if ((st = dev_alloc(r, (void **)&r->table, total)) != IE_OK) return bail(st);
        if (hipMemcpy(r->table, t.data(), total, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(nullptr, IE_ERR_HIP, "table upload failed"));
        char *base = (char *)r->table;
        r->P.n = c.tab_n;
        r->P.x = r->table;
        r->P.v = r->table + n;
        r->P.dv = r->table + 2 * n;
        r->P.T = ie::tab_view(H, (const ie::TabRec *)(base + o_rec), (const ie::TabBucket *)(base + o_l1),
                              (const ie::TabSub *)(base + o_l2), r->P.x, c.tab_n);
    }
    DevBg &b = *r->bg_host;
    memset(&b, 0, sizeof(b));
    b.a = 1.0;
    b.astep = 1.0;
    b.dt = c.dt;
    b.af = c.af;
    b.rescale_s = c.rescale_s;
    b.dx = dx;
    // fused loops of the hilltop sum f and f^2 for the Friedmann step (V is
    // quadratic in f); the close turns them into sum V and sum V'
    b.hill_sums = c.numerical_potential ? 0 : (c.precision == IE_SPLIT ? 2 : 1);
    b.hV = r->P.hV;
    b.hV4 = r->P.hV4;
    b.hV2 = r->P.hV2;
    if (push_bg(r) != IE_OK) return bail(IE_ERR_HIP);
    r->fft = fft_create(c.N);
    if (!r->fft) return bail(fail(nullptr, IE_ERR_OOM, "FFT workspace allocation failed"));
    *out = r;
    return IE_OK;
}

ie_status ie_get_config(const ie_run *r, ie_config *out)
{
    if (!r || !out) return IE_ERR_ARG;
    *out = r->cfg;
    return IE_OK;
}

ie_status ie_destroy(ie_run *r)
{
    if (!r) return IE_OK;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    if (r->fft) fft_destroy(r->fft);
    rk_free(r->rk4);
    rk_free(r->rk45);
    for (void *p : r->allocs) (void)hipFree(p);
    if (r->hist) (void)hipFree(r->hist);
    if (r->bg_host) (void)hipHostFree(r->bg_host);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
    return IE_OK;
}

ie_status ie_upload_fields(ie_run *r, const double *f, const double *fd)
{
    if (!r || !f || !fd) return fail(r, IE_ERR_ARG, "null argument");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    r->bg_host->parity = 0;
    if ((st = push_bg(r)) != IE_OK) return st;
    const size_t bytes = r->gs * 8;
    if (r->cfg.precision == IE_FP64) {
        IE_HIP(r, hipMemcpyAsync(r->F.f[0], f, bytes, hipMemcpyHostToDevice, r->stream));
        IE_HIP(r, hipMemcpyAsync(r->F.fd[0], fd, bytes, hipMemcpyHostToDevice, r->stream));
        return sync(r);
    }
    // split: stage both fields in FP64 (scratch and the FP64 state buffer of the
    // deltaN phase, unused before it), compute the homogeneous mode on the device,
    // store fluctuations
    IE_HIP(r, hipMemcpyAsync(r->scratch, f, bytes, hipMemcpyHostToDevice, r->stream));
    IE_HIP(r, hipMemcpyAsync(r->dn_fd, fd, bytes, hipMemcpyHostToDevice, r->stream));
    IE_HIP(r, launch_split_upload(r->scratch, r->dn_fd, r->F, r->bg, r->partials, r->stream));
    return sync(r);
}

ie_status ie_download_field(ie_run *r, int which, double *out)
{
    if (!r || !out) return fail(r, IE_ERR_ARG, "null argument");
    if (which == IE_FIELD_DELTAN) {
        IE_HIP(r, hipMemcpyAsync(out, r->F.dn, r->gs * 8, hipMemcpyDeviceToHost, r->stream));
        return sync(r);
    }
    if (which != IE_FIELD_F && which != IE_FIELD_FD) return fail(r, IE_ERR_ARG, "bad field");
    if (r->in_deltaN) {
        ie_status st = pull_bg(r);
        if (st != IE_OK) return st;
        const int split = r->cfg.precision == IE_SPLIT;
        const double *src = which == IE_FIELD_F ? (split ? r->dn_f : (const double *)r->F.f[r->bg_host->parity])
                                                : (split ? r->dn_fd : (const double *)r->F.fd[r->bg_host->parity]);
        IE_HIP(r, hipMemcpyAsync(out, src, r->gs * 8, hipMemcpyDeviceToHost, r->stream));
        return sync(r);
    }
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    const int par = r->bg_host->parity;
    if (r->cfg.precision == IE_FP64) {
        const void *src = which == IE_FIELD_F ? r->F.f[par] : r->F.fd[par];
        IE_HIP(r, hipMemcpyAsync(out, src, r->gs * 8, hipMemcpyDeviceToHost, r->stream));
        return sync(r);
    }
    // split: phibar + the fluctuation in FP64, staged in the deltaN phase's FP64
    // buffer (unused before that phase)
    IE_HIP(r, launch_reconstitute(r->F, 1, which, r->bg, r->dn_f, r->stream));
    IE_HIP(r, hipMemcpyAsync(out, r->dn_f, r->gs * 8, hipMemcpyDeviceToHost, r->stream));
    return sync(r);
}

static ie_status dn_state(ie_run *r, double **f, double **fd);

ie_status ie_download_plane(ie_run *r, int which, int i, double *out)
{
    if (!r || !out) return fail(r, IE_ERR_ARG, "null argument");
    if (i < 0 || i >= r->N) return fail(r, IE_ERR_ARG, "plane index out of range");
    const size_t n = (size_t)r->N * r->N, off = (size_t)i * n;
    if (which == IE_FIELD_DELTAN) {
        IE_HIP(r, hipMemcpyAsync(out, r->F.dn + off, n * 8, hipMemcpyDeviceToHost, r->stream));
        return sync(r);
    }
    if (which != IE_FIELD_F && which != IE_FIELD_FD) return fail(r, IE_ERR_ARG, "bad field");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    const DevBg &b = *r->bg_host;
    if (r->in_deltaN || r->cfg.precision == IE_FP64) {
        double *f, *fd;
        if (r->in_deltaN) {
            if ((st = dn_state(r, &f, &fd)) != IE_OK) return st;
        } else {
            f = (double *)r->F.f[b.parity];
            fd = (double *)r->F.fd[b.parity];
        }
        IE_HIP(r, hipMemcpyAsync(out, (which == IE_FIELD_F ? f : fd) + off, n * 8, hipMemcpyDeviceToHost, r->stream));
        return sync(r);
    }
    // split storage: the plane's FP32 fluctuations plus the FP64 mean, as
    // ie_download_field reconstitutes them
    std::vector<float> v(n);
    const float *src = (const float *)(which == IE_FIELD_F ? r->F.f[b.parity] : r->F.fd[b.parity]);
    IE_HIP(r, hipMemcpyAsync(v.data(), src + off, n * 4, hipMemcpyDeviceToHost, r->stream));
    if ((st = sync(r)) != IE_OK) return st;
    const double base = which == IE_FIELD_F ? b.phibar : b.pibar;
    for (size_t q = 0; q < n; ++q) out[q] = base + (double)v[q];
    return IE_OK;
}

ie_status ie_upload_field(ie_run *r, int which, const double *src)
{
    if (!r || !src) return fail(r, IE_ERR_ARG, "null argument");
    double *dst = nullptr;
    if (which == IE_FIELD_DELTAN) {
        dst = r->F.dn;
    } else if (which == IE_FIELD_F || which == IE_FIELD_FD) {
        if (!r->in_deltaN) return fail(r, IE_ERR_STATE, "inflation phase: use ie_upload_fields");
        double *f, *fd;
        ie_status st = dn_state(r, &f, &fd);
        if (st != IE_OK) return st;
        dst = which == IE_FIELD_F ? f : fd;
    } else {
        return fail(r, IE_ERR_ARG, "bad field");
    }
    IE_HIP(r, hipMemcpyAsync(dst, src, r->gs * 8, hipMemcpyHostToDevice, r->stream));
    return sync(r);
}

ie_status ie_upload_tensor(ie_run *r, int comp, const float *h)
{
    if (!r || !h || comp < 0 || comp > 11) return fail(r, IE_ERR_ARG, "bad argument");
    if (!r->F.sigw) return fail(r, IE_ERR_STATE, "run has no tensors (calculate_sigw = 0)");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    float *dst = comp < 6 ? r->F.h[r->bg_host->parity][comp] : r->F.hd[comp - 6];
    IE_HIP(r, hipMemcpyAsync(dst, h, r->gs * 4, hipMemcpyHostToDevice, r->stream));
    return sync(r);
}

ie_status ie_download_tensor(ie_run *r, int comp, float *out)
{
    if (!r || !out || comp < 0 || comp > 11) return fail(r, IE_ERR_ARG, "bad argument");
    if (!r->F.sigw) return fail(r, IE_ERR_STATE, "run has no tensors (calculate_sigw = 0)");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    const float *src = comp < 6 ? r->F.h[r->bg_host->parity][comp] : r->F.hd[comp - 6];
    IE_HIP(r, hipMemcpyAsync(out, src, r->gs * 4, hipMemcpyDeviceToHost, r->stream));
    return sync(r);
}

// The host copy of the background as an ie_background (no read-back).
static void background_out(const ie_run *r, ie_background *out)
{
    const DevBg &b = *r->bg_host;
    out->t = b.t;
    out->a = b.a;
    out->ad = b.ad;
    out->ad2 = b.ad2;
    out->astep = b.astep;
    out->Ne = b.Ne;
    out->phiref = b.phiref;
    out->hubble_init = b.hubble_init;
    out->dt_last = b.d;
    out->steps = b.steps;
    out->done = b.done;
    out->error = b.error ? IE_ERR_TABLE_RANGE : IE_OK;
}

ie_status ie_get_background(ie_run *r, ie_background *out)
{
    if (!r || !out) return fail(r, IE_ERR_ARG, "null argument");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    background_out(r, out);
    return IE_OK;
}

ie_status ie_set_background(ie_run *r, const ie_background *in)
{
    if (!r || !in) return fail(r, IE_ERR_ARG, "null argument");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    DevBg &b = *r->bg_host;
    b.t = in->t;
    b.a = in->a;
    b.ad = in->ad;
    b.ad2 = in->ad2;
    b.astep = in->astep;
    b.Ne = in->Ne;
    b.phiref = in->phiref;
    b.hubble_init = in->hubble_init;
    b.steps = in->steps;
    b.done = in->done;
    b.error = 0;
    return push_bg(r);
}

static ie_status sums(ie_run *r)
{
    IE_HIP(r, launch_field_sums(r->F, r->cfg.precision == IE_SPLIT, r->P, r->bg, r->partials, r->stream));
    return check_kernel_error(r);
}

ie_status ie_gradient_energy(ie_run *r, double *out)
{
    if (!r || !out) return fail(r, IE_ERR_ARG, "null argument");
    ie_status st = sums(r);
    if (st != IE_OK) return st;
    const DevBg &b = *r->bg_host;
    const double norm = (1.0 / (b.a * r->dx)) * (1.0 / (b.a * r->dx));
    *out = 0.5 * b.red[4] * norm / (double)r->gs;
    return IE_OK;
}

ie_status ie_kin_energy(ie_run *r, double *out)
{
    if (!r || !out) return fail(r, IE_ERR_ARG, "null argument");
    ie_status st = sums(r);
    if (st != IE_OK) return st;
    const DevBg &b = *r->bg_host;
    double s2 = b.red[3];
    if (r->cfg.precision == IE_SPLIT) {
        // sum (pibar + dpi)^2 = N^3 pibar^2 + 2 pibar sum dpi + sum dpi^2
        s2 = (double)r->gs * b.pibar * b.pibar + 2.0 * b.pibar * b.red[1] + b.red[3];
    }
    s2 /= (double)r->gs;
    *out = 0.5 * pow(b.a, 2.0 * r->cfg.rescale_s - 2.0) * s2;
    return IE_OK;
}

ie_status ie_potential_energy(ie_run *r, double *out)
{
    if (!r || !out) return fail(r, IE_ERR_ARG, "null argument");
    ie_status st = sums(r);
    if (st != IE_OK) return st;
    *out = r->bg_host->red[5] / (double)r->gs;
    return IE_OK;
}

ie_status ie_field_sums(ie_run *r, double out[3])
{
    if (!r || !out) return fail(r, IE_ERR_ARG, "null argument");
    if (r->in_deltaN) {
        // FP64 state of the deltaN and post-inflation phases
        double *f, *fd;
        ie_status st = dn_state(r, &f, &fd);
        if (st != IE_OK) return st;
        IE_HIP(r, post_sums3(f, fd, r->gs, r->work, r->work + 4000, r->stream));
        IE_HIP(r, hipMemcpyAsync(out, r->work + 4000, 3 * 8, hipMemcpyDeviceToHost, r->stream));
        return sync(r);
    }
    ie_status st = sums(r);
    if (st != IE_OK) return st;
    const DevBg &b = *r->bg_host;
    if (r->cfg.precision == IE_SPLIT) {
        const double n = (double)r->gs;
        out[0] = n * b.phibar + b.red[0];
        out[1] = n * b.pibar + b.red[1];
        out[2] = n * b.phibar * b.phibar + 2.0 * b.phibar * b.red[0] + b.red[2];
    } else {
        out[0] = b.red[0];
        out[1] = b.red[1];
        out[2] = b.red[2];
    }
    return IE_OK;
}

// the inflation steps need the inflation state
static ie_status inflation_phase(ie_run *r)
{
    if (r->in_post) return fail(r, IE_ERR_STATE, "run is in its post-inflation phase");
    if (r->in_deltaN) return fail(r, IE_ERR_STATE, "run is in its deltaN phase");
    return IE_OK;
}

ie_status ie_evolve_derivs(ie_run *r, double d)
{
    if (!r || !(d != 0.0)) return fail(r, IE_ERR_ARG, "bad step");
    if (inflation_phase(r) != IE_OK) return IE_ERR_STATE;
    const int split = r->cfg.precision == IE_SPLIT;
    IE_HIP(r, launch_friedmann_reduce(r->F, split, r->P, r->bg, r->partials, d, 0, r->stream));
    IE_HIP(r, launch_kick_drift(r->F, split, r->P, r->bg, KICK, d, 0, r->stream));
    return check_kernel_error(r);
}

static ie_status post_drift_call(ie_run *r, double d);

ie_status ie_evolve_fields(ie_run *r, double d)
{
    if (!r) return fail(r, IE_ERR_ARG, "null argument");
    if (d == 0.0) return IE_OK;
    if (r->in_post) return post_drift_call(r, d);
    const int split = r->cfg.precision == IE_SPLIT;
    IE_HIP(r, launch_kick_drift(r->F, split, r->P, r->bg, DRIFT, d, 0, r->stream));
    return sync(r);
}

ie_status ie_leapfrog_steps(ie_run *r, int max_steps, ie_background *out)
{
    if (!r || max_steps < 0) return fail(r, IE_ERR_ARG, "bad argument");
    if (inflation_phase(r) != IE_OK) return IE_ERR_STATE;
    const int split = r->cfg.precision == IE_SPLIT;
    // first long loop of a kind: tune the kick / tensor shapes on its first steps
    const Shapes sh = tuned_shapes(r->N, split);
    const bool untuned = march_ok(r->N) ? (sh.march_pf < 0 || (r->F.sigw && sh.tensor_e == 0))
                                        : (sh.kick_e == 0 || (r->F.sigw && sh.tensor_e == 0));
    if (max_steps >= 256 && row_elems(r->N) >= 2 && untuned && !r->pinned) {
        ie_tuning tun;
        ie_status st = ie_tune(r, 8, &tun);
        if (st != IE_OK) return st;
        max_steps -= (int)std::min<int64_t>(tun.steps, max_steps);
    }
    // the steps go out in chunks with the loop test read back in between, so a
    // call with a large max_steps returns soon after a > af (the step kernels of a
    // finished loop return at once, but each launch still costs its dispatch);
    // a loop in chunks gives the same bits as one loop
    constexpr int kChunk = 1024;
    bool fresh = false;   // bg_host read back after the last launch
    for (int done = 0; done < max_steps;) {
        const int n_chunk = std::min(kChunk, max_steps - done);
        k_loop_begin<<<1, 64, 0, r->stream>>>(r->bg);
        IE_HIP(r, hipGetLastError());
        if (march_ok(r->N)) {
            // fused steps: the march launch of a step closes the previous one and
            // computes the next step's sums
            IE_HIP(r, launch_march_begin(r->F, split, r->P, r->bg, r->partials, r->stream));
            for (int n = 0; n < n_chunk; ++n)
                IE_HIP(r, launch_fused_step(r->F, split, r->P, r->bg, r->partials, n, r->stream));
            IE_HIP(r, launch_loop_end(r->F, split, r->bg, r->partials, n_chunk, r->stream));
        } else {
            for (int n = 0; n < n_chunk; ++n) {
                IE_HIP(r, launch_friedmann_reduce(r->F, split, r->P, r->bg, r->partials, 0.0, 1, r->stream));
                IE_HIP(r, launch_kick_drift(r->F, split, r->P, r->bg, KICKDRIFT, 0.0, 1, r->stream));
            }
        }
        done += n_chunk;
        ie_status st = check_kernel_error(r);
        if (st != IE_OK) return st;
        fresh = true;
        if (r->bg_host->done) break;
    }
    if (max_steps == 0) {
        k_loop_begin<<<1, 64, 0, r->stream>>>(r->bg);
        IE_HIP(r, hipGetLastError());
        fresh = false;
    }
    if (!fresh) {
        ie_status st = check_kernel_error(r);
        if (st != IE_OK) return st;
    }
    if (out) background_out(r, out);
    return IE_OK;
}

ie_status ie_profile_steps(ie_run *r, int nsteps, ie_profile *out)
{
    if (!r || nsteps < 1 || !out) return fail(r, IE_ERR_ARG, "bad argument");
    if (inflation_phase(r) != IE_OK) return IE_ERR_STATE;
    const int split = r->cfg.precision == IE_SPLIT;
    // events bracket each launch of a step: reduce | finish | scalar | tensor | finish
    const int nev = 6;
    std::vector<hipEvent_t> ev((size_t)nsteps * nev);
    for (auto &e : ev) IE_HIP(r, hipEventCreate(&e));
    k_loop_begin<<<1, 64, 0, r->stream>>>(r->bg);
    const bool fused = march_ok(r->N);
    if (fused) IE_HIP(r, launch_march_begin(r->F, split, r->P, r->bg, r->partials, r->stream));
    for (int n = 0; n < nsteps; ++n) {
        hipEvent_t *e = &ev[(size_t)n * nev];
        if (fused) {
            // the march kernel is the previous step's close, the kick, the drift
            // and the next step's sums; e[1], e[2] stay unrecorded (no reduce or
            // Friedmann launch), e[5] follows e[4] (no close launch)
            IE_HIP(r, hipEventRecord(e[0], r->stream));
            IE_HIP(r, launch_march(r->F, split, r->P, r->bg, r->partials, n, r->stream));
            IE_HIP(r, hipEventRecord(e[3], r->stream));
            if (r->F.sigw)
                IE_HIP(r, launch_tensor_step(r->F, split, r->bg, KICKDRIFT, 0.0, 1, r->stream, step_record(n)));
            IE_HIP(r, hipEventRecord(e[4], r->stream));
            IE_HIP(r, hipEventRecord(e[5], r->stream));
            continue;
        }
        IE_HIP(r, hipEventRecord(e[0], r->stream));
        IE_HIP(r, launch_reduce(r->F, split, r->P, r->bg, r->partials, 1, r->stream));
        IE_HIP(r, hipEventRecord(e[1], r->stream));
        IE_HIP(r, launch_friedmann_finish(r->F, r->bg, r->partials, 0.0, 1, r->stream));
        IE_HIP(r, hipEventRecord(e[2], r->stream));
        IE_HIP(r, launch_scalar_step(r->F, split, r->P, r->bg, KICKDRIFT, 0.0, 1, r->stream));
        IE_HIP(r, hipEventRecord(e[3], r->stream));
        if (r->F.sigw) IE_HIP(r, launch_tensor_step(r->F, split, r->bg, KICKDRIFT, 0.0, 1, r->stream));
        IE_HIP(r, hipEventRecord(e[4], r->stream));
        IE_HIP(r, launch_step_finish(r->bg, split, KICKDRIFT, 0.0, 1, r->stream));
        IE_HIP(r, hipEventRecord(e[5], r->stream));
    }
    if (fused) IE_HIP(r, launch_loop_end(r->F, split, r->bg, r->partials, nsteps, r->stream));
    ie_status st = sync(r);
    if (st != IE_OK) return st;
    double red = 0, kd = 0, ten = 0, step = 0;
    for (int n = 0; n < nsteps; ++n) {
        hipEvent_t *e = &ev[(size_t)n * nev];
        float ms = 0.f;
        if (!fused) {
            IE_HIP(r, hipEventElapsedTime(&ms, e[0], e[1]));
            red += ms;
        }
        IE_HIP(r, hipEventElapsedTime(&ms, fused ? e[0] : e[2], e[3]));
        kd += ms;
        if (r->F.sigw) {
            IE_HIP(r, hipEventElapsedTime(&ms, e[3], e[4]));
            ten += ms;
        }
        IE_HIP(r, hipEventElapsedTime(&ms, e[0], e[5]));
        step += ms;
    }
    // the loop alone, events around it only
    k_loop_begin<<<1, 64, 0, r->stream>>>(r->bg);
    if (fused) IE_HIP(r, launch_march_begin(r->F, split, r->P, r->bg, r->partials, r->stream));
    IE_HIP(r, hipEventRecord(ev[0], r->stream));
    for (int n = 0; n < nsteps; ++n) {
        if (fused) {
            IE_HIP(r, launch_march(r->F, split, r->P, r->bg, r->partials, n, r->stream));
            if (r->F.sigw)
                IE_HIP(r, launch_tensor_step(r->F, split, r->bg, KICKDRIFT, 0.0, 1, r->stream, step_record(n)));
            continue;
        }
        IE_HIP(r, launch_reduce(r->F, split, r->P, r->bg, r->partials, 1, r->stream));
        IE_HIP(r, launch_friedmann_finish(r->F, r->bg, r->partials, 0.0, 1, r->stream));
        IE_HIP(r, launch_scalar_step(r->F, split, r->P, r->bg, KICKDRIFT, 0.0, 1, r->stream));
        if (r->F.sigw) IE_HIP(r, launch_tensor_step(r->F, split, r->bg, KICKDRIFT, 0.0, 1, r->stream));
        IE_HIP(r, launch_step_finish(r->bg, split, KICKDRIFT, 0.0, 1, r->stream));
    }
    IE_HIP(r, hipEventRecord(ev[1], r->stream));
    if (fused) IE_HIP(r, launch_loop_end(r->F, split, r->bg, r->partials, nsteps, r->stream));
    if ((st = sync(r)) != IE_OK) return st;
    float loop = 0.f;
    IE_HIP(r, hipEventElapsedTime(&loop, ev[0], ev[1]));
    for (auto &e : ev) (void)hipEventDestroy(e);
    if ((st = check_kernel_error(r)) != IE_OK) return st;
    out->loop_ms = loop / nsteps;
    out->reduce_ms = red / nsteps;
    out->kick_drift_ms = kd / nsteps;
    out->tensor_ms = ten / nsteps;
    out->step_ms = step / nsteps;
    out->steps = r->bg_host->steps;
    return IE_OK;
}

ie_status ie_tune(ie_run *r, int K, ie_tuning *out)
{
    if (!r || K < 1 || !out) return fail(r, IE_ERR_ARG, "bad argument");
    if (inflation_phase(r) != IE_OK) return IE_ERR_STATE;
    memset(out, 0, sizeof(*out));
    const int split = r->cfg.precision == IE_SPLIT;
    const int def = row_elems(r->N);
    const bool fused = march_ok(r->N);
    out->kick_elems = def;
    for (int e = 1; e <= 4 && def > 0; e *= 2)
        if (e <= def) out->elems[out->ncand++] = e;
    if (r->F.sigw && def > 0) {
        // row kernel shapes, then march kernel shapes (E = 4 with FP32 rows only)
        // without and with the early loads of the next plane
        for (int e = 1; e <= 4; e *= 2)
            if (e <= def) out->tensor_code[out->ntensor++] = e;
        for (int e = 1; e <= 4; e *= 2)
            if (e <= def && (e < 4 || split)) out->tensor_code[out->ntensor++] = kTensorMarch | e;
        for (int e = 1; e <= 2; e *= 2)
            if (e <= def) out->tensor_code[out->ntensor++] = kTensorMarch | kTensorPf | e;
    }
    // fused steps have one kick shape and two kernel variants; unfused steps have
    // kick shapes
    if (fused) out->nmarch = 2;
    const int nk = fused ? out->nmarch : out->ncand, nt = out->ntensor;
    const int rounds = std::max(nk, nt);
    if (rounds <= 1) {
        r->F.kick_e = def;
        r->F.tensor_e = nt == 1 ? out->tensor_code[0] : 0;
        r->F.march_pf = 0;
        out->tensor_code_chosen = r->F.tensor_e;
        store_shapes(r->N, split, def, r->F.sigw ? std::max(r->F.tensor_e, 1) : 0, 0);
        return IE_OK;
    }
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    const int64_t steps0 = r->bg_host->steps;
    std::vector<hipEvent_t> ev((size_t)rounds * K * 4);
    for (auto &e : ev) IE_HIP(r, hipEventCreate(&e));
    k_loop_begin<<<1, 64, 0, r->stream>>>(r->bg);
    if (fused) IE_HIP(r, launch_march_begin(r->F, split, r->P, r->bg, r->partials, r->stream));
    // candidates in turn, one step each, so that every candidate samples the
    // same stretch of the run (the cost of a step drifts in the first steps
    // after the initial conditions)
    for (int q = 0; q < K; ++q) {
        for (int c = 0; c < rounds; ++c) {
            r->F.kick_e = fused ? def : out->elems[c % nk];
            r->F.march_pf = fused ? c % nk : 0;
            r->F.tensor_e = nt ? out->tensor_code[c % nt] : 0;
            const long long n = (long long)q * rounds + c;
            hipEvent_t *e = &ev[(size_t)n * 4];
            if (!fused) IE_HIP(r, launch_friedmann_reduce(r->F, split, r->P, r->bg, r->partials, 0.0, 1, r->stream));
            IE_HIP(r, hipEventRecord(e[0], r->stream));
            if (fused) IE_HIP(r, launch_march(r->F, split, r->P, r->bg, r->partials, n, r->stream));
            else IE_HIP(r, launch_scalar_step(r->F, split, r->P, r->bg, KICKDRIFT, 0.0, 1, r->stream));
            IE_HIP(r, hipEventRecord(e[1], r->stream));
            if (r->F.sigw)
                IE_HIP(r, launch_tensor_step(r->F, split, r->bg, KICKDRIFT, 0.0, 1, r->stream,
                                             fused ? step_record(n) : 0));
            IE_HIP(r, hipEventRecord(e[2], r->stream));
            if (!fused) IE_HIP(r, launch_step_finish(r->bg, split, KICKDRIFT, 0.0, 1, r->stream));
            IE_HIP(r, hipEventRecord(e[3], r->stream));
        }
    }
    if (fused) IE_HIP(r, launch_loop_end(r->F, split, r->bg, r->partials, (long long)rounds * K, r->stream));
    if ((st = sync(r)) != IE_OK) return st;
    std::vector<double> ksum(3, 0.0), tsum(8, 0.0);
    std::vector<int> kcnt(3, 0), tcnt(8, 0);
    // each candidate's first launch is not counted (its kernel's first use)
    for (int c = 0; c < rounds; ++c)
        for (int q = K > 1 ? 1 : 0; q < K; ++q) {
            hipEvent_t *e = &ev[((size_t)q * rounds + c) * 4];
            float ms = 0.f;
            IE_HIP(r, hipEventElapsedTime(&ms, e[0], e[1]));
            ksum[c % nk] += ms;
            kcnt[c % nk] += 1;
            if (nt) {
                IE_HIP(r, hipEventElapsedTime(&ms, e[1], e[2]));
                tsum[c % nt] += ms;
                tcnt[c % nt] += 1;
            }
        }
    for (auto &e : ev) (void)hipEventDestroy(e);
    if ((st = check_kernel_error(r)) != IE_OK) return st;
    int bk = 0, bt = 0, bm = 0;
    if (fused) {
        for (int c = 0; c < out->nmarch; ++c) {
            out->march_ms[c] = kcnt[c] ? ksum[c] / kcnt[c] : 0.0;
            if (out->march_ms[c] < out->march_ms[bm]) bm = c;
        }
    } else {
        for (int c = 0; c < out->ncand; ++c) {
            out->kick_ms[c] = kcnt[c] ? ksum[c] / kcnt[c] : 0.0;
            if (out->kick_ms[c] < out->kick_ms[bk]) bk = c;
        }
    }
    for (int c = 0; c < nt; ++c) {
        out->tensor_ms[c] = tcnt[c] ? tsum[c] / tcnt[c] : 0.0;
        if (out->tensor_ms[c] < out->tensor_ms[bt]) bt = c;
    }
    out->steps = r->bg_host->steps - steps0;
    out->kick_elems = fused ? def : out->elems[bk];
    out->tensor_code_chosen = nt ? out->tensor_code[bt] : 0;
    out->march_variant = fused ? bm : 0;
    r->F.kick_e = out->kick_elems;
    r->F.tensor_e = out->tensor_code_chosen;
    r->F.march_pf = out->march_variant;
    // timings of a loop that reached af part way are not kept
    if (out->steps == (int64_t)rounds * K)
        store_shapes(r->N, split, out->kick_elems, out->tensor_code_chosen, fused ? bm : -1);
    return IE_OK;
}

ie_status ie_set_shapes(ie_run *r, int kick_elems, int tensor_code, int march_variant)
{
    if (!r) return fail(r, IE_ERR_ARG, "null run");
    const int split = r->cfg.precision == IE_SPLIT;
    const int def = row_elems(r->N);
    const auto elems_ok = [&](int e) { return e == 0 || ((e == 1 || e == 2 || e == 4) && e <= def); };
    if (!elems_ok(kick_elems)) return fail(r, IE_ERR_ARG, "kick_elems must be 0 or 1, 2, 4 within a row");
    const int te = tensor_code & 7, tflags = tensor_code & ~7;
    bool tensor_ok = tensor_code >= 0 && elems_ok(te);
    if (tflags == kTensorMarch) tensor_ok = tensor_ok && te > 0 && (te < 4 || split);
    else if (tflags == (kTensorMarch | kTensorPf)) tensor_ok = tensor_ok && te > 0 && te < 4;
    else tensor_ok = tensor_ok && tflags == 0;
    if (!tensor_ok) return fail(r, IE_ERR_ARG, "tensor_code is not one of the tuner's codes for this run");
    if (march_variant != 0 && march_variant != 1) return fail(r, IE_ERR_ARG, "march_variant must be 0 or 1");
    r->F.kick_e = kick_elems;
    r->F.tensor_e = tensor_code;
    r->F.march_pf = march_variant;
    r->pinned = 1;
    return IE_OK;
}

int ie_spectrum_numbins(int N) { return (int)(1.73205 * (N / 2)) + 1; }

ie_status ie_modes(int N, double L, double B, double *out)
{
    if (!is_pow2(N) || !out) return fail(nullptr, IE_ERR_ARG, "bad argument");
    const int nb = (int)(sqrt(3.0) * (N / 2)) + 1;
    // the result depends on (N, L, B) alone: computed once and kept (a pass over
    // the N^2 (N/2 + 1) modes takes about 0.1 s at N = 256)
    static std::mutex mu;
    static std::map<std::tuple<int, double, double>, std::vector<double>> cache;
    std::lock_guard<std::mutex> lock(mu);
    std::vector<double> &res = cache[std::make_tuple(N, L, B)];
    if (res.empty()) {
        // get_modes (src/output.cpp:674-738): closed form per shell, host side, in
        // the reference's mode order; sin(p pi / N) from a table of the same values
        const double pi = 2.0 * asin(1.0);
        std::vector<double> sn(N + 1);
        for (int p = -N / 2; p <= N / 2; ++p) sn[p + N / 2] = sin((double)p * pi / N);
        std::vector<int> np(nb, 0);
        std::vector<double> pp(nb, 0.0);
        auto shell = [&](int px, int py, int pz, int w) {
            const double pm = sqrt((double)px * px + (double)py * py + (double)pz * pz);
            const double sx = sn[px + N / 2], sy = sn[py + N / 2], sz = sn[pz + N / 2];
            const double ph = sqrt(4.0 * (((double)N / L) * ((double)N / L)) * (sx * sx + sy * sy + sz * sz));
            np[(int)pm] += w;
            pp[(int)pm] += w == 2 ? 2.0 * ph : ph;
        };
        for (int i = 0; i < N; ++i) {
            const int px = i <= N / 2 ? i : i - N;
            for (int j = 0; j < N; ++j) {
                const int py = j <= N / 2 ? j : j - N;
                for (int k = 1; k < N / 2; ++k) shell(px, py, k, 2);   // weight-2 interior modes
                shell(px, py, 0, 1);                                  // kz = 0
                shell(px, py, N / 2, 1);                              // kz = N/2
            }
        }
        res.resize(nb);
        for (int b = 0; b < nb; ++b) res[b] = B * (np[b] > 0 ? pp[b] / np[b] : pp[b]);
    }
    memcpy(out, res.data(), sizeof(double) * (size_t)nb);
    return IE_OK;
}

/* ------------------------------------------------------------- spectra */

ie_status ie_spectrum(ie_run *r, int which, double norm1, double *p, int *counts, double *vals)
{
    if (!r || !p || !counts || !vals) return fail(r, IE_ERR_ARG, "null argument");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    const int split = r->cfg.precision == IE_SPLIT;
    const int nb = fft_numbins(r->N);
    std::vector<double> sums(nb);
    if (which == IE_FIELD_DELTAN) {
        IE_HIP(r, fft_spectrum_double(r->fft, r->F.dn, sums.data(), r->stream));
    } else if (which == IE_FIELD_F || which == IE_FIELD_FD) {
        const int par = r->bg_host->parity;
        if (r->in_deltaN) {
            const double *src = which == IE_FIELD_F ? (split ? r->dn_f : (const double *)r->F.f[par])
                                                    : (split ? r->dn_fd : (const double *)r->F.fd[par]);
            IE_HIP(r, fft_spectrum_double(r->fft, src, sums.data(), r->stream));
        } else if (split) {
            const float *src = (const float *)(which == IE_FIELD_F ? r->F.f[par] : r->F.fd[par]);
            const double dc = (double)r->gs * (which == IE_FIELD_F ? r->bg_host->phibar : r->bg_host->pibar);
            IE_HIP(r, fft_spectrum_float(r->fft, src, dc, sums.data(), r->stream));
        } else {
            const double *src = (const double *)(which == IE_FIELD_F ? r->F.f[par] : r->F.fd[par]);
            IE_HIP(r, fft_spectrum_double(r->fft, src, sums.data(), r->stream));
        }
    } else {
        return fail(r, IE_ERR_ARG, "bad field");
    }
    memcpy(counts, fft_counts_cached(r->N, 0), sizeof(int) * (size_t)nb);
    const double pi = 2.0 * asin(1.0), dp = 2.0 * pi / r->cfg.L;
    for (int b = 0; b < nb; ++b) {
        p[b] = dp * b;
        double v = sums[b];
        if (counts[b] > 0) v /= (double)counts[b];
        vals[b] = norm1 * v;
    }
    return IE_OK;
}

ie_status ie_gw_spectrum(ie_run *r, int hdot, double *p, int *counts, double *vals)
{
    if (!r || !p || !counts || !vals) return fail(r, IE_ERR_ARG, "null argument");
    if (!r->F.sigw) return fail(r, IE_ERR_STATE, "run has no tensors (calculate_sigw = 0)");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    const int par = r->bg_host->parity;
    for (int c = 0; c < 6; ++c)
        IE_HIP(r, fft_forward_float(r->fft, hdot ? r->F.hd[c] : r->F.h[par][c], 0.0, c, r->stream));
    // spectraGWdot (src/output.cpp:644-648): B a^{s-1} in double; its
    // post-inflation copy rounds the factor to float (:662-669)
    double to_phys = hdot ? r->cfg.rescale_B * pow(r->bg_host->a, r->cfg.rescale_s - 1.0) : 1.0;
    if (hdot && r->in_post) to_phys = (double)(float)(r->cfg.rescale_B * pow(r->bg_host->a, r->cfg.rescale_s - 1.f));
    const int nb = fft_gw_numbins(r->N);
    std::vector<double> sums(nb);
    IE_HIP(r, fft_bin_tt(r->fft, r->dx, to_phys, sums.data(), r->stream));
    memcpy(counts, fft_counts_cached(r->N, 1), sizeof(int) * (size_t)nb);
    // output units follow src/output.cpp:475,618-621 (float p and norm1)
    const float dp = 2.f * (float)(2.0 * asin(1.0)) / (float)r->cfg.L;
    const float norm1 = 0.5f * pow(r->cfg.L / r->cfg.rescale_B, 3) / pow((float)r->N, 6);
    for (int b = 0; b < nb; ++b) {
        p[b] = (double)(dp * b);
        double v = sums[b];
        if (counts[b] > 0) v /= (double)counts[b];
        vals[b] = (double)norm1 * v;
    }
    return IE_OK;
}

ie_status ie_forcing_cutoff(ie_run *r, double high, double low)
{
    if (!r) return fail(r, IE_ERR_ARG, "null argument");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    const int N = r->N;
    DevBg &b = *r->bg_host;
    const int par = b.parity;
    if (r->in_deltaN) {
        // FP64 state of the deltaN and post-inflation phases
        double *f, *fd;
        if ((st = dn_state(r, &f, &fd)) != IE_OK) return st;
        IE_HIP(r, fft_forward_double(r->fft, f, 0, r->stream));
        IE_HIP(r, fft_forward_double(r->fft, fd, 1, r->stream));
        for (int q = 0; q < 2; ++q) IE_HIP(r, fft_kill_modes(r->fft, q, high, low, r->stream));
        IE_HIP(r, fft_inverse_slot(r->fft, 0, 0.0, f, r->stream));
        IE_HIP(r, fft_inverse_slot(r->fft, 1, 0.0, fd, r->stream));
        return sync(r);
    }
    if (r->cfg.precision == IE_FP64) {
        double *f = (double *)r->F.f[par], *fd = (double *)r->F.fd[par];
        IE_HIP(r, fft_forward_double(r->fft, f, 0, r->stream));
        IE_HIP(r, fft_forward_double(r->fft, fd, 1, r->stream));
        for (int q = 0; q < 2; ++q) IE_HIP(r, fft_kill_modes(r->fft, q, high, low, r->stream));
        IE_HIP(r, fft_inverse_slot(r->fft, 0, 0.0, f, r->stream));
        IE_HIP(r, fft_inverse_slot(r->fft, 1, 0.0, fd, r->stream));
        return sync(r);
    }
    // split storage: the fluctuations carry every mode but k = 0's homogeneous
    // part, which goes with the FP64 mean when the low cutoff removes k = 0
    float *f = (float *)r->F.f[par], *fd = (float *)r->F.fd[par];
    IE_HIP(r, fft_forward_float(r->fft, f, 0.0, 0, r->stream));
    IE_HIP(r, fft_forward_float(r->fft, fd, 0.0, 1, r->stream));
    for (int q = 0; q < 2; ++q) IE_HIP(r, fft_kill_modes(r->fft, q, high, low, r->stream));
    IE_HIP(r, fft_inverse_slot(r->fft, 0, 0.0, r->scratch, r->stream));
    IE_HIP(r, store_float(r->scratch, f, r->gs, r->stream));
    IE_HIP(r, fft_inverse_slot(r->fft, 1, 0.0, r->scratch, r->stream));
    IE_HIP(r, store_float(r->scratch, fd, r->gs, r->stream));
    if (0.0 > high || 0.0 < low) {
        b.phibar = 0.0;
        b.pibar = 0.0;
        return push_bg(r);
    }
    (void)N;
    return sync(r);
}

ie_status ie_bispectrum(ie_run *r, double *p, int *counts, double *re, double *im)
{
    if (!r || !p || !counts || !re || !im) return fail(r, IE_ERR_ARG, "null argument");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    const int par = r->bg_host->parity;
    if (r->in_deltaN) {
        double *f, *fd;
        if ((st = dn_state(r, &f, &fd)) != IE_OK) return st;
        IE_HIP(r, fft_forward_double(r->fft, f, 0, r->stream));
    } else if (r->cfg.precision == IE_SPLIT) {
        IE_HIP(r, fft_forward_float(r->fft, (const float *)r->F.f[par], 0.0, 0, r->stream));
    } else {
        IE_HIP(r, fft_forward_double(r->fft, (const double *)r->F.f[par], 0, r->stream));
    }
    // the reference transforms f minus its mean (src/output.cpp:766-771)
    IE_HIP(r, fft_zero_dc(r->fft, 0, r->stream));
    const int nb = (int)(sqrt(3.0) * (r->N / 2)) + 1;
    std::vector<double> sums((size_t)nb * 3);
    IE_HIP(r, bispectrum_sums(r->fft, 0, r->N, sums.data(), r->stream));
    const double pi = 2.0 * asin(1.0), dp = 2.0 * pi / r->cfg.L;
    const double norm1 = pow(r->cfg.L / r->cfg.rescale_B, 6) / pow((double)r->N, 9);
    for (int k = 0; k < nb; ++k) {
        p[k] = dp * k;
        counts[k] = (int)sums[3 * k + 2];
        double a = sums[3 * k], c = sums[3 * k + 1];
        if (counts[k] > 0) {
            a = a / counts[k];
            c = c / counts[k];
        }
        re[k] = norm1 * a;
        im[k] = norm1 * c;
    }
    return IE_OK;
}

ie_status ie_deltaN_log(ie_run *r, double eta_log, int nbins, double *bmin, double *df, double *freq, int *valid)
{
    if (!r || !bmin || !df || !freq || !valid || nbins < 1 || !(eta_log != 0.0))
        return fail(r, IE_ERR_ARG, "bad argument");
    if (r->in_deltaN) return fail(r, IE_ERR_STATE, "the LOG map is formed from the inflation fields");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    const DevBg &b = *r->bg_host;
    const int par = b.parity, split = r->cfg.precision == IE_SPLIT;
    if (r->hist_bins < nbins) {
        if (r->hist) (void)hipFree(r->hist);
        IE_HIP(r, hipMalloc((void **)&r->hist, (size_t)nbins * 8));
        r->hist_bins = nbins;
    }
    LogMapResult res{};
    IE_HIP(r, log_map(r->F.f[par], r->F.fd[par], split, r->gs, split ? b.phibar : 0.0, split ? b.pibar : 0.0, b.a,
                      b.ad, r->cfg.rescale_s, eta_log, nbins, r->F.dn, r->work, r->hist, &res, r->stream));
    *valid = res.valid;
    *bmin = res.bmin;
    *df = res.df;
    if (!res.valid) {
        for (int q = 0; q < nbins; ++q) freq[q] = 0.0;
        return IE_OK;
    }
    std::vector<unsigned long long> c(nbins);
    IE_HIP(r, hipMemcpyAsync(c.data(), r->hist, (size_t)nbins * 8, hipMemcpyDeviceToHost, r->stream));
    if ((st = sync(r)) != IE_OK) return st;
    double numpts = 0;
    for (int q = 0; q < nbins; ++q) numpts += (double)c[q];
    if (numpts == 0) numpts = 1;
    for (int q = 0; q < nbins; ++q) freq[q] = (double)c[q] / numpts;
    return IE_OK;
}

ie_status ie_upload_modes(ie_run *r, const double *fk, const double *fny, const double *fdk, const double *fdny,
                          double phi0, double dphi0)
{
    if (!r || !fk || !fny || !fdk || !fdny) return fail(r, IE_ERR_ARG, "null argument");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    r->bg_host->parity = 0;
    if ((st = push_bg(r)) != IE_OK) return st;
    if (r->cfg.precision == IE_FP64) {
        IE_HIP(r, fft_inverse_ffteasy(r->fft, fk, fny, phi0, (double *)r->F.f[0], r->stream));
        IE_HIP(r, fft_inverse_ffteasy(r->fft, fdk, fdny, dphi0, (double *)r->F.fd[0], r->stream));
        return sync(r);
    }
    // split: both fields in FP64 in scratch and the deltaN phase's FP64 buffer
    // (unused before that phase), then the homogeneous mode and the fluctuations
    IE_HIP(r, fft_inverse_ffteasy(r->fft, fk, fny, phi0, r->scratch, r->stream));
    IE_HIP(r, fft_inverse_ffteasy(r->fft, fdk, fdny, dphi0, r->dn_fd, r->stream));
    IE_HIP(r, launch_split_upload(r->scratch, r->dn_fd, r->F, r->bg, r->partials, r->stream));
    return sync(r);
}

ie_status ie_initialize(ie_run *r, const ie_ic_params *ic, ie_background *out)
{
    if (!r || !ic) return fail(r, IE_ERR_ARG, "null argument");
    const size_t ny = (size_t)r->N * 2 * r->N;
    // left uninitialised: ie_fill_modes zeroes them (in parallel)
    std::unique_ptr<double[]> fk(new double[r->gs]), fdk(new double[r->gs]), fny(new double[ny]),
        fdny(new double[ny]);
    double H = 0.0;
    ie_status st = ie_fill_modes(&r->cfg, ic, fk.get(), fny.get(), fdk.get(), fdny.get(), &H);
    if (st == IE_ERR_COURANT) {
        char buf[160];
        snprintf(buf, sizeof(buf), "Time step too large. Courant condition violated: dt/dx = %f", r->cfg.dt / r->dx);
        return fail(r, st, buf);
    }
    if (st == IE_ERR_TABLE_RANGE) return fail(r, st, "initial_field outside the numerical potential range");
    if (st != IE_OK) return fail(r, st, "mode fill failed");
    DevBg &b = *r->bg_host;
    if ((st = pull_bg(r)) != IE_OK) return st;
    b.t = 0.0;
    b.a = 1.0;
    b.ad = H;
    b.ad2 = 0.0;
    b.astep = 1.0;
    b.Ne = 0.0;
    b.phiref = 0.0;
    b.steps = 0;
    b.done = 0;
    b.error = 0;
    if ((st = push_bg(r)) != IE_OK) return st;
    st = ie_upload_modes(r, fk.get(), fny.get(), fdk.get(), fdny.get(), ic->initial_field,
                         ic->initial_derivative);
    if (st != IE_OK) return st;
    // H0 from the lattice energies (src/initialize.cpp:202-212)
    if ((st = sums(r)) != IE_OK) return st;
    double S2 = b.red[3];
    if (r->cfg.precision == IE_SPLIT) S2 = (double)r->gs * b.pibar * b.pibar + 2.0 * b.pibar * b.red[1] + b.red[3];
    const double kin = 0.5 * S2 / (double)r->gs;
    const double pe = b.red[5] / (double)r->gs;
    const double ge = 0.5 * b.red[4] * ((1.0 / r->dx) * (1.0 / r->dx)) / (double)r->gs;
    b.hubble_init = sqrt((kin + pe + ge) / 3.0);
    b.ad = b.hubble_init;
    if ((st = push_bg(r)) != IE_OK) return st;
    if (out) return ie_get_background(r, out);
    return IE_OK;
}

/* ------------------------------------------------------------- deltaN */

// FP64 state of the deltaN and post-inflation phases at the host mirror's parity
static void dn_buffers(ie_run *r, double **f, double **fd)
{
    if (r->cfg.precision == IE_SPLIT) {
        *f = r->dn_f;
        *fd = r->dn_fd;
    } else {
        *f = (double *)r->F.f[r->bg_host->parity];
        *fd = (double *)r->F.fd[r->bg_host->parity];
    }
}

static ie_status dn_state(ie_run *r, double **f, double **fd)
{
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    if (r->cfg.precision == IE_SPLIT) {
        *f = r->dn_f;
        *fd = r->dn_fd;
    } else {
        *f = (double *)r->F.f[r->bg_host->parity];
        *fd = (double *)r->F.fd[r->bg_host->parity];
    }
    return IE_OK;
}

static int dn_error(ie_run *r)
{
    int e = 0;
    (void)hipMemcpyAsync(&e, &r->bg->error, sizeof(int), hipMemcpyDeviceToHost, r->stream);
    (void)hipStreamSynchronize(r->stream);
    return e;
}

static double ref_potential(const ie_run *r, double v, bool *ok)
{
    // potential(phiref) (src/potential.cpp:91-98) on the host
    const ie_config &c = r->cfg;
    *ok = true;
    const double B2 = c.rescale_B * c.rescale_B;
    if (!c.numerical_potential) return c.V0 * (1.0 - (1.0 - c.ns) / 4.0 * (v * v)) / B2;
    std::vector<double> x(c.tab_n);
    if (hipMemcpy(x.data(), r->table, c.tab_n * 8, hipMemcpyDeviceToHost) != hipSuccess) {
        *ok = false;
        return 0.0;
    }
    if (v > x[0] || v < x[c.tab_n - 1]) {
        *ok = false;
        return 0.0;
    }
    int lo = 0, hi = c.tab_n - 1;
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        if (x[mid] >= v) lo = mid;
        else hi = mid;
    }
    double y[2];
    if (hipMemcpy(y, r->table + c.tab_n + hi - 1, 16, hipMemcpyDeviceToHost) != hipSuccess) {
        *ok = false;
        return 0.0;
    }
    return (y[1] + (v - x[hi]) * (y[0] - y[1]) / (x[hi - 1] - x[hi])) / B2;
}

ie_status ie_deltaN_init(ie_run *r)
{
    if (!r) return fail(r, IE_ERR_ARG, "null argument");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    if (r->cfg.precision == IE_SPLIT) {
        IE_HIP(r, launch_reconstitute(r->F, 1, 0, r->bg, r->dn_f, r->stream));
        IE_HIP(r, launch_reconstitute(r->F, 1, 1, r->bg, r->dn_fd, r->stream));
    }
    r->in_deltaN = 1;
    double *f, *fd;
    if ((st = dn_state(r, &f, &fd)) != IE_OK) return st;
    // a^{s-1} on the host with the C library's pow, as the reference
    const double apow = pow(r->bg_host->a, r->cfg.rescale_s - 1.0);
    IE_HIP(r, dn_init(r->F, f, fd, r->F.dn, r->P, apow, &r->bg->error, r->stream));
    if (dn_error(r)) return check_kernel_error(r);
    return sync(r);
}

ie_status ie_evolve_fieldsN(ie_run *r, double d)
{
    if (!r) return fail(r, IE_ERR_ARG, "null argument");
    if (!r->in_deltaN || r->in_post)
        return fail(r, IE_ERR_STATE, r->in_post ? "run is in its post-inflation phase" : "deltaN phase not initialised");
    double *f, *fd;
    ie_status st = dn_state(r, &f, &fd);
    if (st != IE_OK) return st;
    DevBg &b = *r->bg_host;
    bool ok = true;
    const double Vref = r->cfg.monotonic == IE_GENERIC ? ref_potential(r, b.phiref, &ok) : 0.0;
    if (!ok) return fail(r, IE_ERR_TABLE_RANGE, "phiref outside the numerical potential range");
    IE_HIP(r, dn_fieldsN(r->gs, f, fd, r->F.dn, r->P, r->cfg.monotonic, b.phiref, Vref, d, r->stream));
    b.t += d;
    b.a += d * b.ad;
    return push_bg(r);
}

ie_status ie_get_phiref(ie_run *r, double *out)
{
    if (!r || !out) return fail(r, IE_ERR_ARG, "null argument");
    if (!r->in_deltaN || r->in_post)
        return fail(r, IE_ERR_STATE, r->in_post ? "run is in its post-inflation phase" : "deltaN phase not initialised");
    double *f, *fd;
    ie_status st = dn_state(r, &f, &fd);
    if (st != IE_OK) return st;
    IE_HIP(r, dn_phiref(r->gs, f, r->P, r->cfg.monotonic, r->work, r->work + 4000, r->stream));
    IE_HIP(r, hipMemcpyAsync(out, r->work + 4000, 8, hipMemcpyDeviceToHost, r->stream));
    return sync(r);
}

ie_status ie_deltaN_steps(ie_run *r, double dN, int64_t nsteps)
{
    if (!r || nsteps < 0) return fail(r, IE_ERR_ARG, "bad argument");
    if (!r->in_deltaN || r->in_post)
        return fail(r, IE_ERR_STATE, r->in_post ? "run is in its post-inflation phase" : "deltaN phase not initialised");
    double *f, *fd;
    ie_status st = dn_state(r, &f, &fd);
    if (st != IE_OK) return st;
    DevBg &b = *r->bg_host;
    bool ok = true;
    const double Vref = r->cfg.monotonic == IE_GENERIC ? ref_potential(r, b.phiref, &ok) : 0.0;
    if (!ok) return fail(r, IE_ERR_TABLE_RANGE, "phiref outside the numerical potential range");
    IE_HIP(r, dn_steps(r->gs, f, fd, r->F.dn, r->P, r->cfg.monotonic, b.phiref, Vref, dN, (long long)nsteps,
                       &r->bg->error, r->stream));
    if ((st = check_kernel_error(r)) != IE_OK) return st;
    // t and a advance with every step (src/evolution.cpp:1017,1036)
    for (int64_t q = 0; q < nsteps; ++q) {
        r->bg_host->t += dN;
        r->bg_host->a += dN * r->bg_host->ad;
    }
    return push_bg(r);
}

ie_status ie_deltaN_rk4_steps(ie_run *r, double h, int64_t nsteps)
{
    if (!r || nsteps < 0) return fail(r, IE_ERR_ARG, "bad argument");
    if (!r->in_deltaN || r->in_post)
        return fail(r, IE_ERR_STATE, r->in_post ? "run is in its post-inflation phase" : "deltaN phase not initialised");
    double *f, *fd;
    ie_status st = dn_state(r, &f, &fd);
    if (st != IE_OK) return st;
    DevBg &b = *r->bg_host;
    bool ok = true;
    const double Vref = r->cfg.monotonic == IE_GENERIC ? ref_potential(r, b.phiref, &ok) : 0.0;
    if (!ok) return fail(r, IE_ERR_TABLE_RANGE, "phiref outside the numerical potential range");
    IE_HIP(r, dn_rk4_steps(r->gs, f, fd, r->F.dn, r->P, r->cfg.monotonic, b.phiref, Vref, h, (long long)nsteps,
                           &r->bg->error, r->stream));
    if ((st = check_kernel_error(r)) != IE_OK) return st;
    // a advances by h sum_s B_s ad per step, t by h (src/evolution.cpp:1168-1173)
    const double B[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    for (int64_t q = 0; q < nsteps; ++q) {
        for (int s = 0; s < 4; ++s) r->bg_host->a += h * B[s] * r->bg_host->ad;
        r->bg_host->t += h;
    }
    return push_bg(r);
}

ie_status ie_deltaN_rk45_step(ie_run *r, double h, double hmax, int *accepted, double *h_next)
{
    if (!r || !accepted || !h_next || !(h > 0.0)) return fail(r, IE_ERR_ARG, "bad argument");
    if (!r->in_deltaN || r->in_post)
        return fail(r, IE_ERR_STATE, r->in_post ? "run is in its post-inflation phase" : "deltaN phase not initialised");
    double *f, *fd;
    ie_status st = dn_state(r, &f, &fd);
    if (st != IE_OK) return st;
    if (!r->dn5[0])
        for (int q = 0; q < 3; ++q)
            if ((st = dev_alloc(r, (void **)&r->dn5[q], r->gs * 8)) != IE_OK) return st;
    const ie_config &c = r->cfg;
    DevBg &b = *r->bg_host;
    bool ok = true;
    const double Vref = c.monotonic == IE_GENERIC ? ref_potential(r, b.phiref, &ok) : 0.0;
    if (!ok) return fail(r, IE_ERR_TABLE_RANGE, "phiref outside the numerical potential range");
    double *total = r->work + 4000;
    IE_HIP(r, dn_rk45_trial(r->gs, f, fd, r->F.dn, r->dn5[0], r->dn5[1], r->dn5[2], r->P, c.monotonic, b.phiref,
                            Vref, h, c.rk45_abs_tol, c.rk45_rel_tol, r->work, total, &r->bg->error, r->stream));
    double acc_sum = 0.0;
    IE_HIP(r, hipMemcpyAsync(&acc_sum, total, 8, hipMemcpyDeviceToHost, r->stream));
    if ((st = check_kernel_error(r)) != IE_OK) return st;
    // the scale factor's error term: a' = ad in every stage (src/evolution.cpp:1113)
    const double Bw[7] = {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0};
    const double Ew[7] = {35.0 / 384.0 - 5179.0 / 57600.0,       0,
                          500.0 / 1113.0 - 7571.0 / 16695.0,     125.0 / 192.0 - 393.0 / 640.0,
                          -2187.0 / 6784.0 + 92097.0 / 339200.0, 11.0 / 84.0 - 187.0 / 2100.0,
                          -1.0 / 40.0};
    double a5 = b.a, erra = 0.0;
    for (int q = 0; q < 7; ++q) {
        if (Bw[q] != 0.0) a5 += h * Bw[q] * b.ad;
        if (Ew[q] != 0.0) erra += h * Ew[q] * b.ad;
    }
    const double sa = c.rk45_abs_tol + c.rk45_rel_tol * std::max(fabs(b.a), fabs(a5));
    acc_sum += (erra / sa) * (erra / sa);
    const double norm = sqrt(acc_sum / (double)(3 * r->gs + 1));
    if (!isfinite(norm)) return fail(r, IE_ERR_NONFINITE, "RK45 error norm is not finite");
    const int acc = rk45_propose(c, h, hmax, norm, h_next);
    *accepted = acc;
    if (!acc) return IE_OK;
    // adopt the trial state: the previous buffers become the next trial's
    if (c.precision == IE_SPLIT) {
        std::swap(r->dn_f, r->dn5[0]);
        std::swap(r->dn_fd, r->dn5[1]);
    } else {
        double *pf = (double *)r->F.f[b.parity], *pfd = (double *)r->F.fd[b.parity];
        r->F.f[b.parity] = r->dn5[0];
        r->F.fd[b.parity] = r->dn5[1];
        r->dn5[0] = pf;
        r->dn5[1] = pfd;
    }
    std::swap(r->F.dn, r->dn5[2]);
    r->scratch = r->F.dn;
    b.a = a5;
    b.t += h;
    return push_bg(r);
}

ie_status ie_deltaN_zeta(ie_run *r, double cutoff, float *nmean)
{
    if (!r || !nmean) return fail(r, IE_ERR_ARG, "null argument");
    IE_HIP(r, dn_float_mean(r->F.dn, r->gs, cutoff, r->nmean, r->stream));
    IE_HIP(r, dn_zeta(r->F.dn, r->gs, r->nmean, cutoff, 0, r->stream));
    IE_HIP(r, hipMemcpyAsync(nmean, r->nmean, 4, hipMemcpyDeviceToHost, r->stream));
    return sync(r);
}

ie_status ie_deltaN_zero_unfinished(ie_run *r, double cutoff)
{
    if (!r) return fail(r, IE_ERR_ARG, "null argument");
    IE_HIP(r, dn_zeta(r->F.dn, r->gs, r->nmean, cutoff, 1, r->stream));
    return sync(r);
}

ie_status ie_deltaN_center(ie_run *r)
{
    if (!r) return fail(r, IE_ERR_ARG, "null argument");
    IE_HIP(r, field_center(r->F.dn, r->gs, r->work, r->work + 4000, r->stream));
    return sync(r);
}

ie_status ie_histogram(ie_run *r, int which, int nbins, double *bmin, double *bmax, double *counts)
{
    if (!r || !bmin || !bmax || !counts || nbins < 1) return fail(r, IE_ERR_ARG, "bad argument");
    const double *g = nullptr;
    ie_status st;
    if (which == IE_FIELD_DELTAN) g = r->F.dn;
    else if (which == IE_FIELD_F) {
        if (r->in_deltaN) {
            double *f, *fd;
            if ((st = dn_state(r, &f, &fd)) != IE_OK) return st;
            g = f;
        } else {
            if (r->cfg.precision == IE_SPLIT) {
                IE_HIP(r, launch_reconstitute(r->F, 1, 0, r->bg, r->scratch, r->stream));
                g = r->scratch;
            } else {
                if ((st = pull_bg(r)) != IE_OK) return st;
                g = (const double *)r->F.f[r->bg_host->parity];
            }
        }
    } else return fail(r, IE_ERR_ARG, "bad field");
    int nb = 0;
    IE_HIP(r, field_minmax(g, r->gs, r->work, r->stream, &nb));
    std::vector<double> mm(2 * nb);
    IE_HIP(r, hipMemcpyAsync(mm.data(), r->work, 2 * nb * 8, hipMemcpyDeviceToHost, r->stream));
    if ((st = sync(r)) != IE_OK) return st;
    double lo = mm[0], hi = mm[nb];
    for (int q = 0; q < nb; ++q) {
        lo = mm[q] < lo ? mm[q] : lo;
        hi = mm[nb + q] > hi ? mm[nb + q] : hi;
    }
    double df = (hi - lo) / (double)nbins;
    if (!std::isfinite(df) || df <= 0.0) df = 1.0;
    if (r->hist_bins < nbins) {
        if (r->hist) (void)hipFree(r->hist);
        IE_HIP(r, hipMalloc((void **)&r->hist, (size_t)nbins * 8));
        r->hist_bins = nbins;
    }
    IE_HIP(r, hipMemsetAsync(r->hist, 0, (size_t)nbins * 8, r->stream));
    IE_HIP(r, field_histogram(g, r->gs, lo, hi, df, nbins, r->hist, r->stream));
    std::vector<unsigned long long> c(nbins);
    IE_HIP(r, hipMemcpyAsync(c.data(), r->hist, (size_t)nbins * 8, hipMemcpyDeviceToHost, r->stream));
    if ((st = sync(r)) != IE_OK) return st;
    for (int q = 0; q < nbins; ++q) counts[q] = (double)c[q];
    *bmin = lo;
    *bmax = hi;
    return IE_OK;
}

/* ------------------------------------------------------ post-inflation */

// Phi, Phi' (FP64) and h of the current step and the buffers its results go to,
// from the host mirror of the background (the callers pull it once, and the
// leapfrog loop advances it between launches)
static ie_status post_state(ie_run *r, PostState *S)
{
    double *f, *fd;
    dn_buffers(r, &f, &fd);
    const int par = r->bg_host->parity;
    S->f = f;
    S->fd = fd;
    if (r->cfg.precision == IE_SPLIT) {
        S->f_out = r->post_f2;
        S->fd_out = r->post_fd2;
    } else {
        S->f_out = (double *)r->F.f[par ^ 1];
        S->fd_out = (double *)r->F.fd[par ^ 1];
    }
    for (int c = 0; c < 6; ++c) {
        S->h[c] = r->F.h[par][c];
        S->h_out[c] = r->F.h[par ^ 1][c];
        S->hd[c] = r->F.hd[c];
    }
    return IE_OK;
}

// after a step: the written buffers become current
static void post_flip(ie_run *r)
{
    r->bg_host->parity ^= 1;
    if (r->cfg.precision == IE_SPLIT) {
        std::swap(r->dn_f, r->post_f2);
        std::swap(r->dn_fd, r->post_fd2);
    }
}

// Scalars of one kick of evolve_derivs_post_inflation (src/evolution.cpp:1440-1464)
// at the background b, whose ad2 and first half kick of ad it applies.
static PostStep post_kick_scalars(ie_run *r, DevBg &b, double d, double dd)
{
    const double s = r->cfg.rescale_s, w = r->post.omega, dx = r->dx;
    PostStep P{};
    P.d = d;
    P.dd = dd;
    P.laplnorm = pow(b.a, -2.0 * s) / (dx * dx);
    b.ad2 = -(s - 0.5 * (1.0 - 3.0 * w)) * (b.ad * b.ad) / b.a;
    b.ad += 0.5 * d * b.ad2;
    P.A = w * P.laplnorm;
    P.c1ad = (3.0 * (1.0 + w) + s) * b.ad;
    P.a = b.a;
    const double Htilde = b.ad / b.a;
    P.invH = 1.0 / Htilde;
    P.coeffU = 4.0 / (3.0 * (1.0 + w));
    P.srcamp = 2.0 * pow(b.a, -2.0 * s);
    P.c2ad = (2.0 + s) * b.ad;
    P.hx = 0.5 / dx;
    P.inv_dx2 = 1.0 / (dx * dx);
    P.cross_den = 4.0 * dx * dx;
    return P;
}

static ie_status post_drift_call(ie_run *r, double d)
{
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    PostState S;
    if ((st = post_state(r, &S)) != IE_OK) return st;
    IE_HIP(r, post_drift(S, r->gs, d, r->stream));
    DevBg &b = *r->bg_host;
    b.t += d;
    b.a += d * b.ad;
    return push_bg(r);
}

extern "C" {

ie_status ie_post_init(ie_run *r, const ie_post_params *p, ie_background *out)
{
    if (!r || !p) return fail(r, IE_ERR_ARG, "null argument");
    if (!r->F.sigw) return fail(r, IE_ERR_STATE, "the post-inflation phase needs calculate_sigw = 1");
    if (!(p->dt > 0.0) || !(1.0 + p->omega != 0.0) || !(5.0 + 3.0 * p->omega != 0.0) ||
        (p->integrator != IE_LEAPFROG && p->integrator != IE_RK4 && p->integrator != IE_RK45))
        return fail(r, IE_ERR_ARG, "bad post-inflation parameters");
    if (r->in_post) return fail(r, IE_ERR_STATE, "run is already in its post-inflation phase");
    if (p->from_deltaN && !r->in_deltaN) return fail(r, IE_ERR_STATE, "deltaN phase not initialised");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    const int split = r->cfg.precision == IE_SPLIT;
    if (split) {
        if (!r->in_deltaN) {
            IE_HIP(r, launch_reconstitute(r->F, 1, 0, r->bg, r->dn_f, r->stream));
            IE_HIP(r, launch_reconstitute(r->F, 1, 1, r->bg, r->dn_fd, r->stream));
        }
        if (!r->post_f2) {
            if ((st = dev_alloc(r, (void **)&r->post_f2, r->gs * 8)) != IE_OK) return st;
            if ((st = dev_alloc(r, (void **)&r->post_fd2, r->gs * 8)) != IE_OK) return st;
        }
    }
    r->in_deltaN = 1;
    double *f, *fd;
    if ((st = dn_state(r, &f, &fd)) != IE_OK) return st;
    for (int q = 0; q < 2; ++q)
        for (int c = 0; c < 6; ++c) IE_HIP(r, hipMemsetAsync(r->F.h[q][c], 0, r->gs * 4, r->stream));
    for (int c = 0; c < 6; ++c) IE_HIP(r, hipMemsetAsync(r->F.hd[c], 0, r->gs * 4, r->stream));
    const double w = p->omega, s = r->cfg.rescale_s;
    const double Phi_from_zeta = 3.0 * (1.0 + w) / (5.0 + 3.0 * w);
    double sums2[2];
    IE_HIP(r, serial_sums(p->from_deltaN ? r->F.dn : f, p->from_deltaN ? nullptr : fd, r->gs, r->work + 4000,
                          r->stream));
    IE_HIP(r, hipMemcpyAsync(sums2, r->work + 4000, 16, hipMemcpyDeviceToHost, r->stream));
    if ((st = sync(r)) != IE_OK) return st;
    DevBg &b = *r->bg_host;
    if (p->from_deltaN) {
        const double Nmean = sums2[0] / (double)r->gs;
        IE_HIP(r, post_phi(f, r->F.dn, r->gs, Phi_from_zeta, Nmean, 1.0, 0, r->stream));
    } else {
        const double fmean = sums2[0] / (double)r->gs, fdmean = sums2[1] / (double)r->gs;
        const double den = fdmean * pow(b.a, s - 1.0) / (b.ad * pow(b.a, s - 2.0));
        IE_HIP(r, post_phi(f, f, r->gs, Phi_from_zeta, fmean, den, 1, r->stream));
    }
    r->post = *p;
    if (!(r->post.af > 0.0)) r->post.af = 2.0 * r->N;
    r->in_post = 1;
    b.t = 0.0;
    b.a = 1.0;
    b.ad = p->horizon_factor * r->N * 2.0 * (2.0 * asin(1.0)) / r->cfg.L;
    b.steps = 0;
    b.done = 0;
    if ((st = push_bg(r)) != IE_OK) return st;
    if (out) return ie_get_background(r, out);
    return IE_OK;
}


// The post-inflation tensor step runs as the plane-marching kernel where the
// lattice allows it, unless the run's pinned tensor shape is a row kernel code
// (1, 2, 4), which keeps the row kernel (the path of N >= 512 lattices).
static bool post_tensor_rows(const ie_run *r)
{
    return r->pinned && r->F.tensor_e > 0 && r->F.tensor_e <= 4;
}

ie_status ie_post_evolve_derivs(ie_run *r, double d)
{
    if (!r || !(d != 0.0)) return fail(r, IE_ERR_ARG, "bad step");
    if (!r->in_post) return fail(r, IE_ERR_STATE, "post-inflation phase not initialised");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    PostState S;
    if ((st = post_state(r, &S)) != IE_OK) return st;
    DevBg &b = *r->bg_host;
    const PostStep P = post_kick_scalars(r, b, d, 0.0);
    IE_HIP(r, post_step(S, r->N, r->logN, P, r->stream, post_tensor_rows(r)));
    b.ad += 0.5 * d * b.ad2;
    post_flip(r);
    return push_bg(r);
}

ie_status ie_post_leapfrog_steps(ie_run *r, int max_steps, ie_background *out)
{
    if (!r || max_steps < 0) return fail(r, IE_ERR_ARG, "bad argument");
    if (!r->in_post) return fail(r, IE_ERR_STATE, "post-inflation phase not initialised");
    ie_status st = pull_bg(r);
    if (st != IE_OK) return st;
    DevBg &b = *r->bg_host;
    const double d = r->post.dt;
    for (int n = 0; n < max_steps && b.a <= r->post.af; ++n) {
        PostState S;
        if ((st = post_state(r, &S)) != IE_OK) return st;
        const PostStep P = post_kick_scalars(r, b, d, d);
        IE_HIP(r, post_step(S, r->N, r->logN, P, r->stream, post_tensor_rows(r)));
        b.ad += 0.5 * d * b.ad2;
        b.t += d;
        b.a += d * b.ad;
        b.steps += 1;
        b.d = d;
        post_flip(r);
    }
    b.done = !(b.a <= r->post.af) ? 1 : 0;
    if ((st = push_bg(r)) != IE_OK) return st;
    if (out) background_out(r, out);   // the host copy just written
    return IE_OK;
}

}  // extern "C"

/* ------------------------------------------------------------- batches */

}  // extern "C"

struct ie_batch {
    std::vector<ie_run *> runs;
    BatchDev *dev = nullptr;
    hipStream_t stream = nullptr;
    // captured runs of graph_steps loop steps, by the parity of their first step
    // (the steps' kernels take the step records' slots from it)
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    int graph_steps = 0;
    std::string err;
};

static ie_status bfail(ie_batch *b, ie_status st, const std::string &m)
{
    if (b) b->err = m;
    else g_create_error = m;
    return st;
}

#define IE_BHIP(b, call)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return bfail((b), IE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" {

const char *ie_batch_last_error(const ie_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

ie_status ie_batch_create(ie_run *const *runs, int nruns, ie_batch **out)
{
    if (!runs || nruns < 1 || !out) return bfail(nullptr, IE_ERR_ARG, "bad argument");
    *out = nullptr;
    const int split = runs[0]->cfg.precision == IE_SPLIT;
    std::vector<BatchRun> br(nruns);
    for (int q = 0; q < nruns; ++q) {
        ie_run *r = runs[q];
        if (!r) return bfail(nullptr, IE_ERR_ARG, "null run");
        if ((r->cfg.precision == IE_SPLIT) != split) return bfail(nullptr, IE_ERR_ARG, "mixed storage in one batch");
        if (r->in_deltaN) return bfail(nullptr, IE_ERR_STATE, "run is in its deltaN phase");
        if (hipStreamSynchronize(r->stream) != hipSuccess) return bfail(nullptr, IE_ERR_HIP, "run stream");
        br[q].F = r->F;
        br[q].P = r->P;
        br[q].bg = r->bg;
        br[q].partials = r->partials;
    }
    ie_batch *b = new ie_batch();
    b->runs.assign(runs, runs + nruns);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return bfail(nullptr, IE_ERR_HIP, "hipStreamCreate failed");
    }
    if (batch_build(br.data(), nruns, split, &b->dev) != hipSuccess) {
        ie_batch_destroy(b);
        return bfail(nullptr, IE_ERR_OOM, "batch descriptor upload failed");
    }
    *out = b;
    return IE_OK;
}

static void drop_graphs(ie_batch *b)
{
    for (int p = 0; p < 2; ++p) {
        if (b->exec[p]) (void)hipGraphExecDestroy(b->exec[p]);
        if (b->graph[p]) (void)hipGraphDestroy(b->graph[p]);
        b->exec[p] = nullptr;
        b->graph[p] = nullptr;
    }
}

ie_status ie_batch_launch(ie_batch *b, int nsteps, int graph_steps)
{
    if (!b || nsteps < 0 || graph_steps < 0) return bfail(b, IE_ERR_ARG, "bad argument");
    IE_BHIP(b, batch_loop_begin(b->dev, b->stream));
    int done = 0;
    if (graph_steps > 0 && nsteps >= graph_steps) {
        if (b->graph_steps != graph_steps) {
            drop_graphs(b);
            b->graph_steps = graph_steps;
        }
        for (; done + graph_steps <= nsteps; done += graph_steps) {
            const int p = done & 1;
            if (!b->exec[p]) {
                IE_BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                const hipError_t e = batch_steps(b->dev, p, graph_steps, b->stream);
                if (e != hipSuccess) {
                    hipGraph_t g;
                    (void)hipStreamEndCapture(b->stream, &g);
                    return bfail(b, IE_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
                }
                IE_BHIP(b, hipStreamEndCapture(b->stream, &b->graph[p]));
                IE_BHIP(b, hipGraphInstantiate(&b->exec[p], b->graph[p], nullptr, nullptr, 0));
            }
            IE_BHIP(b, hipGraphLaunch(b->exec[p], b->stream));
        }
    }
    if (done < nsteps) IE_BHIP(b, batch_steps(b->dev, done, nsteps - done, b->stream));
    if (nsteps > 0) IE_BHIP(b, batch_loop_end(b->dev, nsteps, b->stream));
    return IE_OK;
}

ie_status ie_batch_wait(ie_batch *b)
{
    if (!b) return IE_ERR_ARG;
    IE_BHIP(b, hipStreamSynchronize(b->stream));
    int raised = 0;
    for (ie_run *r : b->runs) {
        if (pull_bg(r) != IE_OK) return bfail(b, IE_ERR_HIP, r->err);
        if (r->bg_host->error) raised = 1;
    }
    if (raised) return bfail(b, IE_ERR_TABLE_RANGE, "field value outside the numerical potential range");
    return IE_OK;
}

ie_status ie_batch_set_resident_min(int min_hilltop, int min_tabulated)
{
    if (min_hilltop < 0 || min_tabulated < 0) return bfail(nullptr, IE_ERR_ARG, "negative minimum");
    set_resident_min(min_hilltop, min_tabulated);
    return IE_OK;
}

ie_status ie_batch_steps(ie_batch *b, int nsteps, int graph_steps)
{
    const ie_status st = ie_batch_launch(b, nsteps, graph_steps);
    return st != IE_OK ? st : ie_batch_wait(b);
}

ie_status ie_batch_destroy(ie_batch *b)
{
    if (!b) return IE_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    batch_free(b->dev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return IE_OK;
}

static ie_status rk_ready(ie_run *r, int stages, RkWork **w)
{
    if (!r->in_post) {
        if (r->cfg.precision != IE_FP64) return fail(r, IE_ERR_STATE, "RK integrators need FP64 storage");
        if (r->in_deltaN) return fail(r, IE_ERR_STATE, "run is in its deltaN phase");
    }
    RkWork **slot = stages == 4 ? &r->rk4 : &r->rk45;
    if (!*slot) {
        const hipError_t e = rk_create(r->N, r->logN, r->F.sigw, stages, slot);
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
            return fail(r, IE_ERR_OOM, "RK stage buffers do not fit in device memory");
        IE_HIP(r, e);
    }
    *w = *slot;
    rk_set_post(*w, r->in_post, r->post.omega);
    return pull_bg(r);
}

// y of an RK step: the inflation state, or the FP64 post-inflation state
static ie_status rk_fields(ie_run *r, double **f, double **fd)
{
    if (r->in_post) return dn_state(r, f, fd);
    *f = (double *)r->F.f[r->bg_host->parity];
    *fd = (double *)r->F.fd[r->bg_host->parity];
    return IE_OK;
}

static void rk_tensors(ie_run *r, float **h, float **hd)
{
    const int par = r->bg_host->parity;
    for (int c = 0; c < 6; ++c) {
        h[c] = r->F.sigw ? r->F.h[par][c] : nullptr;
        hd[c] = r->F.sigw ? r->F.hd[c] : nullptr;
    }
}

ie_status ie_rk4_step(ie_run *r, double h)
{
    if (!r || !(h > 0.0)) return fail(r, IE_ERR_ARG, "bad argument");
    RkWork *w = nullptr;
    ie_status st = rk_ready(r, 4, &w);
    if (st != IE_OK) return st;
    float *hh[6], *hd[6];
    rk_tensors(r, hh, hd);
    double *f, *fd;
    if ((st = rk_fields(r, &f, &fd)) != IE_OK) return st;
    IE_HIP(r, rk4_step(w, r->P, r->bg, r->bg_host, f, fd, hh, hd, h, r->stream));
    const DevBg advanced = *r->bg_host;
    if ((st = check_kernel_error(r)) != IE_OK) return st;
    DevBg &b = *r->bg_host;
    b.t = advanced.t;
    b.a = advanced.a;
    b.ad = advanced.ad;
    b.steps += 1;
    b.d = h;
    return push_bg(r);
}

ie_status ie_rk45_step(ie_run *r, double h, double hmax, int *accepted, double *h_next)
{
    if (!r || !accepted || !h_next || !(h > 0.0)) return fail(r, IE_ERR_ARG, "bad argument");
    RkWork *w = nullptr;
    ie_status st = rk_ready(r, 7, &w);
    if (st != IE_OK) return st;
    float *hh[6], *hd[6];
    rk_tensors(r, hh, hd);
    double *f, *fd;
    if ((st = rk_fields(r, &f, &fd)) != IE_OK) return st;
    const ie_config &c = r->cfg;
    double factor = 1.0, norm = 0.0;
    int acc = 0;
    IE_HIP(r, rk45_step(w, r->P, r->bg, r->bg_host, f, fd, hh, hd, h, c.rk45_abs_tol, c.rk45_rel_tol,
                        c.rk45_safety, &acc, &factor, &norm, r->stream));
    const DevBg advanced = *r->bg_host;
    if ((st = check_kernel_error(r)) != IE_OK) return st;
    if (!isfinite(norm)) return fail(r, IE_ERR_NONFINITE, "RK45 error norm is not finite");
    (void)factor;
    acc = rk45_propose(c, h, hmax, norm, h_next);
    *accepted = acc;
    if (!acc) return IE_OK;
    DevBg &b = *r->bg_host;
    b.t = advanced.t;
    b.a = advanced.a;
    b.ad = advanced.ad;
    b.steps += 1;
    b.d = h;
    return push_bg(r);
}

}  // extern "C"

// accessors for the other translation units of the library
namespace ie {
ie_status run_internal(ie_run *r, Fields **F, Pot **P, DevBg **bg, DevBg **bgh, double **partials,
                       double **scratch, FftPlan **fft, hipStream_t *s)
{
    *F = &r->F;
    *P = &r->P;
    *bg = r->bg;
    *bgh = r->bg_host;
    *partials = r->partials;
    *scratch = r->scratch;
    *fft = r->fft;
    *s = r->stream;
    return IE_OK;
}
ie_status run_fail(ie_run *r, ie_status st, const std::string &msg) { return fail(r, st, msg); }
const ie_config &run_config(const ie_run *r) { return r->cfg; }
ie_status run_pull_bg(ie_run *r) { return pull_bg(r); }
ie_status run_push_bg(ie_run *r) { return push_bg(r); }
}  // namespace ie
