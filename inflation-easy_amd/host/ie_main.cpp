// ie_main.cpp -- drop-in driver of the gfx950 engine for InflationEasy.
//
// Restates the reference program's host side (src/main.cpp:103-146) above the C
// ABI of include/ie_api.h: the params.txt parser (src/runtime_parameters.cpp),
// the phase order (initial conditions, inflation loop, deltaN loop, post-inflation
// loop), the output cadence and re-synchronisation of save()/saveN()/
// save_post_inflation() (src/output.cpp:1469-1560,1839-1866) and the results/*.dat
// and results/post_inflation/*.dat formats.  The compile-time toggles of
// src/parameters.h are run-time keys here (N, numerical_potential, perform_deltaN,
// calculate_SIGW, post_inflation, monotonic_potential, antimonotonic_potential) plus
// `precision` (fp64 | split).
// Lattice state stays on the GPU; only means, spectra bins, histograms and the
// requested snapshots come back to the host.
//
// usage: inflation_easy_amd [params.txt]      (run in the directory holding
//        inputs/ and where results/ is written, like the reference binary)
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include <cfloat>
#include <cctype>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "ie_api.h"

namespace {

// ------------------------------------------------------------------ params

struct Params {
    int N = 128, numerical = 1, deltaN = 1, sigw = 1, precision = IE_FP64;
    int monotonic = 1, antimonotonic = 0, mono_set = 0;
    int seed = 8;
    double rescale_s = 0.0, af = 0.0, V0 = 0.0, rescale_B = 0.0, ns = 0.97;
    double initial_field = 0.0, initial_derivative = 0.0, L = 10.0, dt = 0.001;
    int integrator = IE_LEAPFROG, deltaN_integrator = IE_LEAPFROG;
    double rk45_abs_tol = 1e-8, rk45_rel_tol = 1e-6, rk45_min_dt = -1.0, rk45_max_dt = -1.0, rk45_safety = 0.9;
    int output_freq = 500, output_infrequent_freq = 500;
    double dN = 1e-4, Nend = 5.0;
    int use_phiref_manual = 0;
    double phiref_manual_value = 0.0;
    double high_cutoff_index = 0.0, low_cutoff_index = 0.0;
    int forcing_cutoff = 0;
    int output_spectra = 1, output_histogram = 1, output_energy = 1, output_box3D = 0, output_box2D = 0;
    int output_bispectrum = 0, output_LOG = 0;
    double eta_log = -0.5;
    int screen_updates = 1, nbins = 256, int_err = 5, int_errN = 5;
    // post-inflation phase (src/parameters.h:31, src/runtime_parameters.cpp:48-52,79)
    int post = 1, post_integrator = IE_LEAPFROG;
    double horizon_factor = 1.0, omega = 1.0 / 3.0, dt_post = 0.001, af_post = 0.0;
};

std::string trim(const std::string &s)
{
    size_t b = 0, e = s.size();
    while (b < e && isspace((unsigned char)s[b])) ++b;
    while (e > b && isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}

bool parse_int(const std::string &s, int &out)
{
    char *end = nullptr;
    errno = 0;
    const long v = strtol(s.c_str(), &end, 10);
    if (errno || !end || *end) return false;
    out = (int)v;
    return true;
}

bool parse_double(const std::string &s, double &out)
{
    char *end = nullptr;
    errno = 0;
    const double v = strtod(s.c_str(), &end);
    if (errno || !end || *end) return false;
    out = v;
    return true;
}

// integrator names and numbers (src/runtime_parameters.cpp:141-166)
bool parse_integrator(std::string s, int &out)
{
    for (auto &ch : s) ch = (char)tolower((unsigned char)ch);
    if (s == "leapfrog" || s == "lf") return out = IE_LEAPFROG, true;
    if (s == "rk4") return out = IE_RK4, true;
    if (s == "rk45" || s == "rkf45" || s == "dopri5") return out = IE_RK45, true;
    int iv;
    if (parse_int(s, iv) && iv >= 0 && iv <= 2) return out = iv, true;
    return false;
}

std::vector<std::pair<int, std::pair<std::string, std::string>>> read_lines(const char *path, bool *found)
{
    std::vector<std::pair<int, std::pair<std::string, std::string>>> kv;
    std::ifstream in(path);
    *found = in.good();
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const auto hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        line = trim(line);
        if (line.empty()) continue;
        const auto eq = line.find('=');
        if (eq == std::string::npos) {
            fprintf(stderr, "Ignoring malformed line %d in %s: %s\n", lineno, path, line.c_str());
            continue;
        }
        kv.push_back({lineno, {trim(line.substr(0, eq)), trim(line.substr(eq + 1))}});
    }
    return kv;
}

// load_runtime_parameters (src/runtime_parameters.cpp:195-365); the defaults of
// src/runtime_parameters.cpp:17-112 depend on the potential kind, so the toggles
// are read first.
Params load_params(const char *path)
{
    Params p;
    bool found = false;
    auto kv = read_lines(path, &found);
    for (auto &e : kv) {
        int iv;
        if (e.second.first == "numerical_potential" && parse_int(e.second.second, iv)) p.numerical = iv != 0;
        if (e.second.first == "N" && parse_int(e.second.second, iv)) p.N = iv;
    }
    if (p.numerical) {
        p.V0 = 3e-9;
        p.initial_field = 2.9181235049318586;
        p.initial_derivative = -0.06727651095116181;
        p.dt = 0.001;
        p.output_freq = p.output_infrequent_freq = 500;
        p.dN = 0.0001;
        p.Nend = 5.0;
    } else {
        p.V0 = 3.338e-13;
        p.initial_field = 0.0935;
        p.initial_derivative = 0.000796;
        p.dt = 0.0005;
        p.output_freq = p.output_infrequent_freq = 200;
        p.dN = 0.0000001;
        p.Nend = 0.001;
    }
    bool B_set = false, af_set = false, min_set = false, max_set = false, af_post_set = false, post_keys = false;
    for (auto &e : kv) {
        const int lineno = e.first;
        const std::string &key = e.second.first, &val = e.second.second;
        int iv = 0;
        double dv = 0.0;
        if (key == "N" || key == "numerical_potential") continue;
        if (key == "seed" && parse_int(val, iv)) p.seed = iv;
        else if (key == "rescale_s" && parse_double(val, dv)) p.rescale_s = dv;
        else if (key == "af" && parse_double(val, dv)) p.af = dv, af_set = true;
        else if (key == "V0" && parse_double(val, dv)) p.V0 = dv;
        else if (key == "rescale_B" && parse_double(val, dv)) p.rescale_B = dv, B_set = true;
        else if (key == "ns" && !p.numerical && parse_double(val, dv)) p.ns = dv;
        else if (key == "initial_field" && parse_double(val, dv)) p.initial_field = dv;
        else if (key == "initial_derivative" && parse_double(val, dv)) p.initial_derivative = dv;
        else if (key == "L" && parse_double(val, dv)) p.L = dv;
        else if (key == "dt" && parse_double(val, dv)) p.dt = dv;
        else if (key == "inflation_integrator") {
            if (!parse_integrator(val, p.integrator))
                fprintf(stderr, "Ignoring invalid inflation_integrator '%s' on line %d.\n", val.c_str(), lineno);
        } else if (key == "deltaN_integrator") {
            if (!parse_integrator(val, p.deltaN_integrator))
                fprintf(stderr, "Ignoring invalid deltaN_integrator '%s' on line %d.\n", val.c_str(), lineno);
        } else if (key == "rk45_abs_tol" && parse_double(val, dv)) p.rk45_abs_tol = dv;
        else if (key == "rk45_rel_tol" && parse_double(val, dv)) p.rk45_rel_tol = dv;
        else if (key == "rk45_min_dt" && parse_double(val, dv)) p.rk45_min_dt = dv, min_set = true;
        else if (key == "rk45_max_dt" && parse_double(val, dv)) p.rk45_max_dt = dv, max_set = true;
        else if (key == "rk45_safety" && parse_double(val, dv)) p.rk45_safety = dv;
        else if (key == "output_freq" && parse_int(val, iv)) p.output_freq = iv;
        else if (key == "output_infrequent_freq" && parse_int(val, iv)) p.output_infrequent_freq = iv;
        else if (key == "dN" && parse_double(val, dv)) p.dN = dv;
        else if (key == "Nend" && parse_double(val, dv)) p.Nend = dv;
        else if (key == "use_phiref_manual" && parse_int(val, iv)) p.use_phiref_manual = iv != 0;
        else if (key == "phiref_manual_value" && parse_double(val, dv)) p.phiref_manual_value = dv;
        else if (key == "high_cutoff_index" && parse_double(val, dv)) p.high_cutoff_index = dv;
        else if (key == "low_cutoff_index" && parse_double(val, dv)) p.low_cutoff_index = dv;
        else if (key == "forcing_cutoff" && parse_int(val, iv)) p.forcing_cutoff = iv != 0;
        else if (key == "output_spectra" && parse_int(val, iv)) p.output_spectra = iv != 0;
        else if (key == "output_histogram" && parse_int(val, iv)) p.output_histogram = iv != 0;
        else if (key == "output_energy" && parse_int(val, iv)) p.output_energy = iv != 0;
        else if (key == "output_box3D" && parse_int(val, iv)) p.output_box3D = iv != 0;
        else if (key == "output_box2D" && parse_int(val, iv)) p.output_box2D = iv != 0;
        else if (key == "output_bispectrum" && parse_int(val, iv)) p.output_bispectrum = iv != 0;
        else if (key == "output_LOG" && parse_int(val, iv)) p.output_LOG = iv != 0;
        else if (key == "eta_log" && parse_double(val, dv)) p.eta_log = dv;
        else if (key == "screen_updates" && parse_int(val, iv)) p.screen_updates = iv != 0;
        else if (key == "nbins" && parse_int(val, iv)) p.nbins = iv;
        else if (key == "int_err" && parse_int(val, iv)) p.int_err = iv > 1 ? iv : 1;
        else if (key == "int_errN" && parse_int(val, iv)) p.int_errN = iv > 1 ? iv : 1;
        // former compile-time toggles of src/parameters.h, run-time here
        else if (key == "perform_deltaN" && parse_int(val, iv)) p.deltaN = iv != 0;
        else if (key == "calculate_SIGW" && parse_int(val, iv)) p.sigw = iv != 0;
        else if (key == "monotonic_potential" && parse_int(val, iv)) p.monotonic = iv != 0, p.mono_set = 1;
        else if (key == "antimonotonic_potential" && parse_int(val, iv)) p.antimonotonic = iv != 0, p.mono_set = 1;
        else if (key == "precision") {
            if (val == "split") p.precision = IE_SPLIT;
            else if (val == "fp64") p.precision = IE_FP64;
            else fprintf(stderr, "Ignoring invalid precision '%s' on line %d.\n", val.c_str(), lineno);
        } else if (key == "post_inflation" && parse_int(val, iv)) p.post = iv != 0;
        else if (key == "horizon_factor" && parse_double(val, dv)) p.horizon_factor = dv, post_keys = true;
        else if (key == "omega" && parse_double(val, dv)) p.omega = dv, post_keys = true;
        else if (key == "dt_post_inflation" && parse_double(val, dv)) p.dt_post = dv, post_keys = true;
        else if (key == "af_post_inflation" && parse_double(val, dv)) p.af_post = dv, af_post_set = true, post_keys = true;
        else if (key == "post_inflation_integrator") {
            post_keys = true;
            if (!parse_integrator(val, p.post_integrator))
                fprintf(stderr, "Ignoring invalid post_inflation_integrator '%s' on line %d.\n", val.c_str(), lineno);
        } else if (key == "parallel_calculation") {
            // OpenMP toggle of the reference build: nothing to select here
        } else {
            fprintf(stderr, "Ignoring unknown or invalid parameter '%s' on line %d.\n", key.c_str(), lineno);
        }
    }
    if (!p.mono_set) {   // src/parameters.h:60-66 auto-switch
        p.monotonic = p.numerical ? 1 : 0;
        p.antimonotonic = p.numerical ? 0 : 1;
    }
    if (!B_set) p.rescale_B = sqrt(p.V0);
    if (!af_set) p.af = 2.0 * p.N;
    if (!af_post_set) p.af_post = 2.0 * p.N;
    if (!p.post && post_keys)
        fprintf(stderr, "Ignoring post-inflation runtime parameters in %s because post_inflation=0.\n", path);
    if (!min_set || p.rk45_min_dt <= 0.0) p.rk45_min_dt = fabs(p.dt) * 1e-6;
    if (!max_set || p.rk45_max_dt <= 0.0) p.rk45_max_dt = fabs(p.dt);
    (void)found;
    return p;
}

// --------------------------------------------------------------- outputs

struct Out {
    std::map<std::string, FILE *> files;
    FILE *get(const char *name)
    {
        auto it = files.find(name);
        if (it != files.end()) return it->second;
        char path[256];
        snprintf(path, sizeof(path), "results/%s.dat", name);
        FILE *fp = fopen(path, "w");
        if (!fp) {
            fprintf(stderr, "Failed to open output file: %s\n", path);
            exit(1);
        }
        files[name] = fp;
        return fp;
    }
    ~Out()
    {
        for (auto &e : files) fclose(e.second);
    }
};

void die(ie_run *run, ie_status st, const char *what)
{
    fprintf(stderr, "%s failed (%d): %s\n", what, (int)st, ie_last_error(run));
    exit(1);
}

#define CHECK(run, call)                    \
    do {                                    \
        ie_status s_ = (call);              \
        if (s_ != IE_OK) die(run, s_, #call); \
    } while (0)

struct Driver {
    Params p;
    ie_config cfg{};
    ie_run *run = nullptr;
    Out out;
    FILE *log = nullptr;
    int N = 0;
    size_t gs = 0;
    double dx = 0.0;
    time_t t_start = 0;
    std::vector<double> tx, tv, tdv;

    double pw2(double x) const { return x * x; }

    ie_background bg()
    {
        ie_background b;
        CHECK(run, ie_get_background(run, &b));
        return b;
    }

    void info_begin()
    {
        FILE *fp = out.get("info");
        fprintf(fp, "--------------------------\n");
        fprintf(fp, "General Program Information\n");
        fprintf(fp, "-----------------------------\n");
        fprintf(fp, "Grid size=%d^%d\n", N, 3);
        fprintf(fp, "L=%f\n", p.L);
        fprintf(fp, "f0=%f\n", p.initial_field);
        fprintf(fp, "fd0=%f\n", p.initial_derivative);
        fprintf(fp, "dt=%f, dt/dx=%f\n", p.dt, p.dt / dx);
        const char *names[] = {"leapfrog", "rk4", "rk45"};
        fprintf(fp, "inflation_integrator=%s\n", names[p.integrator]);
        if (p.deltaN) fprintf(fp, "deltaN_integrator=%s\n", names[p.deltaN_integrator]);
        fprintf(fp, "rescale_s=%f\n", p.rescale_s);
        fprintf(fp, "rescale_B=%e\n", p.rescale_B);
        fprintf(fp, "precision=%s\n", p.precision == IE_SPLIT ? "split" : "fp64");
        time(&t_start);
        fprintf(fp, "\nRun began at %s", ctime(&t_start));
        fflush(fp);
    }

    void info_end(double t0, double t)
    {
        FILE *fp = out.get("info");
        time_t fin;
        time(&fin);
        fprintf(fp, "Run ended at %s", ctime(&fin));
        fprintf(fp, "\nRun from t=%f to t=%f took %d seconds\n\n", t0, t, (int)(fin - t_start));
        fflush(fp);
    }

    void spectra_block(FILE *fp, const std::vector<double> &pp, const std::vector<int> &c, const std::vector<double> &v)
    {
        for (size_t b = 0; b < pp.size(); ++b) fprintf(fp, "%e %d %e\n", pp[b], c[b], v[b]);
        fprintf(fp, "\n");
        fflush(fp);
    }

    void histogram(int which, const char *hist, const char *times, const ie_background &b, bool times_both)
    {
        double lo, hi;
        std::vector<double> counts(p.nbins > 0 ? p.nbins : 1);
        CHECK(run, ie_histogram(run, which, p.nbins, &lo, &hi, counts.data()));
        FILE *ft = out.get(times);
        if (times_both) fprintf(ft, "%f %e", b.t, b.a);
        else {
            fprintf(ft, "%f", b.t);
            fprintf(ft, " %e", b.a);
        }
        double df = (hi - lo) / (double)p.nbins;
        if (!std::isfinite(df) || df <= 0.0) df = 1.0;
        double numpts = 0;
        for (double c : counts) numpts += c;
        if (numpts == 0) numpts = 1;
        FILE *fh = out.get(hist);
        for (int q = 0; q < p.nbins; ++q) fprintf(fh, "%e\n", counts[q] / numpts);
        fprintf(fh, "\n");
        fflush(fh);
        fprintf(ft, " %e %e\n", lo, df);
        fflush(ft);
    }

    // save(infrequent) of src/output.cpp:1469-1503; `half` is the half step of the
    // leapfrog re-synchronisation (0 at t = 0).
    void save(int infrequent, double half)
    {
        if (half != 0.0) CHECK(run, ie_evolve_fields(run, -half));
        ie_background b = bg();
        double sums[3];
        CHECK(run, ie_field_sums(run, sums));
        const double av = sums[0] / (double)gs;
        double vel = sums[1] / (double)gs;
        vel = vel * pow(b.a, p.rescale_s - 1.0) * p.rescale_B;
        FILE *m = out.get("means"), *v = out.get("variance"), *w = out.get("velocity");
        fprintf(m, "%f %e %e", b.t, b.a, av);
        fprintf(w, "%f %e %e", b.t, b.a, vel);
        fprintf(v, "%f %e %e", b.t, b.a, sums[2] - pw2(av));
        if (av + DBL_MAX == av || (av != 0. && av / av != 1.)) {   // src/output.cpp:251-258
            printf("Unstable solution developed. Scalar field not numerical at t=%f\n", b.t);
            fflush(m);
            fflush(v);
            exit(1);
        }
        fprintf(m, "\n");
        fprintf(v, "\n");
        fprintf(w, "\n");
        if (infrequent) {
            fflush(m);
            fflush(v);
            fflush(w);
        }
        FILE *sf = out.get("sf");
        fprintf(sf, "%f %f %e %e\n", b.t, b.a, b.ad * p.rescale_B * pow(b.a, p.rescale_s - 2.0),
                pw2(p.rescale_B) * pow(b.a, 2.0 * p.rescale_s - 2.0) * (b.ad2 + (p.rescale_s - 1.0) * pw2(b.ad) / b.a));
        if (infrequent) fflush(sf);
        if (infrequent) {
            if (p.output_box3D || p.output_box2D) snapshots();
            if (p.output_energy) energy(b);
            if (p.output_spectra) {
                const int nb = ie_spectrum_numbins(N);
                std::vector<double> pp(nb), vv(nb);
                std::vector<int> cc(nb);
                const double norm1 = pow(p.L / p.rescale_B, 3) / pow((double)N, 6);
                CHECK(run, ie_spectrum(run, IE_FIELD_F, norm1, pp.data(), cc.data(), vv.data()));
                spectra_block(out.get("spectra"), pp, cc, vv);
                FILE *st = out.get("spectratimes");
                fprintf(st, "%f %e\n", b.t, b.a);
                fflush(st);
                forcing_cutoff();
                if (p.sigw) {
                    const int ng = (int)(sqrt(3.0) * (N / 2)) + 1;
                    std::vector<double> pg(ng), vg(ng);
                    std::vector<int> cg(ng);
                    CHECK(run, ie_gw_spectrum(run, 0, pg.data(), cg.data(), vg.data()));
                    spectra_block(out.get("spectraGW"), pg, cg, vg);
                    CHECK(run, ie_gw_spectrum(run, 1, pg.data(), cg.data(), vg.data()));
                    spectra_block(out.get("spectraGWdot"), pg, cg, vg);
                }
            }
            if (p.output_histogram) histogram(IE_FIELD_F, "histogram", "histogramtimes", b, false);
        }
        if (half != 0.0) CHECK(run, ie_evolve_fields(run, half));
    }

    // the optional k-space cutoff at the end of spectraf (src/output.cpp:333-404)
    void forcing_cutoff()
    {
        if (p.high_cutoff_index > 0 && p.forcing_cutoff)
            CHECK(run, ie_forcing_cutoff(run, p.high_cutoff_index, p.low_cutoff_index));
    }

    void energy(const ie_background &b)
    {
        double kin, grad, pot;
        CHECK(run, ie_kin_energy(run, &kin));
        CHECK(run, ie_gradient_energy(run, &grad));
        CHECK(run, ie_potential_energy(run, &pot));
        double total = 0.;
        total += kin;
        total += grad;
        total += pot;
        FILE *e = out.get("energy"), *c = out.get("conservation");
        fprintf(e, "%f %e %e %e %e\n", b.t, b.a, kin * p.rescale_B, grad * p.rescale_B, pot * p.rescale_B);
        fflush(e);
        fprintf(c, "%e %e %e\n", b.t, b.a, 3.0 * pow(b.a, 2.0 * p.rescale_s - 4.0) * pw2(b.ad) / (total));
        fflush(c);
    }

    // box (src/output.cpp:913-931) downloads the lattice; box2d / box2dot
    // (:933-981) one plane each
    void snapshots()
    {
        if (p.output_box3D) {
            std::vector<double> f(gs);
            CHECK(run, ie_download_field(run, IE_FIELD_F, f.data()));
            FILE *fp = out.get("box");
            for (size_t id = 0; id < gs; ++id) fprintf(fp, "%.17g\n", f[id]);
            fprintf(fp, "\n");
            fflush(fp);
        }
        if (p.output_box2D) {
            const int i = N < 64 ? 5 : 50;
            std::vector<double> f((size_t)N * N), fd((size_t)N * N);
            CHECK(run, ie_download_plane(run, IE_FIELD_F, i, f.data()));
            CHECK(run, ie_download_plane(run, IE_FIELD_FD, i, fd.data()));
            FILE *a = out.get("snapshots_2d_phi"), *b = out.get("snapshots_2d_phidot");
            for (size_t q = 0; q < (size_t)N * N; ++q) fprintf(a, "%.17g\n", f[q]);
            fprintf(a, "\n");
            for (size_t q = 0; q < (size_t)N * N; ++q) fprintf(b, "%.17g\n", fd[q] * p.rescale_B);
            fprintf(b, "\n");
            fflush(a);
            fflush(b);
        }
    }

    // RK45 step clamp and cap (src/evolution.cpp:330-345)
    double clamp_rk45(double h, double hmax) const
    {
        const double hmin = p.rk45_min_dt > 1e-16 ? p.rk45_min_dt : 1e-16;
        const double hhi = hmax > hmin ? hmax : hmin;
        return h < hmin ? hmin : (h > hhi ? hhi : h);
    }
    double rk45_hmax(double base) const
    {
        const double lo = p.rk45_min_dt > 2.0 * base ? p.rk45_min_dt : 2.0 * base;
        return p.rk45_max_dt < lo ? p.rk45_max_dt : lo;
    }
    // accept/reject driver (src/evolution.cpp:346-362): trial steps until one is
    // accepted; returns the proposed next step
    double rk45_drive(double h, double hmax, bool deltaN_phase)
    {
        int accepted = 0, attempts = 0;
        while (!accepted) {
            double hs = h;
            if (deltaN_phase) CHECK(run, ie_deltaN_rk45_step(run, h, hmax, &accepted, &hs));
            else CHECK(run, ie_rk45_step(run, h, hmax, &accepted, &hs));
            h = hs;
            ++attempts;
            if (!accepted && (attempts > 25 || h <= p.rk45_min_dt * (1.0 + 1e-12))) {
                fprintf(stderr, "RK45 failed to converge (attempts=%d, h=%e)\n", attempts, h);
                exit(1);
            }
        }
        return h;
    }

    void progress(long long n, double a)
    {
        if (log) {
            fprintf(log, "scale factor a = %f\n", a);
            fprintf(log, "numsteps %lld\n\n", n);
            fflush(log);
        }
        if (p.screen_updates && n % (p.output_freq > 0 ? p.output_freq : 1) == 0) {
            printf("scale factor a = %f\n", a);
            printf("numsteps %lld\n\n", n);
        }
    }

    // run_evolution_loop, RK4 and RK45 branches (src/evolution.cpp:947-989): one
    // host-driven step at a time, dt astep^{s-1} from the previous step's a
    void inflation_rk()
    {
        long long n = 0;
        auto step_base = [&]() { return p.dt * pow(bg().astep, p.rescale_s - 1.0); };
        double h = clamp_rk45(step_base(), p.rk45_max_dt);
        for (;;) {
            ie_background b = bg();
            if (!(b.a <= p.af)) break;
            if (p.integrator == IE_RK4) {
                CHECK(run, ie_rk4_step(run, step_base()));
            } else {
                const double hmax = rk45_hmax(step_base());
                h = clamp_rk45(h, hmax);
                h = rk45_drive(h, hmax, false);
            }
            ++n;
            b = bg();
            progress(n, b.a);
            if (n % p.output_freq == 0 && b.a < p.af) save((n % p.output_infrequent_freq == 0) ? 1 : 0, 0.0);
            b.astep = b.a;
            CHECK(run, ie_set_background(run, &b));
        }
        printf("Saving final inflaton data\n");
        save(1, 0.0);
        save_modes();
    }

    void save_modes()
    {
        // save_last: modes.dat (src/output.cpp:1505-1519)
        const int nb = (int)(sqrt(3.0) * (N / 2)) + 1;
        std::vector<double> md(nb);
        CHECK(run, ie_modes(N, p.L, p.rescale_B, md.data()));
        FILE *fp = out.get("modes");
        for (double x : md) fprintf(fp, "%e\n", x);
        fprintf(fp, "\n");
        fflush(fp);
        if (p.output_bispectrum) {   // bispectraf (src/output.cpp:740-911)
            std::vector<double> bp(nb), re(nb), im(nb);
            std::vector<int> bc(nb);
            CHECK(run, ie_bispectrum(run, bp.data(), bc.data(), re.data(), im.data()));
            FILE *bf = out.get("bispectra");
            for (int k = 0; k < nb; ++k) fprintf(bf, "%e %d %e %e\n", bp[k], bc[k], re[k], im[k]);
            fprintf(bf, "\n");
            fflush(bf);
        }
        if (p.deltaN && p.output_LOG) {   // spectraLOG, histogramsLOG (src/output.cpp:1280-1403)
            const int nbins = p.nbins > 0 ? p.nbins : 1;
            std::vector<double> freq(nbins);
            double bmin = 0.0, df = 1.0;
            int valid = 0;
            CHECK(run, ie_deltaN_log(run, p.eta_log, nbins, &bmin, &df, freq.data(), &valid));
            const int ns = ie_spectrum_numbins(N);
            std::vector<double> pp(ns), vv(ns);
            std::vector<int> cc(ns);
            const double norm1 = pow(p.L / p.rescale_B, 3) / pow((double)N, 6);
            CHECK(run, ie_spectrum(run, IE_FIELD_DELTAN, norm1, pp.data(), cc.data(), vv.data()));
            spectra_block(out.get("spectraLOG"), pp, cc, vv);
            ie_background b = bg();
            FILE *ht = out.get("histogramtimesLOG"), *h = out.get("histogramLOG");
            fprintf(ht, "%f %e", b.t, b.a);
            for (int q = 0; q < p.nbins; ++q) fprintf(h, "%e\n", freq[q]);
            fprintf(h, "\n");
            fflush(h);
            fprintf(ht, " %e %e\n", bmin, df);
            fflush(ht);
        }
    }

    // run_evolution_loop, leapfrog branch (src/evolution.cpp:909-946,990-997)
    void inflation()
    {
        if (p.integrator != IE_LEAPFROG) return inflation_rk();
        CHECK(run, ie_evolve_fields(run, 0.5 * p.dt));
        long long n = 0;
        const int chunk = p.output_freq > 0 ? p.output_freq : 1;
        for (;;) {
            ie_background b;
            CHECK(run, ie_leapfrog_steps(run, chunk, &b));
            const long long taken = b.steps - n;
            n = b.steps;
            if (taken <= 0) break;
            progress(n, b.a);
            if (n % p.output_freq == 0 && b.a < p.af) {
                // re-synchronise by half the step just taken, dt astep_old^{s-1}
                // (src/output.cpp:1471-1472: astep is not yet updated there)
                save((n % p.output_infrequent_freq == 0) ? 1 : 0, 0.5 * b.dt_last);
            }
            if (b.done) break;
        }
        printf("Saving final inflaton data\n");
        ie_background b = bg();
        const double half = 0.5 * p.dt * pow(b.astep, p.rescale_s - 1.0);
        save(1, half);
        CHECK(run, ie_evolve_fields(run, -half));
        save_modes();
    }

    // run_deltaN_loop (src/evolution.cpp:1291-1375) with saveN; the leapfrog is
    // staggered by half a step, RK4 and RK45 are not
    void deltaN()
    {
        const bool staggered = p.deltaN_integrator == IE_LEAPFROG;
        printf("Starting deltaN calculation\n");
        if (log) fprintf(log, "Starting deltaN calculation\n");
        CHECK(run, ie_deltaN_init(run));
        ie_background b = bg();
        b.phiref = 0.0;
        b.Ne = 0.0;
        CHECK(run, ie_set_background(run, &b));
        if (staggered) CHECK(run, ie_evolve_fieldsN(run, 0.5 * p.dN));
        double phiref = p.phiref_manual_value;
        if (!p.use_phiref_manual) CHECK(run, ie_get_phiref(run, &phiref));
        b = bg();
        b.phiref = phiref;
        CHECK(run, ie_set_background(run, &b));
        double Ne = 0.0;
        if (p.deltaN_integrator == IE_RK45) {
            double h = clamp_rk45(p.dN, p.rk45_max_dt);
            while (Ne <= p.Nend) {
                const double hmax = rk45_hmax(p.dN);
                h = clamp_rk45(h, hmax);
                h = rk45_drive(h, hmax, true);
                Ne += h;   // the proposed next step, as the reference counts
            }
        } else {
            long long nsteps = 0;
            while (Ne <= p.Nend) {   // the step count of the reference loop, exactly
                Ne += p.dN;
                ++nsteps;
            }
            if (staggered) CHECK(run, ie_deltaN_steps(run, p.dN, nsteps));
            else CHECK(run, ie_deltaN_rk4_steps(run, p.dN, nsteps));
        }
        if (log) fprintf(log, "N = %f\n\n", Ne);
        // saveN (src/output.cpp:1522-1560)
        if (staggered) CHECK(run, ie_evolve_fieldsN(run, -.5 * p.dN));
        const double cutoff = p.Nend - (staggered ? p.dN : 0.0);
        float nmean = 0.f;
        CHECK(run, ie_deltaN_zeta(run, cutoff, &nmean));
        b = bg();
        histogram(IE_FIELD_DELTAN, "histogramN", "histogramtimesN", b, true);
        if (p.output_box2D) {   // box2dN (src/output.cpp:1195-1218)
            std::vector<double> dn((size_t)N * N);
            CHECK(run, ie_download_plane(run, IE_FIELD_DELTAN, N < 64 ? 5 : 50, dn.data()));
            FILE *fp = out.get("snapshots_2d_deltaN");
            for (double x : dn) fprintf(fp, "%.17g\n", x);
            fprintf(fp, "\n");
            fflush(fp);
        }
        CHECK(run, ie_deltaN_zero_unfinished(run, cutoff));
        if (p.output_spectra) {
            CHECK(run, ie_deltaN_center(run));
            const int nb = ie_spectrum_numbins(N);
            std::vector<double> pp(nb), vv(nb);
            std::vector<int> cc(nb);
            const double norm1 = pow(p.L / p.rescale_B, 3) / pow((double)N, 6);
            CHECK(run, ie_spectrum(run, IE_FIELD_DELTAN, norm1, pp.data(), cc.data(), vv.data()));
            spectra_block(out.get("spectraN"), pp, cc, vv);
        }
        if (staggered) {
            CHECK(run, ie_evolve_fieldsN(run, .5 * p.dN));
            CHECK(run, ie_evolve_fieldsN(run, -0.5 * p.dN));
        }
    }

    // save_post_inflation (src/output.cpp:1839-1866): meansvars_post_inflation,
    // scale_post_inflation, spectraf_post_inflation, the tensor spectra and
    // histograms_post_inflation into results/post_inflation/
    void save_post(int infrequent)
    {
        const bool staggered = p.post_integrator == IE_LEAPFROG;
        if (staggered && bg().t > 0.) CHECK(run, ie_evolve_fields(run, -.5 * p.dt_post));
        ie_background b = bg();
        double sums[3];
        CHECK(run, ie_field_sums(run, sums));
        const double av = sums[0] / (double)gs;
        double vel = sums[1] / (double)gs;
        vel = vel * pow(b.a, p.rescale_s - 1) * p.rescale_B;
        FILE *m = out.get("post_inflation/means"), *v = out.get("post_inflation/variance"),
             *w = out.get("post_inflation/velocity");
        fprintf(m, "%f %e %e", b.t, b.a, av);
        fprintf(w, "%f %e %e", b.t, b.a, vel);
        fprintf(v, "%f %e %e", b.t, b.a, sums[2] - pw2(av));
        if (av + FLT_MAX == av || (av != 0. && av / av != 1.)) {   // src/output.cpp:1621-1628
            printf("Unstable solution developed. Scalar field not numerical at t=%f\n", b.t);
            fflush(m);
            fflush(v);
            exit(1);
        }
        fprintf(m, "\n");
        fprintf(v, "\n");
        fprintf(w, "\n");
        if (infrequent) {
            fflush(m);
            fflush(v);
            fflush(w);
        }
        FILE *sf = out.get("post_inflation/sf");
        fprintf(sf, "%f %f %e %e\n", b.t, b.a, b.ad * p.rescale_B * pow(b.a, p.rescale_s - 2.),
                pw2(p.rescale_B) * pow(b.a, 2. * p.rescale_s - 2.) * (b.ad2 + (p.rescale_s - 1.) * pw2(b.ad) / b.a));
        if (infrequent) fflush(sf);
        if (infrequent) {
            if (p.output_spectra) {
                const int nb = ie_spectrum_numbins(N);
                std::vector<double> pp(nb), vv(nb);
                std::vector<int> cc(nb);
                const double norm1 = pow(p.L / p.rescale_B, 3) / pow((double)N, 6);
                CHECK(run, ie_spectrum(run, IE_FIELD_F, norm1, pp.data(), cc.data(), vv.data()));
                spectra_block(out.get("post_inflation/spectra"), pp, cc, vv);
                FILE *st = out.get("post_inflation/spectratimes");
                fprintf(st, "%f %e\n", b.t, b.a);
                fflush(st);
                forcing_cutoff();
                const int ng = (int)(sqrt(3.0) * (N / 2)) + 1;
                std::vector<double> pg(ng), vg(ng);
                std::vector<int> cg(ng);
                CHECK(run, ie_gw_spectrum(run, 0, pg.data(), cg.data(), vg.data()));
                spectra_block(out.get("post_inflation/spectraGW"), pg, cg, vg);
                CHECK(run, ie_gw_spectrum(run, 1, pg.data(), cg.data(), vg.data()));
                spectra_block(out.get("post_inflation/spectraGWdot"), pg, cg, vg);
            }
            if (p.output_histogram)
                histogram(IE_FIELD_F, "post_inflation/histogram", "post_inflation/histogramtimes", b, false);
        }
        if (staggered && bg().t > 0.) CHECK(run, ie_evolve_fields(run, .5 * p.dt_post));
    }

    // run_post_inflation_loop (src/evolution.cpp:1568-1651): the leapfrog on the
    // device in chunks of output_freq steps, RK4 / RK45 one host-driven step at a time
    void post_inflation()
    {
        ie_post_params pp{p.horizon_factor, p.omega, p.dt_post, p.af_post, p.post_integrator, p.deltaN};
        CHECK(run, ie_post_init(run, &pp, nullptr));
        save_post(1);
        long long n = 0;
        auto report = [&](long long step, double a) {
            if (step % p.output_freq == 0) save_post((step % p.output_infrequent_freq == 0) ? 1 : 0);
            progress(step, a);
        };
        if (p.post_integrator == IE_LEAPFROG) {
            CHECK(run, ie_evolve_fields(run, 0.5 * p.dt_post));
            const int chunk = p.output_freq > 0 ? p.output_freq : 1;
            for (;;) {
                ie_background b;
                CHECK(run, ie_post_leapfrog_steps(run, chunk, &b));
                const long long taken = b.steps - n;
                n = b.steps;
                if (taken <= 0) break;
                report(n, b.a);
                if (b.done) break;
            }
        } else if (p.post_integrator == IE_RK4) {
            while (bg().a <= p.af_post) {
                CHECK(run, ie_rk4_step(run, p.dt_post));
                ++n;
                report(n, bg().a);
            }
        } else {
            double h = clamp_rk45(p.dt_post, p.rk45_max_dt);
            while (bg().a <= p.af_post) {
                const double hmax = rk45_hmax(p.dt_post);
                h = clamp_rk45(h, hmax);
                h = rk45_drive(h, hmax, false);
                ++n;
                report(n, bg().a);
            }
        }
        printf("Saving final inflaton data\n");
        save_post(1);
    }

    int main(const char *params_path)
    {
        p = load_params(params_path);
        if (p.seed < 1)
            printf("Warning: The parameter seed has been set to %d, which will result in incorrect output.\n", p.seed);
        struct stat info;
        if (stat("results", &info) != 0) {
            if (mkdir("results", 0777) != 0) {
                fprintf(stderr, "Failed to create 'results' directory.\n");
                return 1;
            }
        } else if (!(info.st_mode & S_IFDIR)) {
            fprintf(stderr, "'results' exists but is not a directory.\n");
            return 1;
        }
        if (p.post && !p.sigw) {   // src/parameters.h:33-35
            fprintf(stderr, "post_inflation requires calculate_SIGW = 1\n");
            return 1;
        }
        if (p.post) {   // ensure_post_inflation_directory (src/main.cpp:68-83)
            if (stat("results/post_inflation", &info) != 0) {
                if (mkdir("results/post_inflation", 0777) != 0) {
                    fprintf(stderr, "Failed to create 'results/post_inflation' directory.\n");
                    return 1;
                }
            } else if (!(info.st_mode & S_IFDIR)) {
                fprintf(stderr, "'results/post_inflation' exists but is not a directory.\n");
                return 1;
            }
        }
        log = fopen("results/output.txt", "w");
        if (!log) {
            fprintf(stderr, "Failed to open output file.\n");
            return 1;
        }
        N = p.N;
        gs = (size_t)N * N * N;
        dx = p.L / (double)N;
        cfg.N = N;
        cfg.precision = p.precision;
        cfg.numerical_potential = p.numerical;
        cfg.calculate_sigw = p.sigw;
        cfg.monotonic = p.monotonic ? IE_MONOTONIC : (p.antimonotonic ? IE_ANTIMONOTONIC : IE_GENERIC);
        cfg.integrator = p.integrator;
        cfg.deltaN_integrator = p.deltaN_integrator;
        cfg.rescale_s = p.rescale_s;
        cfg.af = p.af;
        cfg.V0 = p.V0;
        cfg.rescale_B = p.rescale_B;
        cfg.ns = p.ns;
        cfg.L = p.L;
        cfg.dt = p.dt;
        cfg.rk45_abs_tol = p.rk45_abs_tol;
        cfg.rk45_rel_tol = p.rk45_rel_tol;
        cfg.rk45_min_dt = p.rk45_min_dt;
        cfg.rk45_max_dt = p.rk45_max_dt;
        cfg.rk45_safety = p.rk45_safety;
        cfg.dN = p.dN;
        cfg.Nend = p.Nend;
        cfg.use_phiref_manual = p.use_phiref_manual;
        cfg.phiref_manual_value = p.phiref_manual_value;
        if (p.numerical) {   // load_numerical_inputs (src/main.cpp:46-50,85-101)
            auto load = [](const char *path, std::vector<double> &v) {
                std::ifstream in(path);
                if (!in) {
                    fprintf(stderr, "Could not open file: %s\n", path);
                    exit(1);
                }
                double x;
                while (in >> x) v.push_back(x);
                if (v.empty()) {
                    fprintf(stderr, "File appears to be empty or malformed: %s\n", path);
                    exit(1);
                }
            };
            load("inputs/field_values.dat", tx);
            load("inputs/potential.dat", tv);
            load("inputs/potential_derivative.dat", tdv);
            cfg.tab_n = (int)tx.size();
            cfg.tab_x = tx.data();
            cfg.tab_v = tv.data();
            cfg.tab_dv = tdv.data();
        }
        ie_status st = ie_create(&cfg, &run);
        if (st != IE_OK) {
            fprintf(stderr, "%s\n", ie_last_error(nullptr));
            return 1;
        }
        printf("Generating initial conditions for new run at t = 0\n");
        ie_ic_params ic{p.seed, p.initial_field, p.initial_derivative, p.high_cutoff_index, p.low_cutoff_index};
        ie_background b;
        st = ie_initialize(run, &ic, &b);
        if (st != IE_OK) {
            fprintf(stderr, "%s\n", ie_last_error(run));
            return 1;
        }
        printf("Finished initial conditions\n");
        info_begin();
        save(1, 0.0);
        inflation();
        if (p.deltaN) deltaN();
        if (p.post) post_inflation();
        b = bg();
        info_end(0.0, b.t);
        fprintf(log, "InflationEasy program finished\n");
        printf("InflationEasy program finished\n");
        fclose(log);
        ie_destroy(run);
        return 0;
    }
};

}  // namespace

int main(int argc, char **argv)
{
    Driver d;
    return d.main(argc > 1 ? argv[1] : "params.txt");
}
