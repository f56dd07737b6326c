// ie_init.cpp -- host side of the initial conditions (initialize.cpp surface).
//
// The k-space fill of initializef (src/initialize.cpp:118-191) is inherently
// serial: every mode takes the next two draws of one Park-Miller stream
// (src/initialize.cpp:18-31), in the order of the reference's loops.  It runs on
// the host; the inverse transform, the homogeneous shift and the energies that
// fix H0 run on the GPU (ie_initialize in ie_api.hip).  The generator state lives
// in this call, not in function statics, so any number of runs can be generated
// in one process (the reference is one-run-per-process, src/initialize.cpp:20-21,
// 56-58).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "ie_api.h"

namespace {

// Park-Miller minimal standard generator, Schrage's factorisation, 99 discarded
// warm-up draws before the first value (src/initialize.cpp:18-31).
struct ParkMiller {
    int seed;
    int next = 0;
    bool warmed = false;
    explicit ParkMiller(int s) : seed(s), next(s) {}
    double uniform()
    {
        if (seed < 1) return 0.33;
        if (!(next > 0)) next = 1;
        if (!warmed) {
            warmed = true;
            for (int q = 1; q < 100; ++q) step();
        }
        step();
        return (double)next / 2147483647.0;
    }
    void step()
    {
        next = 16807 * (next % 127773) - 2836 * (next / 127773);
        if (next < 0) next += 2147483647;
    }
};

double sq(double x) { return x * x; }

struct ModeFill {
    const ie_config &c;
    const ie_ic_params &ic;
    ParkMiller rng;
    double pi, dx, norm, hbterm, L;
    bool warned = false;

    ModeFill(const ie_config &cfg, const ie_ic_params &p, double hubble)
        : c(cfg), ic(p), rng(p.seed)
    {
        pi = 2.0 * asin(1.0);
        L = cfg.L;
        dx = cfg.L / (double)cfg.N;
        norm = cfg.rescale_B * pow(L / sq(dx), 1.5);   // src/initialize.cpp:56
        hbterm = -hubble;                             // src/initialize.cpp:58
    }

    // set_mode (src/initialize.cpp:53-89), massless: omega = sqrt(p2)
    void set_mode(double p2, double p2lat, double *field, double *deriv, int real)
    {
        const double m2 = 0.0;
        const double ii = L / (2.0 * pi) * sqrt(p2lat);
        const double omega = (p2 + m2 > 0.0) ? sqrt(p2 + m2) : sqrt(p2);
        if (p2 + m2 <= 0.0 && !warned) {
            printf("Warning: Tachyonic mode(s) may be initialized inaccurately\n");
            warned = true;
        }
        double rms = (omega > 0.0) ? norm / sqrt(2.0 * omega) : 0.0;
        if (ic.high_cutoff_index > 0 && (ii > ic.high_cutoff_index || ii < ic.low_cutoff_index)) rms = 0.0;
        const double amplitude = rms * sqrt(log(1.0 / rng.uniform()));
        const double phase = 2.0 * pi * rng.uniform();
        const double re = amplitude * cos(phase), im = amplitude * sin(phase);
        field[0] = re;
        field[1] = im;
        deriv[0] = omega * im + hbterm * field[0];
        deriv[1] = -omega * re + hbterm * field[1];
        if (real == 1) {
            field[1] = 0.0;
            deriv[1] = 0.0;
        }
    }
};

double tab_interp(const ie_config &c, const double *y, double v, int l)
{
    const double x0 = c.tab_x[l - 1], x1 = c.tab_x[l], y0 = y[l - 1], y1 = y[l];
    return y1 + (v - x1) * (y0 - y1) / (x0 - x1);
}

}  // namespace

extern "C" {

ie_status ie_potential_value(const ie_config *c, double v, double *V, double *dV)
{
    if (!c || !V || !dV) return IE_ERR_ARG;
    const double B2 = c->rescale_B * c->rescale_B;
    if (!c->numerical_potential) {
        *V = c->V0 * (1.0 - (1.0 - c->ns) / 4.0 * sq(v)) / B2;
        *dV = -c->V0 * (1.0 - c->ns) / 2.0 * v / B2;
        return IE_OK;
    }
    const int n = c->tab_n;
    if (n < 2 || !c->tab_x || !c->tab_v || !c->tab_dv) return IE_ERR_ARG;
    if (v > c->tab_x[0] || v < c->tab_x[n - 1]) return IE_ERR_TABLE_RANGE;
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        if (c->tab_x[mid] >= v) lo = mid;
        else hi = mid;
    }
    *V = tab_interp(*c, c->tab_v, v, hi) / B2;
    *dV = tab_interp(*c, c->tab_dv, v, hi) / B2;
    return IE_OK;
}

ie_status ie_fill_modes(const ie_config *cfg, const ie_ic_params *ic, double *f, double *fny, double *fd,
                        double *fdny, double *hubble_init)
{
    if (!cfg || !ic || !f || !fny || !fd || !fdny || !hubble_init) return IE_ERR_ARG;
    const ie_config &c = *cfg;
    const int N = c.N;
    if (N < 4 || (N & (N - 1))) return IE_ERR_ARG;
    const double dx = c.L / (double)N;
    if (c.dt > dx / sqrt(3.0)) return IE_ERR_COURANT;   // src/initialize.cpp:94-97
    double V0, dV0;
    ie_status st = ie_potential_value(cfg, ic->initial_field, &V0, &dV0);
    if (st != IE_OK) return st;
    // src/initialize.cpp:135
    const double H = sqrt((0.5 * sq(ic->initial_derivative) + V0) / 3.0);
    *hubble_init = H;
    const size_t gs = (size_t)N * N * N;
    memset(f, 0, gs * sizeof(double));
    memset(fd, 0, gs * sizeof(double));
    memset(fny, 0, (size_t)N * 2 * N * sizeof(double));
    memset(fdny, 0, (size_t)N * 2 * N * sizeof(double));
    ModeFill M(c, *ic, H);
    const double pi = M.pi;
    const double dp2 = sq(2.0 * pi / c.L);
    auto IX = [N](int i, int j, int k) { return (size_t)i * N * N + (size_t)j * N + (size_t)k; };
    auto NY = [N](int i, int j) { return (size_t)i * 2 * N + 2 * (size_t)j; };
    for (int i = 0; i < N; i++) {
        const double px = (i <= N / 2 ? i : i - N);
        const int iconj = (i == 0 ? 0 : N - i);
        for (int j = 0; j < N; j++) {
            const double py = (j <= N / 2 ? j : j - N);
            for (int k = 1; k < N / 2; k++) {   // interior modes, two draws each
                const double pz = k;
                const double p2lat = dp2 * (sq(px) + sq(py) + sq(pz));
                const double p2 = 4.0 * sq(N / c.L) * (sq(sin(px * pi / N)) + sq(sin(py * pi / N)) + sq(sin(pz * pi / N)));
                M.set_mode(p2, p2lat, &f[IX(i, j, 2 * k)], &fd[IX(i, j, 2 * k)], 0);
            }
            if (j > N / 2 || (i > N / 2 && (j == 0 || j == N / 2))) {
                // kz = 0 and kz = N/2 planes: one of each conjugate pair drawn, the
                // partner mirrored (src/initialize.cpp:156-175)
                const int jconj = (j == 0 ? 0 : N - j);
                double p2lat = dp2 * (sq(px) + sq(py));
                double p2 = 4.0 * sq(N / c.L) * (sq(sin(px * pi / N)) + sq(sin(py * pi / N)));
                M.set_mode(p2, p2lat, &f[IX(i, j, 0)], &fd[IX(i, j, 0)], 0);
                f[IX(iconj, jconj, 0)] = f[IX(i, j, 0)];
                f[IX(iconj, jconj, 1)] = -f[IX(i, j, 1)];
                fd[IX(iconj, jconj, 0)] = fd[IX(i, j, 0)];
                fd[IX(iconj, jconj, 1)] = -fd[IX(i, j, 1)];
                p2lat = dp2 * (sq(px) + sq(py) + sq(N / 2));
                p2 = 4.0 * sq(N / c.L) * (sq(sin(px * pi / N)) + sq(sin(py * pi / N)) + 1.0);
                M.set_mode(p2, p2lat, &fny[NY(i, j)], &fdny[NY(i, j)], 0);
                fny[NY(iconj, jconj)] = fny[NY(i, j)];
                fny[NY(iconj, jconj) + 1] = -fny[NY(i, j) + 1];
                fdny[NY(iconj, jconj)] = fdny[NY(i, j)];
                fdny[NY(iconj, jconj) + 1] = -fdny[NY(i, j) + 1];
            } else if ((i == 0 || i == N / 2) && (j == 0 || j == N / 2)) {
                // self-conjugate modes are real (src/initialize.cpp:176-184)
                double p2lat = dp2 * (sq(px) + sq(py));
                double p2 = 4.0 * sq(N / c.L) * (sq(sin(px * pi / N)) + sq(sin(py * pi / N)));
                if (p2 > 0.0) M.set_mode(p2, p2lat, &f[IX(i, j, 0)], &fd[IX(i, j, 0)], 1);
                p2lat = dp2 * (sq(px) + sq(py) + sq(N / 2));
                p2 = 4.0 * sq(N / c.L) * (sq(sin(px * pi / N)) + sq(sin(py * pi / N)) + 1.0);
                M.set_mode(p2, p2lat, &fny[NY(i, j)], &fdny[NY(i, j)], 1);
            }
        }
    }
    f[0] = f[1] = 0.0;   // no homogeneous fluctuation (src/initialize.cpp:188-191)
    fd[0] = fd[1] = 0.0;
    return IE_OK;
}

}  // extern "C"
