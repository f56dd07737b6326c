// ie_init.cpp -- host side of the initial conditions (initialize.cpp surface).
//
// The k-space fill of initializef (src/initialize.cpp:118-191): every mode takes
// the next two draws of one Park-Miller stream (src/initialize.cpp:18-31), in the
// order of the reference's loops.  The stream is a multiplicative congruential
// generator, so the state before any draw is a^n x0 mod (2^31 - 1) for a known
// n: the planes i of the fill are split among host threads, each starting its
// generator at its first plane's draw (the draws per plane follow from the loop
// conditions), and every mode gets the draws it gets in the serial loop.  The
// inverse transform, the homogeneous shift and the energies that fix H0 run on
// the GPU (ie_initialize in ie_api.hip).  The generator state lives in this
// call, not in function statics, so any number of runs can be generated in one
// process (the reference is one-run-per-process, src/initialize.cpp:20-21, 56-58).
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "ie_api.h"

namespace {

// Park-Miller minimal standard generator, Schrage's factorisation, 99 discarded
// warm-up draws before the first value (src/initialize.cpp:18-31).
struct ParkMiller {
    static constexpr long long kM = 2147483647, kA = 16807;
    int seed;
    int next = 0;
    bool warmed = false;
    explicit ParkMiller(int s) : seed(s), next(s) {}
    // The generator about to make draw d + 1 (0-based count d of earlier draws):
    // the state after the 99 warm-up steps and d draws, a^(99 + d) x0 mod m (one
    // step is x -> a x mod m exactly, Schrage's factorisation being exact).
    static ParkMiller at_draw(int s, long long d)
    {
        ParkMiller g(s);
        if (s < 1) return g;
        long long x = s, e = 99 + d, b = kA, r = 1;
        for (; e > 0; e >>= 1, b = b * b % kM)
            if (e & 1) r = r * b % kM;
        g.next = (int)(r * x % kM);
        g.warmed = true;
        return g;
    }
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
    const double pi = 2.0 * asin(1.0);
    const double dp2 = sq(2.0 * pi / c.L);
    auto IX = [N](int i, int j, int k) { return (size_t)i * N * N + (size_t)j * N + (size_t)k; };
    auto NY = [N](int i, int j) { return (size_t)i * 2 * N + 2 * (size_t)j; };
    // draws of row (i, j) in the loop below: two per set_mode call
    auto row_draws = [N](int i, int j) -> long long {
        long long d = 2LL * (N / 2 - 1);
        if (j > N / 2 || (i > N / 2 && (j == 0 || j == N / 2))) d += 4;
        else if ((i == 0 || i == N / 2) && (j == 0 || j == N / 2)) d += (i == 0 && j == 0) ? 2 : 4;
        return d;
    };
    // planes i0 .. i1 - 1 of the fill, the generator started at draw d0.  The
    // mirrored plane modes a row writes belong to rows that draw none of their own
    // plane modes, so threads never write the same element.
    auto fill = [&](int i0, int i1, long long d0, bool zero) {
        if (zero) {
            const size_t a = IX(i0, 0, 0), b = IX(i1, 0, 0);
            memset(f + a, 0, (b - a) * sizeof(double));
            memset(fd + a, 0, (b - a) * sizeof(double));
            memset(fny + NY(i0, 0), 0, (size_t)(i1 - i0) * 2 * N * sizeof(double));
            memset(fdny + NY(i0, 0), 0, (size_t)(i1 - i0) * 2 * N * sizeof(double));
            return;
        }
        ModeFill M(c, *ic, H);
        M.rng = ParkMiller::at_draw(ic->seed, d0);
        for (int i = i0; i < i1; i++) {
            const double px = (i <= N / 2 ? i : i - N);
            const int iconj = (i == 0 ? 0 : N - i);
            for (int j = 0; j < N; j++) {
                const double py = (j <= N / 2 ? j : j - N);
                for (int k = 1; k < N / 2; k++) {   // interior modes, two draws each
                    const double pz = k;
                    const double p2lat = dp2 * (sq(px) + sq(py) + sq(pz));
                    const double p2 =
                        4.0 * sq(N / c.L) * (sq(sin(px * pi / N)) + sq(sin(py * pi / N)) + sq(sin(pz * pi / N)));
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
    };
    // planes split into one contiguous range per thread; each range's first draw
    // from the draws of the planes before it
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<unsigned>(std::min(hw, 16u), (unsigned)std::max(1, N / 8));
    std::vector<int> cut(nt + 1);
    std::vector<long long> start(nt + 1, 0);
    for (int t = 0; t <= nt; ++t) cut[t] = (int)((long long)N * t / nt);
    for (int t = 0; t < nt; ++t) {
        long long d = 0;
        for (int i = cut[t]; i < cut[t + 1]; ++i)
            for (int j = 0; j < N; ++j) d += row_draws(i, j);
        start[t + 1] = start[t] + d;
    }
    // zero everything first (mirrored writes cross the threads' ranges), then fill
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<std::thread> th;
        for (int t = 1; t < nt; ++t) th.emplace_back(fill, cut[t], cut[t + 1], start[t], pass == 0);
        fill(cut[0], cut[1], start[0], pass == 0);
        for (auto &x : th) x.join();
    }
    f[0] = f[1] = 0.0;   // no homogeneous fluctuation (src/initialize.cpp:188-191)
    fd[0] = fd[1] = 0.0;
    return IE_OK;
}

}  // extern "C"
