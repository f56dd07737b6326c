/*
 * ie_oracle.c -- CPU restatement of the InflationEasy hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and CPU baseline for the gfx950
 * engine (see ie_oracle.h).  Never linked into the engine.
 *
 * Every function names the reference lines it restates.  The reference's
 * floating-point expression order is kept (gcc, x86-64, no FMA contraction) so the
 * results equal the reference build's; the per-site loops carry OpenMP pragmas like
 * the reference (results do not depend on the thread count), while the lattice
 * reductions stay serial and ordered as in the reference.
 */
#include "ie_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IEO_R double
#define IEO_SFX d
#define IEO_ASIN asin
#define IEO_COS cos
#define IEO_SIN sin
#include "ie_oracle_fft.h"
#undef IEO_R
#undef IEO_SFX
#undef IEO_ASIN
#undef IEO_COS
#undef IEO_SIN
#define IEO_R float
#define IEO_SFX f
#define IEO_ASIN asinf
#define IEO_COS cosf
#define IEO_SIN sinf
#include "ie_oracle_fft.h"
#undef IEO_R
#undef IEO_SFX
#undef IEO_ASIN
#undef IEO_COS
#undef IEO_SIN

static double sq(double x) { return x * x; }

struct ieo_state {
    ieo_config cfg;
    int N;
    size_t gs;
    double dx, pi;
    double t, t0, a, ad, ad2, astep, Ne, phiref, hubble_init;
    double phiref_potential;
    int post_rhs;              /* RK stages use the post-inflation right-hand side */
    double *f, *fd, *dn;
    double *fny, *fdny;
    float *h[6], *hd[6];
    float *hny[6], *hdny[6];
    int *lstart;
    /* rand_uniform statics (src/initialize.cpp:20-21) */
    int rng_i, rng_next, rng_seeded;
    /* set_mode statics (src/initialize.cpp:56-59) */
    int sm_ready, sm_tachyonic;
    double sm_norm, sm_hbterm;
};

int ieo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ config */

void ieo_config_defaults(ieo_config *c, int N, int numerical)
{
    memset(c, 0, sizeof(*c));
    c->N = N;
    c->numerical_potential = numerical;
    c->calculate_sigw = 1;
    c->perform_deltaN = 1;
    c->monotonic = numerical ? IEO_MONOTONIC : IEO_ANTIMONOTONIC;
    c->seed = 8;
    c->rescale_s = 0.0;
    c->af = 2.0 * N;
    c->V0 = numerical ? 3e-9 : 3.338e-13;
    c->rescale_B = sqrt(c->V0);
    c->ns = 0.97;
    if (numerical) {
        c->initial_field = 2.9181235049318586;
        c->initial_derivative = -0.06727651095116181;
        c->dt = 0.001;
        c->output_freq = 500;
        c->output_infrequent_freq = 500;
        c->dN = 0.0001;
        c->Nend = 5.0;
    } else {
        c->initial_field = 0.0935;
        c->initial_derivative = 0.000796;
        c->dt = 0.0005;
        c->output_freq = 200;
        c->output_infrequent_freq = 200;
        c->dN = 0.0000001;
        c->Nend = 0.001;
    }
    c->L = 10.0;
    c->integrator = IEO_LEAPFROG;
    c->deltaN_integrator = IEO_LEAPFROG;
    c->rk45_abs_tol = 1e-8;
    c->rk45_rel_tol = 1e-6;
    c->rk45_min_dt = -1.0;
    c->rk45_max_dt = -1.0;
    c->rk45_safety = 0.9;
    c->output_spectra = 1;
    c->output_histogram = 1;
    c->output_energy = 1;
    c->nbins = 256;
    c->int_err = 5;
    c->int_errN = 5;
    c->post_inflation = 0;
    c->horizon_factor = 1.0;
    c->omega = 1.0 / 3.0;
    c->dt_post = 0.001;
    c->af_post = 2.0 * N;
    c->post_integrator = IEO_LEAPFROG;
    c->output_bispectrum = 0;
    c->output_LOG = 0;
    c->eta_log = -0.5;
}

/* src/runtime_parameters.cpp:342-353 and :184-192 */
void ieo_config_finalize(ieo_config *c, int rescale_B_set, int af_set)
{
    if (!rescale_B_set) c->rescale_B = sqrt(c->V0);
    if (!af_set) c->af = 2.0 * c->N;
    if (!(c->rk45_min_dt > 0.0)) c->rk45_min_dt = fabs(c->dt) * 1e-6;
    if (!(c->rk45_max_dt > 0.0)) c->rk45_max_dt = fabs(c->dt);
    if (c->rk45_min_dt <= 0.0) c->rk45_min_dt = 1e-16;
    if (c->rk45_max_dt < c->rk45_min_dt) c->rk45_max_dt = c->rk45_min_dt;
    if (!(c->rk45_abs_tol > 0.0)) c->rk45_abs_tol = 1e-8;
    if (!(c->rk45_rel_tol > 0.0)) c->rk45_rel_tol = 1e-6;
    if (!(c->rk45_safety > 0.0 && c->rk45_safety < 1.0)) c->rk45_safety = 0.9;
}

/* ------------------------------------------------------------------- state */

ieo_state *ieo_create(const ieo_config *cfg)
{
    ieo_state *st = (ieo_state *)calloc(1, sizeof(*st));
    if (!st) return NULL;
    st->cfg = *cfg;
    st->N = cfg->N;
    st->gs = (size_t)cfg->N * cfg->N * cfg->N;
    st->dx = cfg->L / (double)cfg->N;
    st->pi = 2.0 * asin(1.0);
    st->a = 1.0;
    st->astep = 1.0;
    const size_t gs = st->gs, ny = (size_t)cfg->N * 2 * cfg->N;
    st->f = (double *)calloc(gs, sizeof(double));
    st->fd = (double *)calloc(gs, sizeof(double));
    st->dn = (double *)calloc(gs, sizeof(double));
    st->fny = (double *)calloc(ny, sizeof(double));
    st->fdny = (double *)calloc(ny, sizeof(double));
    st->lstart = (int *)calloc(gs, sizeof(int));
    int ok = st->f && st->fd && st->dn && st->fny && st->fdny && st->lstart;
    for (int c = 0; c < 6; ++c) {
        st->h[c] = (float *)calloc(gs, sizeof(float));
        st->hd[c] = (float *)calloc(gs, sizeof(float));
        st->hny[c] = (float *)calloc(ny, sizeof(float));
        st->hdny[c] = (float *)calloc(ny, sizeof(float));
        ok = ok && st->h[c] && st->hd[c] && st->hny[c] && st->hdny[c];
    }
    if (!ok) { ieo_destroy(st); return NULL; }
    return st;
}

void ieo_destroy(ieo_state *st)
{
    if (!st) return;
    free(st->f); free(st->fd); free(st->dn); free(st->fny); free(st->fdny); free(st->lstart);
    for (int c = 0; c < 6; ++c) { free(st->h[c]); free(st->hd[c]); free(st->hny[c]); free(st->hdny[c]); }
    free(st);
}

double *ieo_field(ieo_state *st, int which)
{
    switch (which) {
    case 0: return st->f;
    case 1: return st->fd;
    case 2: return st->dn;
    case 3: return st->fny;
    case 4: return st->fdny;
    default: return NULL;
    }
}

float *ieo_tensor(ieo_state *st, int comp)
{
    if (comp >= 0 && comp < 6) return st->h[comp];
    if (comp >= 6 && comp < 12) return st->hd[comp - 6];
    return NULL;
}

int *ieo_lstart(ieo_state *st) { return st->lstart; }

void ieo_get_background(const ieo_state *st, double o[8])
{
    o[0] = st->t; o[1] = st->a; o[2] = st->ad; o[3] = st->ad2;
    o[4] = st->astep; o[5] = st->Ne; o[6] = st->phiref; o[7] = st->hubble_init;
}

void ieo_set_background(ieo_state *st, const double in[8])
{
    st->t = in[0]; st->a = in[1]; st->ad = in[2]; st->ad2 = in[3];
    st->astep = in[4]; st->Ne = in[5]; st->phiref = in[6]; st->hubble_init = in[7];
}

static inline size_t IX(int N, int i, int j, int k)
{
    return (size_t)i * N * N + (size_t)j * N + (size_t)k;
}

/* --------------------------------------------------------------------- RNG */

/* Park-Miller minimal standard generator with Schrage factorisation and 99
 * discarded warm-up draws (src/initialize.cpp:18-31).  State is per run here. */
double ieo_rand_uniform(ieo_state *st)
{
    const int A = 16807, M = 2147483647, Q = 127773, R = 2836;
    if (st->cfg.seed < 1) return 0.33;
    if (!st->rng_seeded) { st->rng_next = st->cfg.seed; st->rng_seeded = 1; }
    if (!(st->rng_next > 0)) {
        printf("Invalid seed used in random number function. Using seed=1\n");
        st->rng_next = 1;
    }
    if (st->rng_i == 0)
        for (st->rng_i = 1; st->rng_i < 100; st->rng_i++) ieo_rand_uniform(st);
    st->rng_next = A * (st->rng_next % Q) - R * (st->rng_next / Q);
    if (st->rng_next < 0) st->rng_next += M;
    return (double)st->rng_next / (double)M;
}

/* --------------------------------------------------------------- potential */

/* Bracket search over the descending table: returns l with x[l-1] >= v > x[l].
 * Binary search for start <= 1, hinted walk otherwise (src/potential.cpp:11-50).
 * *err: 0 ok, 1 table too small, 2 out of range, 3 walk out of bounds. */
int ieo_find_bracketing_index(const ieo_state *st, double v, int start, int *err)
{
    const int n = st->cfg.tab_n;
    const double *x = st->cfg.tab_x;
    *err = 0;
    if (n < 2) { *err = 1; return 1; }
    if (v > x[0] || v < x[n - 1]) { *err = 2; return 1; }
    if (start <= 1) {
        int lo = 0, hi = n - 1;
        while (hi - lo > 1) {
            const int mid = lo + (hi - lo) / 2;
            if (x[mid] >= v) lo = mid;
            else hi = mid;
        }
        return hi;
    }
    int l = start < 1 ? 1 : (start > n - 1 ? n - 1 : start);
    while (l < n && x[l] >= v) ++l;
    while (l > 1 && x[l - 1] < v) --l;
    if (l <= 0 || l >= n) { *err = 3; return 1; }
    return l;
}

static void ieo_die_range(double v, const ieo_state *st, int err)
{
    if (err == 2)
        fprintf(stderr, "Field value %.17g is outside the numerical potential range [%.17g, %.17g].\n",
                v, st->cfg.tab_x[st->cfg.tab_n - 1], st->cfg.tab_x[0]);
    else if (err == 3)
        fprintf(stderr, "Interpolation index out of bounds for field value %.17g.\n", v);
    else
        fprintf(stderr, "Numerical potential table must contain at least two points.\n");
    exit(1);
}

/* Linear interpolation between l-1 and l (src/potential.cpp:52-68). */
static double interp(const ieo_state *st, double v, int l, const double *y)
{
    const double x0 = st->cfg.tab_x[l - 1], x1 = st->cfg.tab_x[l];
    const double y0 = y[l - 1], y1 = y[l];
    if (x0 == x1) {
        fprintf(stderr, "Interpolation failed due to repeated field grid points.\n");
        exit(1);
    }
    return y1 + (v - x1) * (y0 - y1) / (x0 - x1);
}

/* Analytic hilltop V/B^2 and V'/B^2 (src/potential.cpp:77-83). */
static double hill_V(const ieo_config *c, double v)
{
    return c->V0 * (1.0 - (1.0 - c->ns) / 4.0 * sq(v)) / sq(c->rescale_B);
}
static double hill_dV(const ieo_config *c, double v)
{
    return -c->V0 * (1.0 - c->ns) / 2.0 * v / sq(c->rescale_B);
}

static int bracket_or_die(const ieo_state *st, double v, int start)
{
    int err;
    const int l = ieo_find_bracketing_index(st, v, start, &err);
    if (err) ieo_die_range(v, st, err);
    return l;
}

/* potential(field) (src/potential.cpp:91-98) */
double ieo_potential(const ieo_state *st, double v)
{
    if (st->cfg.numerical_potential) {
        const int l = bracket_or_die(st, v, 1);
        return interp(st, v, l, st->cfg.tab_v) / sq(st->cfg.rescale_B);
    }
    return hill_V(&st->cfg, v);
}

/* potential_derivative_from_value (src/potential.cpp:114-121) */
double ieo_potential_derivative_value(const ieo_state *st, double v)
{
    if (st->cfg.numerical_potential) {
        const int l = bracket_or_die(st, v, 1);
        return interp(st, v, l, st->cfg.tab_dv) / sq(st->cfg.rescale_B);
    }
    return hill_dV(&st->cfg, v);
}

/* potential_derivative(i,j,k) with the lstart hint (src/potential.cpp:103-112) */
static double pot_deriv_site(ieo_state *st, size_t id)
{
    if (st->cfg.numerical_potential) {
        const int l = bracket_or_die(st, st->f[id], st->lstart[id]);
        const int nh = l - st->cfg.int_err;
        st->lstart[id] = nh > 1 ? nh : 1;
        return interp(st, st->f[id], l, st->cfg.tab_dv) / sq(st->cfg.rescale_B);
    }
    return hill_dV(&st->cfg, st->f[id]);
}

/* evaluate_potential_from_value (src/potential.cpp:123-146) */
static void eval_pot(const ieo_state *st, double v, int hint, int lookback, int *next_hint,
                     double *pot, double *dpot)
{
    if (st->cfg.numerical_potential) {
        const int l = bracket_or_die(st, v, hint);
        *pot = interp(st, v, l, st->cfg.tab_v) / sq(st->cfg.rescale_B);
        *dpot = interp(st, v, l, st->cfg.tab_dv) / sq(st->cfg.rescale_B);
        if (next_hint) {
            const int back = lookback > 1 ? lookback : 1;
            *next_hint = (l - back) > 1 ? (l - back) : 1;
        }
    } else {
        *pot = hill_V(&st->cfg, v);
        *dpot = hill_dV(&st->cfg, v);
        if (next_hint) *next_hint = 1;
    }
}

/* pot_ratio(i,j,k) (src/potential.cpp:177-192) */
static double pot_ratio_site(ieo_state *st, size_t id)
{
    double pot, pot_deriv;
    if (st->cfg.numerical_potential) {
        const int l = bracket_or_die(st, st->f[id], st->lstart[id]);
        pot = interp(st, st->f[id], l, st->cfg.tab_v);
        pot_deriv = interp(st, st->f[id], l, st->cfg.tab_dv);
        const int nh = l - st->cfg.int_errN;
        st->lstart[id] = nh > 1 ? nh : 1;
    } else {
        pot = hill_V(&st->cfg, st->f[id]);
        pot_deriv = hill_dV(&st->cfg, st->f[id]);
    }
    return pot_deriv / pot;
}

/* ---------------------------------------------------------------- stencils */

/* 7-point periodic Laplacian, un-normalised (src/evolution.cpp:24-51). */
static inline double lapl_d(int N, const double *g, int i, int j, int k)
{
    const int ip = i == N - 1 ? 0 : i + 1, im = i == 0 ? N - 1 : i - 1;
    const int jp = j == N - 1 ? 0 : j + 1, jm = j == 0 ? N - 1 : j - 1;
    const int kp = k == N - 1 ? 0 : k + 1, km = k == 0 ? N - 1 : k - 1;
    return g[IX(N, i, j, kp)] + g[IX(N, i, j, km)] + g[IX(N, i, jp, k)] + g[IX(N, i, jm, k)]
         + g[IX(N, ip, j, k)] + g[IX(N, im, j, k)] - 6.0 * g[IX(N, i, j, k)];
}

static inline float lapl_f(int N, const float *g, int i, int j, int k)
{
    const int ip = i == N - 1 ? 0 : i + 1, im = i == 0 ? N - 1 : i - 1;
    const int jp = j == N - 1 ? 0 : j + 1, jm = j == 0 ? N - 1 : j - 1;
    const int kp = k == N - 1 ? 0 : k + 1, km = k == 0 ? N - 1 : k - 1;
    return g[IX(N, i, j, kp)] + g[IX(N, i, j, km)] + g[IX(N, i, jp, k)] + g[IX(N, i, jm, k)]
         + g[IX(N, ip, j, k)] + g[IX(N, im, j, k)] - 6.0f * g[IX(N, i, j, k)];
}

/* Central first derivative of phi / (2 dx) (src/evolution.cpp:55-77). */
static inline void grad_site(int N, double dx, const double *g, int i, int j, int k, double out[3])
{
    const double h = 0.5 / dx;
    const int ip = i == N - 1 ? 0 : i + 1, im = i == 0 ? N - 1 : i - 1;
    const int jp = j == N - 1 ? 0 : j + 1, jm = j == 0 ? N - 1 : j - 1;
    const int kp = k == N - 1 ? 0 : k + 1, km = k == 0 ? N - 1 : k - 1;
    out[0] = (g[IX(N, ip, j, k)] - g[IX(N, im, j, k)]) * h;
    out[1] = (g[IX(N, i, jp, k)] - g[IX(N, i, jm, k)]) * h;
    out[2] = (g[IX(N, i, j, kp)] - g[IX(N, i, j, km)]) * h;
}

/* ---------------------------------------------------------------- energies */

/* src/evolution.cpp:130-136 */
static double gradient_energy_of(const ieo_state *st, const double *g, double a)
{
    const int N = st->N;
    double s = 0.0;
    const double norm = sq(1.0 / (a * st->dx));
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) s -= g[IX(N, i, j, k)] * lapl_d(N, g, i, j, k);
    return 0.5 * s * norm / (double)st->gs;
}

double ieo_gradient_energy(ieo_state *st) { return gradient_energy_of(st, st->f, st->a); }

/* src/evolution.cpp:139-145 */
double ieo_kin_energy(ieo_state *st)
{
    double s = 0.0;
    for (size_t id = 0; id < st->gs; ++id) s += sq(st->fd[id]);
    s /= (double)st->gs;
    return 0.5 * pow(st->a, 2.0 * st->cfg.rescale_s - 2.0) * s;
}

/* src/potential.cpp:151-171 */
double ieo_potential_energy(ieo_state *st)
{
    double pot = 0.0;
    if (st->cfg.numerical_potential) {
        for (size_t id = 0; id < st->gs; ++id) {
            const int l = bracket_or_die(st, st->f[id], st->lstart[id]);
            pot += interp(st, st->f[id], l, st->cfg.tab_v) / sq(st->cfg.rescale_B);
            const int nh = l - st->cfg.int_err;
            st->lstart[id] = nh > 1 ? nh : 1;
        }
    } else {
        for (size_t id = 0; id < st->gs; ++id) pot += ieo_potential(st, st->f[id]);
    }
    return pot / (double)st->gs;
}

/* ---------------------------------------------------------------- leapfrog */

/* Kick: Friedmann half-updates of ad around the phi (and tensor) kicks
 * (src/evolution.cpp:150-257). */
void ieo_evolve_derivs(ieo_state *st, double d)
{
    const int N = st->N;
    const double s = st->cfg.rescale_s;
    const double laplnorm = pow(st->a, -2.0 * s) / sq(st->dx);
    const double sfev1 = s + 1.0;
    const double sfev2 = -2.0 * s + 2.0;
    const double a = st->a, ad = st->ad;

    st->ad2 = (-2.0 * ad - 2.0 * a / d / sfev1 * (
        1.0 - sqrt(1.0 + 2.0 * d * sfev1 * ad / a +
        sq(d) * sfev1 * pow(a, sfev2) *
        (2.0 * ieo_gradient_energy(st) / 3.0 + ieo_potential_energy(st))))) / d;
    st->ad += 0.5 * d * st->ad2;

    const double friction = (2.0 + s) * st->ad / st->a;
    const double potnorm = pow(st->a, 2.0 - 2.0 * s);

#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                const size_t id = IX(N, i, j, k);
                st->fd[id] += d * (laplnorm * lapl_d(N, st->f, i, j, k)
                                   - friction * st->fd[id]
                                   - potnorm * pot_deriv_site(st, id));
            }

    if (st->cfg.calculate_sigw) {
        const double srcAmp = 2.0 * pow(st->a, -2.0 * s);
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                for (int k = 0; k < N; ++k) {
                    const size_t id = IX(N, i, j, k);
                    double g[3];
                    grad_site(N, st->dx, st->f, i, j, k, g);
                    const double T[6] = {g[0] * g[0], g[1] * g[1], g[2] * g[2],
                                         g[0] * g[1], g[0] * g[2], g[1] * g[2]};
                    for (int c = 0; c < 6; ++c) {
                        const double lap_h = (double)lapl_f(N, st->h[c], i, j, k);
                        const double rhs = d * (laplnorm * lap_h
                                                - friction * (double)st->hd[c][id]
                                                + srcAmp * T[c]);
                        st->hd[c][id] += (float)rhs;
                    }
                }
    }
    st->ad += 0.5 * d * st->ad2;
}

/* Drift (src/evolution.cpp:260-284). */
void ieo_evolve_fields(ieo_state *st, double d)
{
    st->t += d;
    const size_t gs = st->gs;
#pragma omp parallel for schedule(static)
    for (size_t id = 0; id < gs; ++id) st->f[id] += d * st->fd[id];
    if (st->cfg.calculate_sigw) {
        for (int c = 0; c < 6; ++c) {
            float *h = st->h[c];
            const float *hd = st->hd[c];
#pragma omp parallel for schedule(static)
            for (size_t id = 0; id < gs; ++id) h[id] += (float)(d * hd[id]);
        }
    }
    st->a += d * st->ad;
}

/* ---------------------------------------------------------- post-inflation */

static inline int wrap_up(int N, int x) { return x == N - 1 ? 0 : x + 1; }
static inline int wrap_dn(int N, int x) { return x == 0 ? N - 1 : x - 1; }

/* d2_same_state (src/evolution.cpp:79-91): second difference along dim. */
static inline double d2_same(int N, double dx, const double *g, int dim, int i, int j, int k)
{
    const double inv_dx2 = 1.0 / (dx * dx);
    const double c = g[IX(N, i, j, k)];
    if (dim == 0) return (g[IX(N, wrap_up(N, i), j, k)] - 2.0 * c + g[IX(N, wrap_dn(N, i), j, k)]) * inv_dx2;
    if (dim == 1) return (g[IX(N, i, wrap_up(N, j), k)] - 2.0 * c + g[IX(N, i, wrap_dn(N, j), k)]) * inv_dx2;
    return (g[IX(N, i, j, wrap_up(N, k))] - 2.0 * c + g[IX(N, i, j, wrap_dn(N, k))]) * inv_dx2;
}

/* d2_cross_state (src/evolution.cpp:93-106).  Each lattice direction named by d1
 * or d2 is stepped "+" in the first two samples' first slot and "-" in the other
 * two, with every named direction moving together: (x, y) and (x, z) are the usual
 * four-corner cross differences, while for (y, z) the samples are f(i, j+1, k+1),
 * f(i, j-1, k-1), f(i, j+1, k+1), f(i, j-1, k-1), whose combination
 * ((A - B) - A) + B is the rounding residue of A - B; it is kept as written. */
static inline double d2_cross(int N, double dx, const double *g, int d1, int d2, int i, int j, int k)
{
    const int xi = d1 == 0 || d2 == 0, yj = d1 == 1 || d2 == 1, zk = d1 == 2 || d2 == 2;
    const int ip = xi ? wrap_up(N, i) : i, im = xi ? wrap_dn(N, i) : i;
    const int jp = yj ? wrap_up(N, j) : j, jm = yj ? wrap_dn(N, j) : j;
    const int kp = zk ? wrap_up(N, k) : k, km = zk ? wrap_dn(N, k) : k;
    const double fpp = g[IX(N, ip, jp, kp)];
    const double fpm = g[IX(N, ip, jm, km)];
    const double fmp = g[IX(N, im, jp, kp)];
    const double fmm = g[IX(N, im, jm, km)];
    return (fpp - fpm - fmp + fmm) / (4.0 * dx * dx);
}

/* Bare tensor source of one site (src/evolution.cpp:1474-1503, 549-570): gradients
 * of Phi and Phi', the Hessian of Phi, in the order xx yy zz xy xz yz. */
static void post_source(int N, double dx, const double *f, const double *fd, int i, int j, int k,
                        double invH, double coeffU, double S[6])
{
    double g[3], gd[3];
    grad_site(N, dx, f, i, j, k, g);
    grad_site(N, dx, fd, i, j, k, gd);
    const double H[6] = {d2_same(N, dx, f, 0, i, j, k), d2_same(N, dx, f, 1, i, j, k),
                         d2_same(N, dx, f, 2, i, j, k), d2_cross(N, dx, f, 0, 1, i, j, k),
                         d2_cross(N, dx, f, 0, 2, i, j, k), d2_cross(N, dx, f, 1, 2, i, j, k)};
    const double Phi = f[IX(N, i, j, k)];
    const double du[3] = {gd[0] * invH + g[0], gd[1] * invH + g[1], gd[2] * invH + g[2]};
    static const int L[6] = {0, 1, 2, 0, 0, 1}, M[6] = {0, 1, 2, 1, 2, 2};
    for (int c = 0; c < 6; ++c)
        S[c] = 4.0 * Phi * H[c] + 2.0 * g[L[c]] * g[M[c]] - coeffU * (du[L[c]] * du[M[c]]);
}

/* initialize_post_inflation (src/initialize.cpp:246-291) without its save: Phi
 * from the deltaN map (or, without deltaN, from f and the mean velocity), fd kept
 * as the previous phase left it, tensors zeroed, t = 0, a = 1,
 * ad = horizon_factor N 2 pi / L. */
void ieo_post_initialize(ieo_state *st)
{
    const ieo_config *c = &st->cfg;
    const size_t gs = st->gs;
    st->t0 = 0.0;
    st->t = 0.0;
    for (int q = 0; q < 6; ++q) {
        memset(st->h[q], 0, gs * sizeof(float));
        memset(st->hd[q], 0, gs * sizeof(float));
    }
    const double Phi_from_zeta = 3.0 * (1.0 + c->omega) / (5.0 + 3.0 * c->omega);
    if (c->perform_deltaN) {
        double Nmean = 0.;
        for (size_t id = 0; id < gs; ++id) Nmean += st->dn[id];
        Nmean = Nmean / (double)gs;
        for (size_t id = 0; id < gs; ++id) st->f[id] = Phi_from_zeta * (st->dn[id] - Nmean);
    } else {
        double fmean = 0.0, fdmean = 0.0;
        for (size_t id = 0; id < gs; ++id) {
            fmean += st->f[id];
            fdmean += st->fd[id];
        }
        fmean = fmean / (double)gs;
        fdmean = fdmean / (double)gs;
        const double s = c->rescale_s;
        for (size_t id = 0; id < gs; ++id)
            st->f[id] = Phi_from_zeta * (st->f[id] - fmean)
                      / (fdmean * pow(st->a, s - 1.0) / (st->ad * pow(st->a, s - 2.0)));
    }
    st->a = 1.0;
    st->ad = c->horizon_factor * st->N * 2.0 * st->pi / c->L;
}

/* evolve_derivs_post_inflation (src/evolution.cpp:1437-1564): half kicks of ad
 * with the fluid ad2, the Phi kick, then the tensor kick whose source uses the
 * kicked Phi' of every site. */
void ieo_post_evolve_derivs(ieo_state *st, double d)
{
    const int N = st->N;
    const double s = st->cfg.rescale_s, w = st->cfg.omega;
    const double laplnorm = pow(st->a, -2.0 * s) / sq(st->dx);
    st->ad2 = -(s - 0.5 * (1.0 - 3.0 * w)) * sq(st->ad) / st->a;
    st->ad += 0.5 * d * st->ad2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                const size_t id = IX(N, i, j, k);
                st->fd[id] += d * (w * laplnorm * lapl_d(N, st->f, i, j, k)
                                   - (3.0 * (1.0 + w) + s) * st->ad * st->fd[id] / st->a);
            }
    const double Htilde = st->ad / st->a;
    const double invH = 1.0 / Htilde;
    const double coeffU = 4.0 / (3.0 * (1.0 + w));
    const double srcAmp = 2.0 * pow(st->a, -2.0 * s);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                const size_t id = IX(N, i, j, k);
                double S[6];
                post_source(N, st->dx, st->f, st->fd, i, j, k, invH, coeffU, S);
                for (int c = 0; c < 6; ++c) {
                    const double lap_h = (double)lapl_f(N, st->h[c], i, j, k);
                    const double rhs = d * (laplnorm * lap_h
                                            - (2.0 + s) * st->ad * (double)st->hd[c][id] / st->a
                                            + srcAmp * S[c]);
                    st->hd[c][id] += (float)rhs;
                }
            }
    st->ad += 0.5 * d * st->ad2;
}

static double rescaled_step(const ieo_state *st, double astep_value)
{
    return st->cfg.dt * pow(astep_value, st->cfg.rescale_s - 1.0);
}

int ieo_leapfrog_steps(ieo_state *st, int nsteps)
{
    int n = 0;
    while (n < nsteps && st->a <= st->cfg.af) {
        const double d = rescaled_step(st, st->astep);
        ieo_evolve_derivs(st, d);
        ieo_evolve_fields(st, d);
        ++n;
        st->astep = st->a;
    }
    return n;
}

/* ---------------------------------------------------------------- RK4/RK45 */

/* Butcher tables (src/evolution.cpp:373-414,595-637). */
static const double RK4A[4][4] = {{0, 0, 0, 0}, {0.5, 0, 0, 0}, {0, 0.5, 0, 0}, {0, 0, 1.0, 0}};
static const double RK4B[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
#define DP_a21 (1.0 / 5.0)
#define DP_a31 (3.0 / 40.0)
#define DP_a32 (9.0 / 40.0)
#define DP_a41 (44.0 / 45.0)
#define DP_a42 (-56.0 / 15.0)
#define DP_a43 (32.0 / 9.0)
#define DP_a51 (19372.0 / 6561.0)
#define DP_a52 (-25360.0 / 2187.0)
#define DP_a53 (64448.0 / 6561.0)
#define DP_a54 (-212.0 / 729.0)
#define DP_a61 (9017.0 / 3168.0)
#define DP_a62 (-355.0 / 33.0)
#define DP_a63 (46732.0 / 5247.0)
#define DP_a64 (49.0 / 176.0)
#define DP_a65 (-5103.0 / 18656.0)
#define DP_b1 (35.0 / 384.0)
#define DP_b3 (500.0 / 1113.0)
#define DP_b4 (125.0 / 192.0)
#define DP_b5 (-2187.0 / 6784.0)
#define DP_b6 (11.0 / 84.0)
#define DP_bs1 (5179.0 / 57600.0)
#define DP_bs3 (7571.0 / 16695.0)
#define DP_bs4 (393.0 / 640.0)
#define DP_bs5 (-92097.0 / 339200.0)
#define DP_bs6 (187.0 / 2100.0)
#define DP_bs7 (1.0 / 40.0)
static const double DPA[7][7] = {
    {0, 0, 0, 0, 0, 0, 0},
    {DP_a21, 0, 0, 0, 0, 0, 0},
    {DP_a31, DP_a32, 0, 0, 0, 0, 0},
    {DP_a41, DP_a42, DP_a43, 0, 0, 0, 0},
    {DP_a51, DP_a52, DP_a53, DP_a54, 0, 0, 0},
    {DP_a61, DP_a62, DP_a63, DP_a64, DP_a65, 0, 0},
    {DP_b1, 0, DP_b3, DP_b4, DP_b5, DP_b6, 0}};
static const double DPB[7] = {DP_b1, 0, DP_b3, DP_b4, DP_b5, DP_b6, 0};
static const double DPE[7] = {DP_b1 - DP_bs1, 0, DP_b3 - DP_bs3, DP_b4 - DP_bs4,
                              DP_b5 - DP_bs5, DP_b6 - DP_bs6, -DP_bs7};

typedef struct {
    size_t gs;
    double *ftmp, *fdtmp, *dntmp;
    double *kf[7], *kfd[7], *kdn[7];
    float *htmp[6], *hdtmp[6];
    float *kh[7][6], *khd[7][6];
} rk_scratch;

static void rk_scratch_init(rk_scratch *s, size_t gs, int sigw, int with_dn)
{
    memset(s, 0, sizeof(*s));
    s->gs = gs;
    s->ftmp = calloc(gs, sizeof(double));
    s->fdtmp = calloc(gs, sizeof(double));
    if (with_dn) s->dntmp = calloc(gs, sizeof(double));
    for (int q = 0; q < 7; ++q) {
        s->kf[q] = calloc(gs, sizeof(double));
        s->kfd[q] = calloc(gs, sizeof(double));
        if (with_dn) s->kdn[q] = calloc(gs, sizeof(double));
        if (sigw)
            for (int c = 0; c < 6; ++c) {
                s->kh[q][c] = calloc(gs, sizeof(float));
                s->khd[q][c] = calloc(gs, sizeof(float));
            }
    }
    if (sigw)
        for (int c = 0; c < 6; ++c) {
            s->htmp[c] = calloc(gs, sizeof(float));
            s->hdtmp[c] = calloc(gs, sizeof(float));
        }
}

static void rk_scratch_free(rk_scratch *s)
{
    free(s->ftmp); free(s->fdtmp); free(s->dntmp);
    for (int q = 0; q < 7; ++q) {
        free(s->kf[q]); free(s->kfd[q]); free(s->kdn[q]);
        for (int c = 0; c < 6; ++c) { free(s->kh[q][c]); free(s->khd[q][c]); }
    }
    for (int c = 0; c < 6; ++c) { free(s->htmp[c]); free(s->hdtmp[c]); }
}

/* Inflation RHS with fused energy sums (src/evolution.cpp:419-518, inflation
 * branch).  The reference sums with an OpenMP reduction; here the sums are serial
 * (equal to the reference run with one thread). */
static void inflation_rhs(ieo_state *st, const double *fs, const double *fds,
                          float *const *hs, float *const *hds,
                          double *dfdt, double *dfddt, float **dhdt, float **dhddt,
                          double a_s, double ad_s, double *dadt, double *daddt)
{
    const int N = st->N;
    const double s = st->cfg.rescale_s;
    const double laplnorm = pow(a_s, -2.0 * s) / sq(st->dx);
    const double friction = (2.0 + s) * ad_s / a_s;
    const double potnorm = pow(a_s, 2.0 - 2.0 * s);
    const double srcAmp = 2.0 * pow(a_s, -2.0 * s);
    double gsum = 0.0, psum = 0.0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                const size_t id = IX(N, i, j, k);
                const double fh = fs[id];
                const double lf = lapl_d(N, fs, i, j, k);
                double pot = 0.0, dpot = 0.0;
                if (st->cfg.numerical_potential) {
                    int nh = 1;
                    eval_pot(st, fh, st->lstart[id], st->cfg.int_err, &nh, &pot, &dpot);
                    st->lstart[id] = nh;
                } else {
                    eval_pot(st, fh, 1, 1, NULL, &pot, &dpot);
                }
                dfdt[id] = fds[id];
                dfddt[id] = laplnorm * lf - friction * fds[id] - potnorm * dpot;
                gsum -= fh * lf;
                psum += pot;
                if (st->cfg.calculate_sigw) {
                    double g[3];
                    grad_site(N, st->dx, fs, i, j, k, g);
                    const double T[6] = {g[0] * g[0], g[1] * g[1], g[2] * g[2],
                                         g[0] * g[1], g[0] * g[2], g[1] * g[2]};
                    for (int c = 0; c < 6; ++c) dhdt[c][id] = hds[c][id];
                    for (int c = 0; c < 6; ++c)
                        dhddt[c][id] = (float)(laplnorm * (double)lapl_f(N, hs[c], i, j, k)
                                               - friction * (double)hds[c][id] + srcAmp * T[c]);
                }
            }
    const double grad = 0.5 * gsum * sq(1.0 / (a_s * st->dx)) / (double)st->gs;
    const double pavg = psum / (double)st->gs;
    const double source = 2.0 * grad / 3.0 + pavg;
    *dadt = ad_s;
    *daddt = pow(a_s, 3.0 - 2.0 * s) * source - (s + 1.0) * sq(ad_s) / a_s;
}

/* Post-inflation right-hand side (src/evolution.cpp:520-592). */
static void post_rhs(ieo_state *st, const double *fs, const double *fds,
                     float *const *hs, float *const *hds,
                     double *dfdt, double *dfddt, float **dhdt, float **dhddt,
                     double a_s, double ad_s, double *dadt, double *daddt)
{
    const int N = st->N;
    const double s = st->cfg.rescale_s, w = st->cfg.omega;
    const double laplnorm = pow(a_s, -2.0 * s) / sq(st->dx);
    const double scalar_friction = (3.0 * (1.0 + w) + s) * ad_s / a_s;
    const double tensor_friction = (2.0 + s) * ad_s / a_s;
    const double Htilde = ad_s / a_s;
    const double invH = 1.0 / Htilde;
    const double coeffU = 4.0 / (3.0 * (1.0 + w));
    const double srcAmp = 2.0 * pow(a_s, -2.0 * s);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                const size_t id = IX(N, i, j, k);
                const double lap_f = lapl_d(N, fs, i, j, k);
                dfdt[id] = fds[id];
                dfddt[id] = w * laplnorm * lap_f - scalar_friction * fds[id];
                double S[6];
                post_source(N, st->dx, fs, fds, i, j, k, invH, coeffU, S);
                for (int c = 0; c < 6; ++c) dhdt[c][id] = hds[c][id];
                for (int c = 0; c < 6; ++c)
                    dhddt[c][id] = (float)(laplnorm * (double)lapl_f(N, hs[c], i, j, k)
                                           - tensor_friction * (double)hds[c][id] + srcAmp * S[c]);
            }
    *dadt = ad_s;
    *daddt = -(s - 0.5 * (1.0 - 3.0 * w)) * sq(ad_s) / a_s;
}

/* Stage state y + h * sum_s A[stage][s] k_s (src/evolution.cpp:639-713). */
static void rk_stage(ieo_state *st, rk_scratch *sc, double h, int stage, const double *A, int ns,
                     const double *ka, const double *kad, double *a_s, double *ad_s)
{
    const size_t gs = st->gs;
    for (size_t id = 0; id < gs; ++id) {
        double sf = 0.0, sfd = 0.0;
        for (int q = 0; q < stage; ++q) {
            const double c = A[stage * ns + q];
            if (c == 0.0) continue;
            sf += c * sc->kf[q][id];
            sfd += c * sc->kfd[q][id];
        }
        sc->ftmp[id] = st->f[id] + h * sf;
        sc->fdtmp[id] = st->fd[id] + h * sfd;
    }
    if (st->cfg.calculate_sigw)
        for (int c = 0; c < 6; ++c)
            for (size_t id = 0; id < gs; ++id) {
                double sh = 0.0, shd = 0.0;
                for (int q = 0; q < stage; ++q) {
                    const double cs = A[stage * ns + q];
                    if (cs == 0.0) continue;
                    sh += cs * (double)sc->kh[q][c][id];
                    shd += cs * (double)sc->khd[q][c][id];
                }
                sc->htmp[c][id] = (float)((double)st->h[c][id] + h * sh);
                sc->hdtmp[c][id] = (float)((double)st->hd[c][id] + h * shd);
            }
    *a_s = st->a;
    *ad_s = st->ad;
    for (int q = 0; q < stage; ++q) {
        const double c = A[stage * ns + q];
        if (c == 0.0) continue;
        *a_s += h * c * ka[q];
        *ad_s += h * c * kad[q];
    }
}

static void rk_eval_stages(ieo_state *st, rk_scratch *sc, double h, const double *A, int ns,
                           double *ka, double *kad)
{
    /* compute_inflation_rhs dispatches on the post-inflation mode flag
     * (src/evolution.cpp:325,364-371,439) */
    void (*rhs)(ieo_state *, const double *, const double *, float *const *, float *const *, double *,
                double *, float **, float **, double, double, double *, double *) =
        st->post_rhs ? post_rhs : inflation_rhs;
    rhs(st, st->f, st->fd, st->h, st->hd, sc->kf[0], sc->kfd[0], sc->kh[0], sc->khd[0],
        st->a, st->ad, &ka[0], &kad[0]);
    for (int stage = 1; stage < ns; ++stage) {
        double a_s, ad_s;
        rk_stage(st, sc, h, stage, A, ns, ka, kad, &a_s, &ad_s);
        rhs(st, sc->ftmp, sc->fdtmp, sc->htmp, sc->hdtmp, sc->kf[stage], sc->kfd[stage],
            sc->kh[stage], sc->khd[stage], a_s, ad_s, &ka[stage], &kad[stage]);
    }
}

static rk_scratch *rk_get(ieo_state *st, int with_dn)
{
    /* scratch is allocated per call: RK is a secondary path exercised at small N */
    rk_scratch *sc = (rk_scratch *)malloc(sizeof(rk_scratch));
    rk_scratch_init(sc, st->gs, st->cfg.calculate_sigw, with_dn);
    return sc;
}

static void rk_put(rk_scratch *sc) { rk_scratch_free(sc); free(sc); }

/* src/evolution.cpp:760-779 (with apply_inflation_weighted_update :715-758) */
void ieo_rk4_step(ieo_state *st, double h)
{
    rk_scratch *sc = rk_get(st, 0);
    double ka[4] = {0}, kad[4] = {0};
    rk_eval_stages(st, sc, h, &RK4A[0][0], 4, ka, kad);
    const size_t gs = st->gs;
    for (size_t id = 0; id < gs; ++id) {
        double sf = 0.0, sfd = 0.0;
        for (int q = 0; q < 4; ++q) {
            sf += RK4B[q] * sc->kf[q][id];
            sfd += RK4B[q] * sc->kfd[q][id];
        }
        st->f[id] += h * sf;
        st->fd[id] += h * sfd;
    }
    if (st->cfg.calculate_sigw)
        for (int c = 0; c < 6; ++c)
            for (size_t id = 0; id < gs; ++id) {
                double sh = 0.0, shd = 0.0;
                for (int q = 0; q < 4; ++q) {
                    sh += RK4B[q] * (double)sc->kh[q][c][id];
                    shd += RK4B[q] * (double)sc->khd[q][c][id];
                }
                st->h[c][id] = (float)((double)st->h[c][id] + h * sh);
                st->hd[c][id] = (float)((double)st->hd[c][id] + h * shd);
            }
    for (int q = 0; q < 4; ++q) {
        st->a += h * RK4B[q] * ka[q];
        st->ad += h * RK4B[q] * kad[q];
    }
    st->t += h;
    rk_put(sc);
}

static double clamp_rk45(const ieo_state *st, double h, double hmax)
{
    const double hmin = st->cfg.rk45_min_dt > 1e-16 ? st->cfg.rk45_min_dt : 1e-16;
    const double hhi = hmax > hmin ? hmax : hmin;
    if (h < hmin) return hmin;
    if (h > hhi) return hhi;
    return h;
}

static double rk45_factor(const ieo_state *st, double err_norm)
{
    const double safe = err_norm > 1e-16 ? err_norm : 1e-16;
    double factor = st->cfg.rk45_safety * pow(safe, -0.2);
    if (factor < 0.1) factor = 0.1;
    if (factor > 5.0) factor = 5.0;
    return factor;
}

/* Dormand-Prince 5(4) trial step (src/evolution.cpp:781-904). Returns 1 when
 * accepted (state committed), 0 when rejected. */
int ieo_rk45_step(ieo_state *st, double h, double hmax, double *h_next)
{
    rk_scratch *sc = rk_get(st, 0);
    double ka[7] = {0}, kad[7] = {0};
    rk_eval_stages(st, sc, h, &DPA[0][0], 7, ka, kad);
    double a5 = st->a, ad5 = st->ad, erra = 0.0, errad = 0.0;
    for (int q = 0; q < 7; ++q) {
        if (DPB[q] != 0.0) { a5 += h * DPB[q] * ka[q]; ad5 += h * DPB[q] * kad[q]; }
        if (DPE[q] != 0.0) { erra += h * DPE[q] * ka[q]; errad += h * DPE[q] * kad[q]; }
    }
    const size_t gs = st->gs;
    const double at = st->cfg.rk45_abs_tol, rt = st->cfg.rk45_rel_tol;
    double acc = 0.0;
    size_t nvars = 2 * gs + 2;
    for (size_t id = 0; id < gs; ++id) {
        double y5f = st->f[id], y5fd = st->fd[id], ef = 0.0, efd = 0.0;
        for (int q = 0; q < 7; ++q) {
            if (DPB[q] != 0.0) { y5f += h * DPB[q] * sc->kf[q][id]; y5fd += h * DPB[q] * sc->kfd[q][id]; }
            if (DPE[q] != 0.0) { ef += h * DPE[q] * sc->kf[q][id]; efd += h * DPE[q] * sc->kfd[q][id]; }
        }
        sc->ftmp[id] = y5f;
        sc->fdtmp[id] = y5fd;
        const double sfv = at + rt * fmax(fabs(st->f[id]), fabs(y5f));
        const double sfdv = at + rt * fmax(fabs(st->fd[id]), fabs(y5fd));
        acc += sq(ef / sfv) + sq(efd / sfdv);
    }
    if (st->cfg.calculate_sigw) {
        nvars += 12 * gs;
        for (int c = 0; c < 6; ++c)
            for (size_t id = 0; id < gs; ++id) {
                double y5h = (double)st->h[c][id], y5hd = (double)st->hd[c][id], eh = 0.0, ehd = 0.0;
                for (int q = 0; q < 7; ++q) {
                    if (DPB[q] != 0.0) {
                        y5h += h * DPB[q] * (double)sc->kh[q][c][id];
                        y5hd += h * DPB[q] * (double)sc->khd[q][c][id];
                    }
                    if (DPE[q] != 0.0) {
                        eh += h * DPE[q] * (double)sc->kh[q][c][id];
                        ehd += h * DPE[q] * (double)sc->khd[q][c][id];
                    }
                }
                sc->htmp[c][id] = (float)y5h;
                sc->hdtmp[c][id] = (float)y5hd;
                const double shv = at + rt * fmax(fabs((double)st->h[c][id]), fabs(y5h));
                const double shdv = at + rt * fmax(fabs((double)st->hd[c][id]), fabs(y5hd));
                acc += sq(eh / shv) + sq(ehd / shdv);
            }
    }
    const double sa = at + rt * fmax(fabs(st->a), fabs(a5));
    const double sad = at + rt * fmax(fabs(st->ad), fabs(ad5));
    acc += sq(erra / sa) + sq(errad / sad);
    const double err_norm = sqrt(acc / (double)nvars);
    const double factor = rk45_factor(st, err_norm);
    int accepted = 0;
    if (err_norm <= 1.0) {
        double *tmp;
        tmp = st->f; st->f = sc->ftmp; sc->ftmp = tmp;
        tmp = st->fd; st->fd = sc->fdtmp; sc->fdtmp = tmp;
        if (st->cfg.calculate_sigw)
            for (int c = 0; c < 6; ++c) {
                float *t2;
                t2 = st->h[c]; st->h[c] = sc->htmp[c]; sc->htmp[c] = t2;
                t2 = st->hd[c]; st->hd[c] = sc->hdtmp[c]; sc->hdtmp[c] = t2;
            }
        st->a = a5;
        st->ad = ad5;
        st->t += h;
        *h_next = clamp_rk45(st, h * factor, hmax);
        accepted = 1;
    } else {
        const double shrink = factor < 0.5 ? factor : 0.5;
        *h_next = clamp_rk45(st, h * (shrink > 0.1 ? shrink : 0.1), hmax);
    }
    rk_put(sc);
    return accepted;
}

/* rk4_step_inflation / rk45_step_inflation under ScopedPostInflationRhsMode
 * (src/evolution.cpp:1606-1640). */
void ieo_post_rk4_step(ieo_state *st, double h)
{
    st->post_rhs = 1;
    ieo_rk4_step(st, h);
    st->post_rhs = 0;
}

int ieo_post_rk45_step(ieo_state *st, double h, double hmax, double *h_next)
{
    st->post_rhs = 1;
    const int acc = ieo_rk45_step(st, h, hmax, h_next);
    st->post_rhs = 0;
    return acc;
}

/* ---------------------------------------------------------------- spectra */

int ieo_spectrum_numbins(int N) { return (int)(1.73205 * (N / 2)) + 1; }

/* src/output.cpp:63-151 */
void ieo_bin_spectrum(int N, double L, const double *F, const double *ny, double norm1,
                      double *p, int *counts, double *vals)
{
    const int nb = ieo_spectrum_numbins(N);
    const double pi = 2.0 * asin(1.0);
    const double dp = 2.0 * pi / L;
    for (int b = 0; b < nb; ++b) { p[b] = dp * b; counts[b] = 0; vals[b] = 0.0; }
    for (int i = 0; i < N; ++i) {
        const int px = i <= N / 2 ? i : i - N;
        for (int j = 0; j < N; ++j) {
            const int py = j <= N / 2 ? j : j - N;
            for (int k = 1; k < N / 2; ++k) {
                const double pm = sqrt(sq((double)px) + sq((double)py) + sq((double)k));
                const int bin = (int)lround(pm);
                if (bin < 0 || bin >= nb) continue;
                const double re = F[IX(N, i, j, 2 * k)], im = F[IX(N, i, j, 2 * k + 1)];
                counts[bin] += 2;
                vals[bin] += 2.0 * (re * re + im * im);
            }
            for (int k = 0; k <= N / 2; k += N / 2) {
                const double pm = sqrt(sq((double)px) + sq((double)py) + sq((double)k));
                const int bin = (int)lround(pm);
                if (bin < 0 || bin >= nb) continue;
                double re, im;
                if (k == 0) { re = F[IX(N, i, j, 0)]; im = F[IX(N, i, j, 1)]; }
                else { re = ny[(size_t)i * 2 * N + 2 * j]; im = ny[(size_t)i * 2 * N + 2 * j + 1]; }
                counts[bin] += 1;
                vals[bin] += re * re + im * im;
            }
        }
    }
    for (int b = 0; b < nb; ++b) {
        if (counts[b] > 0) vals[b] /= (double)counts[b];
        vals[b] = norm1 * vals[b];
    }
}


static double norm_scalar(const ieo_state *st)
{
    return pow(st->cfg.L / st->cfg.rescale_B, 3) / pow((double)st->N, 6);
}

/* spectraf() without the forcing cutoff (src/output.cpp:297-406): forward FFT in
 * place, bin, inverse FFT to restore real space (the round trip is kept since it
 * perturbs f at round-off level, as in the reference). */
void ieo_spectrum_f(ieo_state *st, double *p, int *counts, double *vals)
{
    ieo_fftrn_d(st->f, st->fny, st->N, 1);
    ieo_bin_spectrum(st->N, st->cfg.L, st->f, st->fny, norm_scalar(st), p, counts, vals);
    if (st->cfg.high_cutoff_index > 0 && st->cfg.forcing_cutoff) ieo_cutoff_modes(st);
    ieo_fftrn_d(st->f, st->fny, st->N, -1);
}

static void kill_mode(double *field, double *deriv)
{
    field[0] = 0.;
    field[1] = 0.;
    deriv[0] = 0.;
    deriv[1] = 0.;
}

void ieo_cutoff_modes(ieo_state *st)
{
    const int N = st->N;
    const double hi = st->cfg.high_cutoff_index, lo = st->cfg.low_cutoff_index;
    double *f = st->f, *fd = st->fd, *fny = st->fny, *fdny = st->fdny;
    ieo_fftrn_d(fd, fdny, N, 1);
#define NY(a, i, jj) (a)[(size_t)(i) * 2 * N + (jj)]
    for (int i = 0; i < N; i++) {
        const int px = i <= N / 2 ? i : i - N;
        const int iconj = i == 0 ? 0 : N - i;
        for (int j = 0; j < N; j++) {
            const int py = j <= N / 2 ? j : j - N;
            for (int k = 1; k < N / 2; k++) {
                const double pdisc = sqrt(sq((double)px) + sq((double)py) + sq((double)k));
                if (pdisc > hi || pdisc < lo) kill_mode(&f[IX(N, i, j, 2 * k)], &fd[IX(N, i, j, 2 * k)]);
            }
            if (j > N / 2 || (i > N / 2 && (j == 0 || j == N / 2))) {
                const int jconj = j == 0 ? 0 : N - j;
                double pdisc = sqrt(sq((double)px) + sq((double)py));
                if (pdisc > hi || pdisc < lo) {
                    kill_mode(&f[IX(N, i, j, 0)], &fd[IX(N, i, j, 0)]);
                    f[IX(N, iconj, jconj, 0)] = f[IX(N, i, j, 0)];
                    f[IX(N, iconj, jconj, 1)] = -f[IX(N, i, j, 1)];
                    fd[IX(N, iconj, jconj, 0)] = fd[IX(N, i, j, 0)];
                    fd[IX(N, iconj, jconj, 1)] = -fd[IX(N, i, j, 1)];
                }
                pdisc = sqrt(sq((double)px) + sq((double)py) + sq((double)(N / 2.)));
                if (pdisc > hi || pdisc < lo) {
                    kill_mode(&NY(fny, i, 2 * j), &NY(fdny, i, 2 * j));
                    NY(fny, iconj, 2 * jconj) = NY(fny, i, 2 * j);
                    NY(fny, iconj, 2 * jconj + 1) = -NY(fny, i, 2 * j + 1);
                    NY(fdny, iconj, 2 * jconj) = NY(fdny, i, 2 * j);
                    NY(fdny, iconj, 2 * jconj + 1) = -NY(fdny, i, 2 * j + 1);
                }
            } else if ((i == 0 || i == N / 2) && (j == 0 || j == N / 2)) {
                double pdisc = sqrt(sq((double)px) + sq((double)py));
                if (pdisc > hi || pdisc < lo) kill_mode(&f[IX(N, i, j, 0)], &fd[IX(N, i, j, 0)]);
                pdisc = sqrt(sq((double)px) + sq((double)py) + sq((double)(N / 2.)));
                if (pdisc > hi || pdisc < lo) kill_mode(&NY(fny, i, 2 * j), &NY(fdny, i, 2 * j));
            }
        }
    }
#undef NY
    ieo_fftrn_d(fd, fdny, N, -1);
}

/* One FFTEASY-layout mode of f: k < N/2 from the N^3 array, k = N/2 from the
 * Nyquist plane (the reference reads f at 2k first and then replaces it; the
 * replaced read is skipped here, it can fall past the array). */
static void bis_mode(const ieo_state *st, int i, int j, int k, double *re, double *im)
{
    const int N = st->N;
    if (k == N / 2) {
        *re = st->fny[(size_t)i * 2 * N + 2 * j];
        *im = st->fny[(size_t)i * 2 * N + 2 * j + 1];
    } else {
        *re = st->f[IX(N, i, j, 2 * k)];
        *im = st->f[IX(N, i, j, 2 * k + 1)];
    }
}

void ieo_bispectrum(ieo_state *st, double *p, int *numpoints, double *bisreal, double *bisimag)
{
    const int N = st->N;
    const int numbins = (int)(sqrt(3.0) * (N / 2)) + 1;
    const double dp = 2.0 * st->pi / st->cfg.L;
    const double norm1 = pow(st->cfg.L / st->cfg.rescale_B, 6) / pow((double)N, 9);
    double mean = 0.;
    for (size_t id = 0; id < st->gs; ++id) mean += st->f[id];
    mean = mean / (double)st->gs;
    for (size_t id = 0; id < st->gs; ++id) st->f[id] -= mean;
    for (int k = 0; k < numbins; k++) {
        p[k] = dp * k;
        numpoints[k] = 0;
        bisreal[k] = 0.;
        bisimag[k] = 0.;
    }
    ieo_fftrn_d(st->f, st->fny, N, 1);
    for (int k = 0; k < numbins; k++) {
        const double kf = (double)k;
        for (int i1 = 0; i1 < N; i1++)
            for (int j1 = 0; j1 < N; j1++) {
                const double px1 = i1 <= N / 2 ? i1 : i1 - N;
                const double py1 = j1 <= N / 2 ? j1 : j1 - N;
                const double pzaus1 = sq(kf) - sq(px1) - sq(py1);
                if (!(pzaus1 >= 0)) continue;
                for (int k1 = (int)round(sqrt(pzaus1)) - 1; k1 < (int)round(sqrt(pzaus1)) + 2; k1++) {
                    if (!(fabs(sqrt(sq(px1) + sq(py1) + sq((double)k1)) - kf) < 1.5 && k1 <= N / 2 && k1 >= 0))
                        continue;
                    for (int i2 = 0; i2 < N; i2++)
                        for (int j2 = 0; j2 < N; j2++) {
                            const double px2 = i2 <= N / 2 ? i2 : i2 - N;
                            const double py2 = j2 <= N / 2 ? j2 : j2 - N;
                            const double pzaus2 = sq(kf) - sq(px2) - sq(py2);
                            if (!(pzaus2 >= 0)) continue;
                            for (int k2 = (int)round(sqrt(pzaus2)) - 1; k2 < (int)round(sqrt(pzaus2)) + 2; k2++) {
                                if (!(fabs(sqrt(sq(px2) + sq(py2) + sq((double)k2)) - kf) < 1.5 && k2 <= N / 2 &&
                                      k2 >= 0))
                                    continue;
                                const double px3 = px1 + px2, py3 = py1 + py2;
                                int k3 = k1 + k2;
                                if (!(px3 <= N / 2 && px3 > -N / 2 && py3 <= N / 2 && py3 > -N / 2)) continue;
                                const int i3 = (int)(px3 >= 0 ? px3 : px3 + N);
                                const int j3 = (int)(py3 >= 0 ? py3 : py3 + N);
                                double f1r, f1i, f2r, f2i, f3r, f3i;
                                if (fabs(sqrt(sq(px3) + sq(py3) + sq((double)k3)) - kf) < 1.5 && k3 <= N / 2) {
                                    bis_mode(st, i1, j1, k1, &f1r, &f1i);
                                    bis_mode(st, i2, j2, k2, &f2r, &f2i);
                                    bis_mode(st, i3, j3, k3, &f3r, &f3i);
                                    double counts = 1.;
                                    if (k1 != N / 2 && k2 != N / 2 && (k1 != 0 || k2 != 0)) counts = 2.;
                                    numpoints[k] += (int)counts;
                                    bisreal[k] += counts * (f1r * f2r * f3r - f1i * f2i * f3r + f1i * f2r * f3i
                                                            + f2i * f1r * f3i);
                                    bisimag[k] += counts * (-f1r * f2r * f3i + f1i * f2i * f3i + f1i * f2r * f3r
                                                            + f2i * f1r * f3r);
                                }
                                if (k1 != 0 && k2 != 0) {
                                    k3 = k1 - k2;
                                    if (k3 >= 0 && fabs(sqrt(sq(px3) + sq(py3) + sq((double)k3)) - kf) < 1.5 &&
                                        k3 <= N / 2) {
                                        const int i2n = (int)(-px2 >= 0 ? -px2 : -px2 + N);
                                        const int j2n = (int)(-py2 >= 0 ? -py2 : -py2 + N);
                                        bis_mode(st, i1, j1, k1, &f1r, &f1i);
                                        bis_mode(st, i2n, j2n, k2, &f2r, &f2i);
                                        f2i = -f2i;
                                        bis_mode(st, i3, j3, k3, &f3r, &f3i);
                                        numpoints[k] += 2;
                                        bisreal[k] += 2.0 * (f1r * f2r * f3r - f1i * f2i * f3r + f1i * f2r * f3i
                                                             + f2i * f1r * f3i);
                                        bisimag[k] += 2.0 * (-f1r * f2r * f3i + f1i * f2i * f3i + f1i * f2r * f3r
                                                             + f2i * f1r * f3r);
                                    }
                                }
                            }
                        }
                }
            }
    }
    for (int k = 0; k < numbins; k++) {
        if (numpoints[k] > 0) {
            bisreal[k] = bisreal[k] / numpoints[k];
            bisimag[k] = bisimag[k] / numpoints[k];
        }
        bisreal[k] = norm1 * bisreal[k];
        bisimag[k] = norm1 * bisimag[k];
    }
    ieo_fftrn_d(st->f, st->fny, N, -1);
    for (size_t id = 0; id < st->gs; ++id) st->f[id] += mean;
}

/* factt = <fd a^{s-1} / (ad a^{s-2})>, fmean = <f> as the LOG outputs form them */
static int log_factors(const ieo_state *st, double *factt, double *fmean)
{
    const double s = st->cfg.rescale_s;
    double ft = 0.0, fm = 0.0;
    for (size_t id = 0; id < st->gs; ++id) {
        ft += st->fd[id] * pow(st->a, s - 1.0) / (st->ad * pow(st->a, s - 2.0));
        fm += st->f[id];
    }
    *factt = ft / (double)st->gs;
    *fmean = fm / (double)st->gs;
    return !(!isfinite(*factt) || fabs(*factt) < 1e-30);
}

int ieo_deltaN_log(ieo_state *st)
{
    double factt, fmean;
    const double eta = st->cfg.eta_log;
    if (!log_factors(st, &factt, &fmean)) {
        for (size_t id = 0; id < st->gs; ++id) st->dn[id] = 0.0;
        return 0;
    }
    for (size_t id = 0; id < st->gs; ++id) {
        st->dn[id] = 0.0;
        const double arg = 1.0 - eta * (st->f[id] - fmean) / factt;
        if (arg > 0.0) st->dn[id] = (1.0 / eta) * log(arg);
    }
    return 1;
}

void ieo_spectrum_log(ieo_state *st, double *p, int *counts, double *vals)
{
    ieo_deltaN_log(st);
    ieo_fftrn_d(st->dn, st->fny, st->N, 1);
    ieo_bin_spectrum(st->N, st->cfg.L, st->dn, st->fny, norm_scalar(st), p, counts, vals);
    ieo_fftrn_d(st->dn, st->fny, st->N, -1);
}

int ieo_histogram_log(ieo_state *st, double *bmin_out, double *df_out, double *freq)
{
    const int nbins = st->cfg.nbins;
    const double eta = st->cfg.eta_log;
    double factt, fmean;
    if (!log_factors(st, &factt, &fmean)) {
        for (int b = 0; b < nbins; ++b) freq[b] = 0.0;
        *bmin_out = 0.0;
        *df_out = 1.0;
        return 0;
    }
    for (size_t id = 0; id < st->gs; ++id) {
        st->dn[id] = 0;
        if ((1 - eta * (st->f[id] - fmean) / factt) > 0)
            st->dn[id] = 1. / eta * log(1 - eta * (st->f[id] - fmean) / factt);
    }
    double bmin = st->dn[0], bmax = bmin;
    for (size_t id = 0; id < st->gs; ++id)
        if (1 - eta * (st->f[id] - fmean) / factt > 0) {
            bmin = st->dn[id] < bmin ? st->dn[id] : bmin;
            bmax = st->dn[id] > bmax ? st->dn[id] : bmax;
        }
    double df = (bmax - bmin) / (double)nbins;
    if (!isfinite(df) || df <= 0.0) df = 1.0;
    for (int b = 0; b < nbins; ++b) freq[b] = 0.0;
    int numpts = 0;
    for (size_t id = 0; id < st->gs; ++id)
        if (1 - eta * (st->f[id] - fmean) / factt > 0) {
            int b = (int)((st->dn[id] - bmin) / df);
            if (st->dn[id] == bmax) b = nbins - 1;
            if (b >= 0 && b < nbins) {
                freq[b]++;
                numpts++;
            }
        }
    if (numpts == 0) numpts = 1;
    for (int b = 0; b < nbins; ++b) freq[b] = freq[b] / (double)numpts;
    *bmin_out = bmin;
    *df_out = df;
    return 1;
}

/* Symmetric tensor packing (src/main.h:43-65). */
static int sym_idx(int l, int m)
{
    if (l > m) { const int t = l; l = m; m = t; }
    if (l == m) return l;
    if (l == 0 && m == 1) return 3;
    if (l == 0 && m == 2) return 4;
    return 5;
}

/* TT projection |Lambda C|^2 of one mode, float arithmetic
 * (src/output.cpp:504-543). */
static float tt_power(const float Cr[3][3], const float Ci[3][3], float kx, float ky, float kz, float kt2)
{
    const float inv = 1.0f / sqrtf(kt2);
    const float kh[3] = {kx * inv, ky * inv, kz * inv};
    float P[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) P[a][b] = (a == b ? 1.f : 0.f) - kh[a] * kh[b];
    float Ar[3][3], Ai[3][3];
    float Tr = 0.f, Ti = 0.f;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            float r = 0.f, im = 0.f;
            for (int l = 0; l < 3; ++l)
                for (int m = 0; m < 3; ++m) {
                    const float pal = P[a][l], pbm = P[b][m];
                    r += pal * Cr[l][m] * pbm;
                    im += pal * Ci[l][m] * pbm;
                }
            Ar[a][b] = r;
            Ai[a][b] = im;
        }
    for (int l = 0; l < 3; ++l)
        for (int m = 0; m < 3; ++m) {
            Tr += P[l][m] * Cr[l][m];
            Ti += P[l][m] * Ci[l][m];
        }
    float fp2 = 0.f;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const float Xr = Ar[a][b] - 0.5f * P[a][b] * Tr;
            const float Xi = Ai[a][b] - 0.5f * P[a][b] * Ti;
            fp2 += Xr * Xr + Xi * Xi;
        }
    return fp2;
}

/* write_gw_spectrum_impl (src/output.cpp:466-625) with GwScaleMode None (h) or
 * DoubleThenFloat (hdot, to_phys = B a^{s-1}).  Per-shell sums are float, as in
 * the reference. */
void ieo_gw_spectrum(ieo_state *st, int hdot, double *p_out, int *counts, double *vals)
{
    const int N = st->N;
    const int nb = (int)(sqrt(3.0) * (N / 2)) + 1;
    const int imax = (int)floor(0.80 * nb);
    const float pif = (float)st->pi;
    const float dp = 2.f * pif / (float)st->cfg.L;
    float **h = hdot ? st->hd : st->h;
    float **hn = hdot ? st->hdny : st->hny;
    const double to_phys = hdot ? st->cfg.rescale_B * pow(st->a, st->cfg.rescale_s - 1.0) : 1.0;
    /* the post-inflation variant rounds the factor to float and multiplies in
     * float (spectraGWdot_post_inflation, src/output.cpp:662-669) */
    const float tp = (float)(st->cfg.rescale_B * pow(st->a, st->cfg.rescale_s - 1.f));
    float *f2 = (float *)calloc(nb, sizeof(float));
    for (int b = 0; b < nb; ++b) counts[b] = 0;
    for (int c = 0; c < 6; ++c) ieo_fftrn_f(h[c], hn[c], N, 1);
    const float kpre = 2.f / (float)st->dx;
    for (int i = 0; i < N; ++i) {
        const int px = i <= N / 2 ? i : i - N;
        for (int j = 0; j < N; ++j) {
            const int py = j <= N / 2 ? j : j - N;
            /* interior modes 1 <= kz < N/2 (weight 2), then kz = 0 and kz = N/2
             * (weight 1, the latter from the Nyquist plane), in this order */
            for (int q = 1; q < N / 2 + 2; ++q) {
                const int special = q >= N / 2;
                const int k = special ? (q - N / 2) * (N / 2) : q;
                const float kx = kpre * sinf(pif * px / (float)N);
                const float ky = kpre * sinf(pif * py / (float)N);
                const float kz = kpre * sinf(pif * k / (float)N);
                const float kt2 = kx * kx + ky * ky + kz * kz;
                if (kt2 == 0.f) continue;
                const int bin = (int)lroundf(sqrtf((float)(px * px + py * py + k * k)));
                if (bin >= imax) continue;
                float Cr[3][3], Ci[3][3];
                for (int l = 0; l < 3; ++l)
                    for (int m = l; m < 3; ++m) {
                        const int comp = sym_idx(l, m);
                        float re, im;
                        if (!special || k == 0) {
                            const size_t id = IX(N, i, j, 2 * k);
                            re = h[comp][id];
                            im = h[comp][id + 1];
                        } else {
                            re = hn[comp][(size_t)i * 2 * N + 2 * j];
                            im = hn[comp][(size_t)i * 2 * N + 2 * j + 1];
                        }
                        if (hdot == 2) {
                            re *= tp;
                            im *= tp;
                        } else if (hdot) {
                            re = (float)((double)re * to_phys);
                            im = (float)((double)im * to_phys);
                        }
                        Cr[l][m] = re; Ci[l][m] = im;
                        if (m != l) { Cr[m][l] = re; Ci[m][l] = im; }
                    }
                const float fp2 = tt_power(Cr, Ci, kx, ky, kz, kt2);
                if (special) { counts[bin] += 1; f2[bin] += fp2; }
                else { counts[bin] += 2; f2[bin] += 2.f * fp2; }
            }
        }
    }
    for (int c = 0; c < 6; ++c) ieo_fftrn_f(h[c], hn[c], N, -1);
    const float norm1 = 0.5f * pow(st->cfg.L / st->cfg.rescale_B, 3) / pow((float)N, 6);
    for (int b = 0; b < nb; ++b) {
        if (counts[b] > 0) f2[b] /= (float)counts[b];
        p_out[b] = (double)(dp * b);
        vals[b] = (double)(norm1 * f2[b]);
    }
    free(f2);
}

/* get_modes (src/output.cpp:674-738): mean lattice momentum per truncated shell. */
void ieo_modes(int N, double L, double rescale_B, double *out)
{
    const int nb = (int)(sqrt(3.0) * (N / 2)) + 1;
    const double pi = 2.0 * asin(1.0);
    int *np = (int *)calloc(nb, sizeof(int));
    double *pp = (double *)calloc(nb, sizeof(double));
    for (int i = 0; i < N; ++i) {
        const int px = i <= N / 2 ? i : i - N;
        for (int j = 0; j < N; ++j) {
            const int py = j <= N / 2 ? j : j - N;
            for (int k = 1; k < N / 2; ++k) {
                const double pm = sqrt(sq((double)px) + sq((double)py) + sq((double)k));
                const double ph = sqrt(4.0 * sq((double)N / (double)L) *
                                       (sq(sin((double)px * pi / N)) + sq(sin((double)py * pi / N)) +
                                        sq(sin((double)k * pi / N))));
                np[(int)pm] += 2;
                pp[(int)pm] += 2.0 * ph;
            }
            for (int k = 0; k <= N / 2; k += N / 2) {
                const double pm = sqrt(sq((double)px) + sq((double)py) + sq((double)k));
                const double ph = sqrt(4.0 * sq((double)N / (double)L) *
                                       (sq(sin((double)px * pi / N)) + sq(sin((double)py * pi / N)) +
                                        sq(sin((double)k * pi / N))));
                np[(int)pm]++;
                pp[(int)pm] += ph;
            }
        }
    }
    for (int b = 0; b < nb; ++b) {
        if (np[b] > 0) pp[b] = pp[b] / np[b];
        out[b] = rescale_B * pp[b];
    }
    free(np);
    free(pp);
}

/* ------------------------------------------------------ initial conditions */

/* set_mode: Rayleigh amplitude, uniform phase, Bunch-Davies momentum
 * (src/initialize.cpp:53-89).  The first call fixes norm and hbterm. */
static void set_mode(ieo_state *st, double p2, double p2lat, double m2, double *field, double *deriv, int real)
{
    const ieo_config *c = &st->cfg;
    if (!st->sm_ready) {
        st->sm_norm = c->rescale_B * pow(c->L / sq(st->dx), 1.5);
        st->sm_hbterm = -st->hubble_init;
        st->sm_ready = 1;
    }
    const double ii = c->L / (2.0 * st->pi) * sqrt(p2lat);
    const double omega = (p2 + m2 > 0.0) ? sqrt(p2 + m2) : sqrt(p2);
    if (p2 + m2 <= 0.0 && st->sm_tachyonic == 0) {
        printf("Warning: Tachyonic mode(s) may be initialized inaccurately\n");
        st->sm_tachyonic = 1;
    }
    double rms = (omega > 0.0) ? st->sm_norm / sqrt(2.0 * omega) : 0.0;
    if (c->high_cutoff_index > 0 && (ii > c->high_cutoff_index || ii < c->low_cutoff_index)) rms = 0.0;
    const double amplitude = rms * sqrt(log(1.0 / ieo_rand_uniform(st)));
    const double phase = 2.0 * st->pi * ieo_rand_uniform(st);
    const double re = amplitude * cos(phase), im = amplitude * sin(phase);
    field[0] = re;
    field[1] = im;
    deriv[0] = omega * im + st->sm_hbterm * field[0];
    deriv[1] = -omega * re + st->sm_hbterm * field[1];
    if (real == 1) { field[1] = 0.0; deriv[1] = 0.0; }
}

/* initialize() and the k-space part of initializef() (src/initialize.cpp:93-191):
 * f, fd hold the FFTEASY-layout mode arrays afterwards, fny, fdny the Nyquist
 * planes.  Returns 0, or 1 for a Courant violation. */
int ieo_fill_modes(ieo_state *st)
{
    const ieo_config *c = &st->cfg;
    const int N = st->N;
    const double pi = st->pi;
    if (c->dt > st->dx / sqrt(3.0)) {
        printf("Time step too large. Courant condition violated: dt/dx = %f\n", c->dt / st->dx);
        return 1;
    }
    st->t0 = 0.0;
    st->hubble_init = sqrt(ieo_potential(st, c->initial_field) / 3.0);
    st->ad = st->hubble_init;
    memset(st->lstart, 0, st->gs * sizeof(int));

    double *f = st->f, *fd = st->fd, *fny = st->fny, *fdny = st->fdny;
    memset(f, 0, st->gs * sizeof(double));
    memset(fd, 0, st->gs * sizeof(double));
    const double dp2 = sq(2.0 * pi / c->L);
    const double m2 = 0.0;
    st->hubble_init = sqrt((0.5 * sq(c->initial_derivative) + ieo_potential(st, c->initial_field)) / 3.0);
    st->ad = st->hubble_init;

#define NY(i, j) ((size_t)(i) * 2 * N + 2 * (size_t)(j))
    for (int i = 0; i < N; i++) {
        const double px = (i <= N / 2 ? i : i - N);
        const int iconj = (i == 0 ? 0 : N - i);
        for (int j = 0; j < N; j++) {
            const double py = (j <= N / 2 ? j : j - N);
            for (int k = 1; k < N / 2; k++) {
                const double pz = k;
                const double p2lat = dp2 * (sq(px) + sq(py) + sq(pz));
                const double p2 = 4.0 * sq(N / c->L) *
                                  (sq(sin(px * pi / N)) + sq(sin(py * pi / N)) + sq(sin(pz * pi / N)));
                set_mode(st, p2, p2lat, m2, &f[IX(N, i, j, 2 * k)], &fd[IX(N, i, j, 2 * k)], 0);
            }
            if (j > N / 2 || (i > N / 2 && (j == 0 || j == N / 2))) {
                const int jconj = (j == 0 ? 0 : N - j);
                double p2lat = dp2 * (sq(px) + sq(py));
                double p2 = 4.0 * sq(N / c->L) * (sq(sin(px * pi / N)) + sq(sin(py * pi / N)));
                set_mode(st, p2, p2lat, m2, &f[IX(N, i, j, 0)], &fd[IX(N, i, j, 0)], 0);
                f[IX(N, iconj, jconj, 0)] = f[IX(N, i, j, 0)];
                f[IX(N, iconj, jconj, 1)] = -f[IX(N, i, j, 1)];
                fd[IX(N, iconj, jconj, 0)] = fd[IX(N, i, j, 0)];
                fd[IX(N, iconj, jconj, 1)] = -fd[IX(N, i, j, 1)];
                p2lat = dp2 * (sq(px) + sq(py) + sq(N / 2));
                p2 = 4.0 * sq(N / c->L) * (sq(sin(px * pi / N)) + sq(sin(py * pi / N)) + 1.0);
                set_mode(st, p2, p2lat, m2, &fny[NY(i, j)], &fdny[NY(i, j)], 0);
                fny[NY(iconj, jconj)] = fny[NY(i, j)];
                fny[NY(iconj, jconj) + 1] = -fny[NY(i, j) + 1];
                fdny[NY(iconj, jconj)] = fdny[NY(i, j)];
                fdny[NY(iconj, jconj) + 1] = -fdny[NY(i, j) + 1];
            } else if ((i == 0 || i == N / 2) && (j == 0 || j == N / 2)) {
                double p2lat = dp2 * (sq(px) + sq(py));
                double p2 = 4.0 * sq(N / c->L) * (sq(sin(px * pi / N)) + sq(sin(py * pi / N)));
                if (p2 > 0.0) set_mode(st, p2, p2lat, m2, &f[IX(N, i, j, 0)], &fd[IX(N, i, j, 0)], 1);
                p2lat = dp2 * (sq(px) + sq(py) + sq(N / 2));
                p2 = 4.0 * sq(N / c->L) * (sq(sin(px * pi / N)) + sq(sin(py * pi / N)) + 1.0);
                set_mode(st, p2, p2lat, m2, &fny[NY(i, j)], &fdny[NY(i, j)], 1);
            }
        }
    }
#undef NY
    f[0] = 0.0; f[1] = 0.0; fd[0] = 0.0; fd[1] = 0.0;
    return 0;
}

int ieo_initialize(ieo_state *st)
{
    const ieo_config *c = &st->cfg;
    const int N = st->N;
    double *f = st->f, *fd = st->fd, *fny = st->fny, *fdny = st->fdny;
    const int rc = ieo_fill_modes(st);
    if (rc) return rc;

    ieo_fftrn_d(f, fny, N, -1);
    ieo_fftrn_d(fd, fdny, N, -1);

    for (size_t id = 0; id < st->gs; ++id) {
        f[id] += c->initial_field;
        fd[id] += c->initial_derivative;
    }
    double vard = 0.0;
    for (size_t id = 0; id < st->gs; ++id) vard += sq(fd[id]);
    const double kin = 0.5 * vard / (double)st->gs;
    const double pe = ieo_potential_energy(st);
    const double ge = ieo_gradient_energy(st);
    st->hubble_init = sqrt((kin + pe + ge) / 3.0);
    st->ad = st->hubble_init;
    for (int q = 0; q < 6; ++q) {
        memset(st->h[q], 0, st->gs * sizeof(float));
        memset(st->hd[q], 0, st->gs * sizeof(float));
    }
    return 0;
}

/* ------------------------------------------------------------------ deltaN */

/* src/initialize.cpp:219-232 */
void ieo_initializeN(ieo_state *st)
{
    for (size_t id = 0; id < st->gs; ++id) {
        if (st->cfg.numerical_potential) st->lstart[id] -= 100;
        const double H = sqrt(ieo_potential(st, st->f[id]) / 3.0);
        st->fd[id] = st->fd[id] * pow(st->a, st->cfg.rescale_s - 1.0) / H;
        st->dn[id] = 0.0;
    }
}

/* src/evolution.cpp:1004-1012 */
void ieo_evolve_derivsN(ieo_state *st, double d)
{
    const size_t gs = st->gs;
#pragma omp parallel for schedule(static)
    for (size_t id = 0; id < gs; ++id)
        st->fd[id] += d * (-(3.0 - 0.5 * sq(st->fd[id])) * (st->fd[id] + pot_ratio_site(st, id)));
}

static int deltaN_active(const ieo_state *st, double v)
{
    if (st->cfg.monotonic == IEO_MONOTONIC) return fabs(v) > fabs(st->phiref);
    if (st->cfg.monotonic == IEO_ANTIMONOTONIC) return fabs(v) < fabs(st->phiref);
    return ieo_potential(st, v) > ieo_potential(st, st->phiref);
}

/* src/evolution.cpp:1015-1037 */
void ieo_evolve_fieldsN(ieo_state *st, double d)
{
    st->t += d;
    const size_t gs = st->gs;
#pragma omp parallel for schedule(static)
    for (size_t id = 0; id < gs; ++id) {
        if (deltaN_active(st, st->f[id])) {
            st->dn[id] += d;
            st->f[id] += d * st->fd[id];
        }
    }
    st->a += d * st->ad;
}

/* src/evolution.cpp:1271-1288: ordered search, first strict improvement wins */
double ieo_get_phiref(ieo_state *st)
{
    double fref = st->f[0];
    for (size_t id = 0; id < st->gs; ++id) {
        const double v = st->f[id];
        int better;
        if (st->cfg.monotonic == IEO_MONOTONIC) better = fabs(v) < fabs(fref);
        else if (st->cfg.monotonic == IEO_ANTIMONOTONIC) better = fabs(v) > fabs(fref);
        else better = ieo_potential(st, v) < ieo_potential(st, fref);
        if (better) fref = v;
    }
    return fref;
}

/* deltaN RHS (src/evolution.cpp:1065-1113). */
static void deltaN_rhs(ieo_state *st, const double *fs, const double *fds,
                       double *dfdt, double *dfddt, double *ddn, double *dadt)
{
    for (size_t id = 0; id < st->gs; ++id) {
        const double fh = fs[id], dh = fds[id];
        double pot = 0.0, dpot = 0.0;
        if (st->cfg.numerical_potential) {
            int nh = 1;
            eval_pot(st, fh, st->lstart[id], st->cfg.int_errN, &nh, &pot, &dpot);
            st->lstart[id] = nh;
        } else {
            eval_pot(st, fh, 1, 1, NULL, &pot, &dpot);
        }
        const double ratio = dpot / pot;
        int active;
        if (st->cfg.monotonic == IEO_MONOTONIC) active = fabs(fh) > fabs(st->phiref);
        else if (st->cfg.monotonic == IEO_ANTIMONOTONIC) active = fabs(fh) < fabs(st->phiref);
        else active = pot > st->phiref_potential;
        dfddt[id] = -(3.0 - 0.5 * sq(dh)) * (dh + ratio);
        if (active) { dfdt[id] = dh; ddn[id] = 1.0; }
        else { dfdt[id] = 0.0; ddn[id] = 0.0; }
    }
    *dadt = st->ad;
}

static void deltaN_stages(ieo_state *st, rk_scratch *sc, double h, const double *A, int ns, double *ka)
{
    deltaN_rhs(st, st->f, st->fd, sc->kf[0], sc->kfd[0], sc->kdn[0], &ka[0]);
    for (int stage = 1; stage < ns; ++stage) {
        for (size_t id = 0; id < st->gs; ++id) {
            double sf = 0.0, sfd = 0.0, sdn = 0.0;
            for (int q = 0; q < stage; ++q) {
                const double c = A[stage * ns + q];
                if (c == 0.0) continue;
                sf += c * sc->kf[q][id];
                sfd += c * sc->kfd[q][id];
                sdn += c * sc->kdn[q][id];
            }
            sc->ftmp[id] = st->f[id] + h * sf;
            sc->fdtmp[id] = st->fd[id] + h * sfd;
            sc->dntmp[id] = st->dn[id] + h * sdn;
        }
        deltaN_rhs(st, sc->ftmp, sc->fdtmp, sc->kf[stage], sc->kfd[stage], sc->kdn[stage], &ka[stage]);
    }
}

/* src/evolution.cpp:1176-1189 */
void ieo_rk4_stepN(ieo_state *st, double h)
{
    rk_scratch *sc = rk_get(st, 1);
    double ka[4] = {0};
    deltaN_stages(st, sc, h, &RK4A[0][0], 4, ka);
    for (size_t id = 0; id < st->gs; ++id) {
        double sf = 0.0, sfd = 0.0, sdn = 0.0;
        for (int q = 0; q < 4; ++q) {
            sf += RK4B[q] * sc->kf[q][id];
            sfd += RK4B[q] * sc->kfd[q][id];
            sdn += RK4B[q] * sc->kdn[q][id];
        }
        st->f[id] += h * sf;
        st->fd[id] += h * sfd;
        st->dn[id] += h * sdn;
    }
    for (int q = 0; q < 4; ++q) st->a += h * RK4B[q] * ka[q];
    st->t += h;
    rk_put(sc);
}

/* src/evolution.cpp:1191-1267 */
int ieo_rk45_stepN(ieo_state *st, double h, double hmax, double *h_next)
{
    rk_scratch *sc = rk_get(st, 1);
    double ka[7] = {0};
    deltaN_stages(st, sc, h, &DPA[0][0], 7, ka);
    double a5 = st->a, erra = 0.0;
    for (int q = 0; q < 7; ++q) {
        if (DPB[q] != 0.0) a5 += h * DPB[q] * ka[q];
        if (DPE[q] != 0.0) erra += h * DPE[q] * ka[q];
    }
    const double at = st->cfg.rk45_abs_tol, rt = st->cfg.rk45_rel_tol;
    double acc = 0.0;
    const size_t nvars = 3 * st->gs + 1;
    for (size_t id = 0; id < st->gs; ++id) {
        double y5f = st->f[id], y5fd = st->fd[id], y5dn = st->dn[id];
        double ef = 0.0, efd = 0.0, edn = 0.0;
        for (int q = 0; q < 7; ++q) {
            if (DPB[q] != 0.0) {
                y5f += h * DPB[q] * sc->kf[q][id];
                y5fd += h * DPB[q] * sc->kfd[q][id];
                y5dn += h * DPB[q] * sc->kdn[q][id];
            }
            if (DPE[q] != 0.0) {
                ef += h * DPE[q] * sc->kf[q][id];
                efd += h * DPE[q] * sc->kfd[q][id];
                edn += h * DPE[q] * sc->kdn[q][id];
            }
        }
        sc->ftmp[id] = y5f;
        sc->fdtmp[id] = y5fd;
        sc->dntmp[id] = y5dn;
        const double s1 = at + rt * fmax(fabs(st->f[id]), fabs(y5f));
        const double s2 = at + rt * fmax(fabs(st->fd[id]), fabs(y5fd));
        const double s3 = at + rt * fmax(fabs(st->dn[id]), fabs(y5dn));
        acc += sq(ef / s1) + sq(efd / s2) + sq(edn / s3);
    }
    const double sa = at + rt * fmax(fabs(st->a), fabs(a5));
    acc += sq(erra / sa);
    const double err_norm = sqrt(acc / (double)nvars);
    const double factor = rk45_factor(st, err_norm);
    int accepted = 0;
    if (err_norm <= 1.0) {
        double *tmp;
        tmp = st->f; st->f = sc->ftmp; sc->ftmp = tmp;
        tmp = st->fd; st->fd = sc->fdtmp; sc->fdtmp = tmp;
        tmp = st->dn; st->dn = sc->dntmp; sc->dntmp = tmp;
        st->a = a5;
        st->t += h;
        *h_next = clamp_rk45(st, h * factor, hmax);
        accepted = 1;
    } else {
        const double shrink = factor < 0.5 ? factor : 0.5;
        *h_next = clamp_rk45(st, h * (shrink > 0.1 ? shrink : 0.1), hmax);
    }
    rk_put(sc);
    return accepted;
}

/* ---------------------------------------------------------------- program */

/* Output files of one run, opened on first use like the reference's function
 * statics (src/output.cpp:206-221 and siblings). */
typedef struct {
    char dir[400];
    FILE *means, *vars, *vel, *sf, *energy, *cons, *spectra, *spectratimes, *gw, *gwdot;
    FILE *hist, *histtimes, *box, *box2d, *box2dot, *modes, *spectraN, *histN, *histtimesN, *box2dN;
    FILE *bis, *spectraLOG, *histLOG, *histtimesLOG;
    double *hist_buf;
} ieo_out;

static FILE *out_open(ieo_out *o, FILE **fp, const char *name)
{
    if (!*fp) {
        char path[600];
        snprintf(path, sizeof(path), "%s/%s.dat", o->dir, name);
        *fp = fopen(path, "w");
        if (!*fp) {
            fprintf(stderr, "Failed to open output file: %s\n", path);
            exit(1);
        }
    }
    return *fp;
}

static void out_close(ieo_out *o)
{
    FILE **all[] = {&o->means, &o->vars, &o->vel, &o->sf, &o->energy, &o->cons, &o->spectra,
                    &o->spectratimes, &o->gw, &o->gwdot, &o->hist, &o->histtimes, &o->box, &o->box2d,
                    &o->box2dot, &o->modes, &o->spectraN, &o->histN, &o->histtimesN, &o->box2dN,
                    &o->bis, &o->spectraLOG, &o->histLOG, &o->histtimesLOG};
    for (size_t q = 0; q < sizeof(all) / sizeof(all[0]); ++q)
        if (*all[q]) { fclose(*all[q]); *all[q] = NULL; }
    free(o->hist_buf);
    o->hist_buf = NULL;
}

/* meansvars (src/output.cpp:204-269; the post-inflation copy, :1575-1636, tests
 * against FLT_MAX). Returns 1 when the mean is not finite. */
static int out_meansvars(ieo_state *st, ieo_out *o, double big)
{
    FILE *m = out_open(o, &o->means, "means"), *v = out_open(o, &o->vars, "variance"),
         *w = out_open(o, &o->vel, "velocity");
    fprintf(m, "%f", st->t); fprintf(m, " %e", st->a);
    fprintf(w, "%f", st->t); fprintf(w, " %e", st->a);
    fprintf(v, "%f", st->t); fprintf(v, " %e", st->a);
    double av = 0., vel = 0., var = 0.;
    for (size_t id = 0; id < st->gs; ++id) {
        av += st->f[id];
        vel += st->fd[id];
        var += sq(st->f[id]);
    }
    av = av / (double)st->gs;
    vel = vel / (double)st->gs;
    vel = vel * pow(st->a, st->cfg.rescale_s - 1.0) * st->cfg.rescale_B;
    fprintf(m, " %e", av);
    fprintf(w, " %e", vel);
    fprintf(v, " %e", var - sq(av));
    if (av + big == av || (av != 0. && av / av != 1.)) {
        printf("Unstable solution developed. Scalar field not numerical at t=%f\n", st->t);
        fflush(m); fflush(v);
        return 1;
    }
    fprintf(m, "\n"); fprintf(v, "\n"); fprintf(w, "\n");
    return 0;
}

/* scale (src/output.cpp:272-293) */
static void out_scale(ieo_state *st, ieo_out *o)
{
    const double s = st->cfg.rescale_s, B = st->cfg.rescale_B;
    fprintf(out_open(o, &o->sf, "sf"), "%f %f %e %e\n", st->t, st->a,
            st->ad * B * pow(st->a, s - 2.0),
            sq(B) * pow(st->a, 2.0 * s - 2.0) * (st->ad2 + (s - 1.0) * sq(st->ad) / st->a));
}

/* energy (src/output.cpp:983-1026) */
static void out_energy(ieo_state *st, ieo_out *o)
{
    FILE *e = out_open(o, &o->energy, "energy"), *c = out_open(o, &o->cons, "conservation");
    const double B = st->cfg.rescale_B;
    double total = 0.;
    fprintf(e, "%f", st->t);
    fprintf(e, " %e", st->a);
    const double kin = ieo_kin_energy(st);
    total += kin;
    fprintf(e, " %e", kin * B);
    const double grad = ieo_gradient_energy(st);
    total += grad;
    fprintf(e, " %e", grad * B);
    const double pot = ieo_potential_energy(st);
    total += pot;
    fprintf(e, " %e", pot * B);
    fprintf(e, "\n");
    fprintf(c, "%e %e %e\n", st->t, st->a,
            3.0 * pow(st->a, 2.0 * st->cfg.rescale_s - 4.0) * sq(st->ad) / (total));
}

static void out_bins(FILE *fp, int nb, const double *p, const int *cnt, const double *val)
{
    for (int b = 0; b < nb; ++b) fprintf(fp, "%e %d %e\n", p[b], cnt[b], val[b]);
    fprintf(fp, "\n");
}

/* histogram of a field (src/output.cpp:1070-1135, 1220-1277) */
static void out_histogram(ieo_state *st, ieo_out *o, const double *g, FILE *hist, FILE *times)
{
    const int nbins = st->cfg.nbins;
    double bmin = g[0], bmax = bmin;
    for (size_t id = 0; id < st->gs; ++id) {
        bmin = g[id] < bmin ? g[id] : bmin;
        bmax = g[id] > bmax ? g[id] : bmax;
    }
    double df = (bmax - bmin) / (double)nbins;
    if (!isfinite(df) || df <= 0.0) df = 1.0;
    free(o->hist_buf);
    o->hist_buf = (double *)calloc(nbins > 0 ? nbins : 1, sizeof(double));
    int numpts = 0;
    for (size_t id = 0; id < st->gs; ++id) {
        int b = (int)((g[id] - bmin) / df);
        if (g[id] == bmax) b = nbins - 1;
        if (b >= 0 && b < nbins) { o->hist_buf[b]++; numpts++; }
    }
    if (numpts == 0) numpts = 1;
    for (int b = 0; b < nbins; ++b) fprintf(hist, "%e\n", o->hist_buf[b] / (double)numpts);
    fprintf(hist, "\n");
    fprintf(times, " %e %e", bmin, df);
    fprintf(times, "\n");
}

static int plane_index(int N) { return N < 64 ? 5 : 50; }

/* save(infrequent) for the inflation phase (src/output.cpp:1469-1503) */
static int program_save(ieo_state *st, ieo_out *o, int infrequent)
{
    const ieo_config *c = &st->cfg;
    const int N = st->N;
    const int staggered = c->integrator == IEO_LEAPFROG;
    if (staggered && st->t > 0.) ieo_evolve_fields(st, -.5 * c->dt * pow(st->astep, c->rescale_s - 1));
    if (out_meansvars(st, o, DBL_MAX)) return 1;
    out_scale(st, o);
    if (infrequent) {
        if (c->output_box3D) {
            FILE *b = out_open(o, &o->box, "box");
            for (size_t id = 0; id < st->gs; ++id) fprintf(b, "%.17g\n", st->f[id]);
            fprintf(b, "\n");
        }
        if (c->output_box2D) {
            const int i = plane_index(N);
            FILE *b = out_open(o, &o->box2d, "snapshots_2d_phi");
            for (int j = 0; j < N; ++j)
                for (int k = 0; k < N; ++k) fprintf(b, "%.17g\n", st->f[IX(N, i, j, k)]);
            fprintf(b, "\n");
            FILE *bd = out_open(o, &o->box2dot, "snapshots_2d_phidot");
            for (int j = 0; j < N; ++j)
                for (int k = 0; k < N; ++k) fprintf(bd, "%.17g\n", st->fd[IX(N, i, j, k)] * c->rescale_B);
            fprintf(bd, "\n");
        }
        if (c->output_energy) out_energy(st, o);
        if (c->output_spectra) {
            const int nb = ieo_spectrum_numbins(N);
            double *p = calloc(nb, sizeof(double)), *v = calloc(nb, sizeof(double));
            int *cnt = calloc(nb, sizeof(int));
            ieo_spectrum_f(st, p, cnt, v);
            out_bins(out_open(o, &o->spectra, "spectra"), nb, p, cnt, v);
            fprintf(out_open(o, &o->spectratimes, "spectratimes"), "%f %e\n", st->t, st->a);
            if (c->calculate_sigw) {
                const int ngw = (int)(sqrt(3.0) * (N / 2)) + 1;
                double *pg = calloc(ngw, sizeof(double)), *vg = calloc(ngw, sizeof(double));
                int *cg = calloc(ngw, sizeof(int));
                ieo_gw_spectrum(st, 0, pg, cg, vg);
                out_bins(out_open(o, &o->gw, "spectraGW"), ngw, pg, cg, vg);
                ieo_gw_spectrum(st, 1, pg, cg, vg);
                out_bins(out_open(o, &o->gwdot, "spectraGWdot"), ngw, pg, cg, vg);
                free(pg); free(vg); free(cg);
            }
            free(p); free(v); free(cnt);
        }
        if (c->output_histogram) {
            FILE *ht = out_open(o, &o->histtimes, "histogramtimes");
            fprintf(ht, "%f", st->t);
            fprintf(ht, " %e", st->a);
            out_histogram(st, o, st->f, out_open(o, &o->hist, "histogram"), ht);
        }
    }
    if (staggered && st->t > 0.) ieo_evolve_fields(st, .5 * c->dt * pow(st->astep, c->rescale_s - 1));
    return 0;
}

/* RK45 accept/reject driver (src/evolution.cpp:346-362). Returns the next step or
 * a negative value on failure. */
static double rk45_drive(ieo_state *st, double h, double hmax, int phase)
{
    int accepted = 0, attempts = 0;
    while (!accepted) {
        double hs = h;
        accepted = phase == 1 ? ieo_rk45_stepN(st, h, hmax, &hs)
                 : phase == 2 ? ieo_post_rk45_step(st, h, hmax, &hs) : ieo_rk45_step(st, h, hmax, &hs);
        h = hs;
        ++attempts;
        if (!accepted && (attempts > 25 || h <= st->cfg.rk45_min_dt * (1.0 + 1e-12))) {
            fprintf(stderr, "RK45 failed to converge (attempts=%d, h=%e)\n", attempts, h);
            return -1.0;
        }
    }
    return h;
}

static double rk45_hmax(const ieo_config *c, double base)
{
    const double lo = c->rk45_min_dt > 2.0 * base ? c->rk45_min_dt : 2.0 * base;
    return c->rk45_max_dt < lo ? c->rk45_max_dt : lo;
}

/* run_evolution_loop (src/evolution.cpp:909-997) */
static int program_inflation(ieo_state *st, ieo_out *o)
{
    const ieo_config *c = &st->cfg;
    int n = 0;
#define REPORT()                                                                      \
    do {                                                                              \
        if (n % c->output_freq == 0 && st->a < c->af)                                 \
            if (program_save(st, o, (n % c->output_infrequent_freq == 0) ? 1 : 0))   \
                return 1;                                                             \
    } while (0)
    if (c->integrator == IEO_LEAPFROG) {
        ieo_evolve_fields(st, 0.5 * c->dt);
        while (st->a <= c->af) {
            const double d = rescaled_step(st, st->astep);
            ieo_evolve_derivs(st, d);
            ieo_evolve_fields(st, d);
            n++;
            REPORT();
            st->astep = st->a;
        }
    } else if (c->integrator == IEO_RK4) {
        while (st->a <= c->af) {
            ieo_rk4_step(st, rescaled_step(st, st->astep));
            n++;
            REPORT();
            st->astep = st->a;
        }
    } else {
        double h = clamp_rk45(st, rescaled_step(st, st->astep), c->rk45_max_dt);
        while (st->a <= c->af) {
            const double hmax = rk45_hmax(c, rescaled_step(st, st->astep));
            h = clamp_rk45(st, h, hmax);
            h = rk45_drive(st, h, hmax, 0);
            if (h < 0) return 2;
            n++;
            REPORT();
            st->astep = st->a;
        }
    }
#undef REPORT
    if (program_save(st, o, 1)) return 1;
    if (c->integrator == IEO_LEAPFROG) ieo_evolve_fields(st, -0.5 * rescaled_step(st, st->astep));
    {
        const int nb = (int)(sqrt(3.0) * (st->N / 2)) + 1;
        double *m = calloc(nb, sizeof(double));
        ieo_modes(st->N, c->L, c->rescale_B, m);
        FILE *fp = out_open(o, &o->modes, "modes");
        for (int b = 0; b < nb; ++b) fprintf(fp, "%e\n", m[b]);
        fprintf(fp, "\n");
        free(m);
    }
    /* the rest of save_last (src/output.cpp:1505-1519) */
    if (c->output_bispectrum) {
        const int nb = (int)(sqrt(3.0) * (st->N / 2)) + 1;
        double *p = calloc(nb, sizeof(double)), *re = calloc(nb, sizeof(double)), *im = calloc(nb, sizeof(double));
        int *cnt = calloc(nb, sizeof(int));
        ieo_bispectrum(st, p, cnt, re, im);
        FILE *fp = out_open(o, &o->bis, "bispectra");
        for (int k = 0; k < nb; ++k) fprintf(fp, "%e %d %e %e\n", p[k], cnt[k], re[k], im[k]);
        fprintf(fp, "\n");
        free(p); free(re); free(im); free(cnt);
    }
    if (c->perform_deltaN && c->output_LOG) {
        const int nb = ieo_spectrum_numbins(st->N);
        double *p = calloc(nb, sizeof(double)), *v = calloc(nb, sizeof(double));
        int *cnt = calloc(nb, sizeof(int));
        ieo_spectrum_log(st, p, cnt, v);
        out_bins(out_open(o, &o->spectraLOG, "spectraLOG"), nb, p, cnt, v);
        free(p); free(v); free(cnt);
        double *freq = calloc(c->nbins > 0 ? c->nbins : 1, sizeof(double)), bmin, df;
        FILE *ht = out_open(o, &o->histtimesLOG, "histogramtimesLOG");
        FILE *h = out_open(o, &o->histLOG, "histogramLOG");
        fprintf(ht, "%f %e", st->t, st->a);
        ieo_histogram_log(st, &bmin, &df, freq);
        for (int b = 0; b < c->nbins; ++b) fprintf(h, "%e\n", freq[b]);
        fprintf(h, "\n");
        fprintf(ht, " %e %e\n", bmin, df);
        free(freq);
    }
    return 0;
}

/* saveN's mean (src/output.cpp:1527-1537): a float accumulator over the sites
 * with deltaN <= cutoff in lattice order, float += double, then / gridsize. */
float ieo_float_mean(const double *dn, size_t gs, double cutoff)
{
    float Nmean = 0.;
    for (size_t id = 0; id < gs; ++id)
        if (dn[id] <= cutoff) Nmean += dn[id];
    return Nmean / (int)gs;
}

int ieo_deltaN_zeta_map(ieo_state *st, float *nmean_out)
{
    const ieo_config *c = &st->cfg;
    st->Ne = 0.0;
    ieo_initializeN(st);
    ieo_evolve_fieldsN(st, 0.5 * c->dN);
    st->phiref = c->use_phiref_manual ? c->phiref_manual_value : ieo_get_phiref(st);
    st->phiref_potential = ieo_potential(st, st->phiref);
    int n = 0;
    while (st->Ne <= c->Nend) {
        ieo_evolve_derivsN(st, c->dN);
        ieo_evolve_fieldsN(st, c->dN);
        st->Ne += c->dN;
        ++n;
    }
    ieo_evolve_fieldsN(st, -.5 * c->dN);
    const double cutoff = c->Nend - c->dN;
    const float Nmean = ieo_float_mean(st->dn, st->gs, cutoff);
    for (size_t id = 0; id < st->gs; ++id) st->dn[id] -= Nmean;
    for (size_t id = 0; id < st->gs; ++id)
        if (st->dn[id] > (cutoff - Nmean)) st->dn[id] = 0;
    double mean = 0.0;
    for (size_t id = 0; id < st->gs; ++id) mean += st->dn[id];
    mean /= (double)st->gs;
    for (size_t id = 0; id < st->gs; ++id) st->dn[id] -= mean;
    if (nmean_out) *nmean_out = Nmean;
    return n;
}

/* saveN (src/output.cpp:1522-1560) with spectraN (:1140-1169) */
static void program_saveN(ieo_state *st, ieo_out *o)
{
    const ieo_config *c = &st->cfg;
    const int N = st->N;
    const int staggered = c->deltaN_integrator == IEO_LEAPFROG;
    if (staggered && st->t > 0.) ieo_evolve_fieldsN(st, -.5 * c->dN);
    const double cutoff = c->Nend - (staggered ? c->dN : 0.0);
    const float Nmean = ieo_float_mean(st->dn, st->gs, cutoff);
    for (size_t id = 0; id < st->gs; ++id) st->dn[id] -= Nmean;
    {
        FILE *ht = out_open(o, &o->histtimesN, "histogramtimesN");
        fprintf(ht, "%f %e", st->t, st->a);
        out_histogram(st, o, st->dn, out_open(o, &o->histN, "histogramN"), ht);
    }
    if (c->output_box2D) {
        const int i = plane_index(N);
        FILE *b = out_open(o, &o->box2dN, "snapshots_2d_deltaN");
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) fprintf(b, "%.17g\n", st->dn[IX(N, i, j, k)]);
        fprintf(b, "\n");
    }
    for (size_t id = 0; id < st->gs; ++id)
        if (st->dn[id] > (cutoff - Nmean)) st->dn[id] = 0;
    if (c->output_spectra) {
        double mean = 0.0;
        for (size_t id = 0; id < st->gs; ++id) mean += st->dn[id];
        mean /= (double)st->gs;
        for (size_t id = 0; id < st->gs; ++id) st->dn[id] -= mean;
        const int nb = ieo_spectrum_numbins(N);
        double *p = calloc(nb, sizeof(double)), *v = calloc(nb, sizeof(double));
        int *cnt = calloc(nb, sizeof(int));
        ieo_fftrn_d(st->dn, st->fny, N, 1);
        ieo_bin_spectrum(N, c->L, st->dn, st->fny, norm_scalar(st), p, cnt, v);
        ieo_fftrn_d(st->dn, st->fny, N, -1);
        out_bins(out_open(o, &o->spectraN, "spectraN"), nb, p, cnt, v);
        free(p); free(v); free(cnt);
    }
    if (staggered && st->t > 0.) ieo_evolve_fieldsN(st, .5 * c->dN);
}

/* run_deltaN_loop (src/evolution.cpp:1291-1375) */
static int program_deltaN(ieo_state *st, ieo_out *o)
{
    const ieo_config *c = &st->cfg;
    st->Ne = 0.0;
    ieo_initializeN(st);
    const int staggered = c->deltaN_integrator == IEO_LEAPFROG;
    if (staggered) ieo_evolve_fieldsN(st, 0.5 * c->dN);
    st->phiref = c->use_phiref_manual ? c->phiref_manual_value : ieo_get_phiref(st);
    st->phiref_potential = ieo_potential(st, st->phiref);
    if (c->deltaN_integrator == IEO_LEAPFROG) {
        while (st->Ne <= c->Nend) {
            ieo_evolve_derivsN(st, c->dN);
            ieo_evolve_fieldsN(st, c->dN);
            st->Ne += c->dN;
        }
    } else if (c->deltaN_integrator == IEO_RK4) {
        while (st->Ne <= c->Nend) {
            ieo_rk4_stepN(st, c->dN);
            st->Ne += c->dN;
        }
    } else {
        double h = clamp_rk45(st, c->dN, c->rk45_max_dt);
        while (st->Ne <= c->Nend) {
            const double hmax = rk45_hmax(c, c->dN);
            h = clamp_rk45(st, h, hmax);
            h = rk45_drive(st, h, hmax, 1);
            if (h < 0) return 2;
            st->Ne += h;
        }
    }
    program_saveN(st, o);
    if (staggered) ieo_evolve_fieldsN(st, -0.5 * c->dN);
    return 0;
}

/* save_post_inflation (src/output.cpp:1839-1866): the same outputs as save()
 * under post_inflation/, with the float factor of spectraGWdot_post_inflation. */
static int program_save_post(ieo_state *st, ieo_out *o, int infrequent)
{
    const ieo_config *c = &st->cfg;
    const int N = st->N;
    const int staggered = c->post_integrator == IEO_LEAPFROG;
    if (staggered && st->t > 0.) ieo_evolve_fields(st, -.5 * c->dt_post);
    if (out_meansvars(st, o, FLT_MAX)) return 1;
    out_scale(st, o);
    if (infrequent) {
        if (c->output_spectra) {
            const int nb = ieo_spectrum_numbins(N);
            double *p = calloc(nb, sizeof(double)), *v = calloc(nb, sizeof(double));
            int *cnt = calloc(nb, sizeof(int));
            ieo_spectrum_f(st, p, cnt, v);
            out_bins(out_open(o, &o->spectra, "spectra"), nb, p, cnt, v);
            fprintf(out_open(o, &o->spectratimes, "spectratimes"), "%f %e\n", st->t, st->a);
            const int ngw = (int)(sqrt(3.0) * (N / 2)) + 1;
            double *pg = calloc(ngw, sizeof(double)), *vg = calloc(ngw, sizeof(double));
            int *cg = calloc(ngw, sizeof(int));
            ieo_gw_spectrum(st, 0, pg, cg, vg);
            out_bins(out_open(o, &o->gw, "spectraGW"), ngw, pg, cg, vg);
            ieo_gw_spectrum(st, 2, pg, cg, vg);
            out_bins(out_open(o, &o->gwdot, "spectraGWdot"), ngw, pg, cg, vg);
            free(pg); free(vg); free(cg);
            free(p); free(v); free(cnt);
        }
        if (c->output_histogram) {
            FILE *ht = out_open(o, &o->histtimes, "histogramtimes");
            fprintf(ht, "%f", st->t);
            fprintf(ht, " %e", st->a);
            out_histogram(st, o, st->f, out_open(o, &o->hist, "histogram"), ht);
        }
    }
    if (staggered && st->t > 0.) ieo_evolve_fields(st, .5 * c->dt_post);
    return 0;
}

/* run_post_inflation_loop (src/evolution.cpp:1568-1651) */
static int program_post(ieo_state *st, ieo_out *o)
{
    const ieo_config *c = &st->cfg;
    ieo_post_initialize(st);
    if (program_save_post(st, o, 1)) return 1;
    int n = 0;
#define REPORT_POST()                                                                     \
    do {                                                                                  \
        if (n % c->output_freq == 0)                                                      \
            if (program_save_post(st, o, (n % c->output_infrequent_freq == 0) ? 1 : 0))  \
                return 1;                                                                 \
    } while (0)
    if (c->post_integrator == IEO_LEAPFROG) {
        ieo_evolve_fields(st, 0.5 * c->dt_post);
        while (st->a <= c->af_post) {
            ieo_post_evolve_derivs(st, c->dt_post);
            ieo_evolve_fields(st, c->dt_post);
            n++;
            REPORT_POST();
        }
    } else if (c->post_integrator == IEO_RK4) {
        while (st->a <= c->af_post) {
            ieo_post_rk4_step(st, c->dt_post);
            n++;
            REPORT_POST();
        }
    } else {
        double h = clamp_rk45(st, c->dt_post, c->rk45_max_dt);
        while (st->a <= c->af_post) {
            const double hmax = rk45_hmax(c, c->dt_post);
            h = clamp_rk45(st, h, hmax);
            h = rk45_drive(st, h, hmax, 2);
            if (h < 0) return 2;
            n++;
            REPORT_POST();
        }
    }
#undef REPORT_POST
    return program_save_post(st, o, 1);
}

int ieo_program(const ieo_config *cfg, const char *results_dir)
{
    ieo_state *st = ieo_create(cfg);
    if (!st) return 10;
    ieo_out o;
    memset(&o, 0, sizeof(o));
    snprintf(o.dir, sizeof(o.dir), "%s", results_dir);
    mkdir(results_dir, 0777);
    int rc = ieo_initialize(st);
    if (!rc) rc = program_save(st, &o, 1);
    if (!rc) { st->t = st->t0; rc = program_inflation(st, &o); }
    if (!rc && cfg->perform_deltaN) rc = program_deltaN(st, &o);
    out_close(&o);
    if (!rc && cfg->post_inflation) {
        snprintf(o.dir, sizeof(o.dir), "%s/post_inflation", results_dir);
        mkdir(o.dir, 0777);
        rc = program_post(st, &o);
        out_close(&o);
    }
    ieo_destroy(st);
    return rc;
}
