/*
 * ie_oracle.h -- CPU restatement of the InflationEasy hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the gfx950
 * engine: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the CPU baseline.  The engine never links,
 * loads or falls back to it.
 *
 * It restates, in plain C with an explicit context instead of globals, the
 * reference algorithm of caravangelo/inflation-easy for:
 *   - initial conditions     src/initialize.cpp:18-232
 *   - potential              src/potential.cpp:11-204
 *   - leapfrog / RK stepping src/evolution.cpp:21-284,286-997
 *   - deltaN loop            src/evolution.cpp:999-1376
 *   - post-inflation loop    src/evolution.cpp:79-106,520-592,1378-1651,
 *                            src/initialize.cpp:246-291
 *   - FFT (FFTEASY layout)   src/ffteasy.hpp:21-204
 *   - outputs                src/output.cpp:63-151,156-164,204-293,297-406,
 *                            434-648,674-738,983-1026,1070-1135,1140-1277,1469-1560
 * The floating-point operation order of the reference is kept so that, built with
 * gcc on x86-64 (no FMA contraction), the restatement reproduces the reference
 * build bit for bit; tests/test_oracle_golden.py checks that against fixtures
 * written by the reference itself.
 */
#ifndef IE_ORACLE_H
#define IE_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { IEO_LEAPFROG = 0, IEO_RK4 = 1, IEO_RK45 = 2 };
enum { IEO_GENERIC = 0, IEO_MONOTONIC = 1, IEO_ANTIMONOTONIC = -1 };

/* Run configuration: the reference's compile-time toggles (src/parameters.h:22-70)
 * plus its run-time parameters (src/runtime_parameters.cpp:17-112). */
typedef struct ieo_config {
    int N;
    int numerical_potential;   /* 1: tabulated, 0: analytic hilltop */
    int calculate_sigw;
    int perform_deltaN;
    int monotonic;             /* IEO_MONOTONIC / IEO_ANTIMONOTONIC / IEO_GENERIC */
    int seed;
    double rescale_s, af, V0, rescale_B, ns;
    double initial_field, initial_derivative, L, dt;
    int integrator, deltaN_integrator;
    double rk45_abs_tol, rk45_rel_tol, rk45_min_dt, rk45_max_dt, rk45_safety;
    int output_freq, output_infrequent_freq;
    double dN, Nend;
    int use_phiref_manual;
    double phiref_manual_value;
    double high_cutoff_index, low_cutoff_index;
    int forcing_cutoff;
    int output_spectra, output_histogram, output_energy, output_box3D, output_box2D;
    int nbins;
    int int_err, int_errN;
    /* tabulated potential (descending field grid) */
    int tab_n;
    const double *tab_x, *tab_v, *tab_dv;
    /* post-inflation phase (src/parameters.h:31,148-155; off unless set) */
    int post_inflation;
    double horizon_factor, omega, dt_post, af_post;
    int post_integrator;
    /* outputs off by default (src/runtime_parameters.cpp:97-101) */
    int output_bispectrum, output_LOG;
    double eta_log;
} ieo_config;

typedef struct ieo_state ieo_state;

/* Fill cfg with the reference defaults for the given potential kind
 * (src/runtime_parameters.cpp:17-112; af = 2N, rescale_B = sqrt(V0)). */
void ieo_config_defaults(ieo_config *cfg, int N, int numerical_potential);
/* Derived quantities after overrides (src/runtime_parameters.cpp:342-353). */
void ieo_config_finalize(ieo_config *cfg, int rescale_B_set, int af_set);

ieo_state *ieo_create(const ieo_config *cfg);
void ieo_destroy(ieo_state *st);

/* Field access: which = 0 f, 1 fd, 2 deltaN (double, N^3);
 * 3 f Nyquist plane, 4 fd Nyquist plane (double, N*2N). */
double *ieo_field(ieo_state *st, int which);
/* tensors: comp 0..5 h_ij, 6..11 hdot_ij (float, N^3) */
float *ieo_tensor(ieo_state *st, int comp);
int *ieo_lstart(ieo_state *st);
/* background: t, a, ad, ad2, astep, Ne, phiref, hubble_init */
void ieo_get_background(const ieo_state *st, double out[8]);
void ieo_set_background(ieo_state *st, const double in[8]);

/* RNG (src/initialize.cpp:18-31) */
double ieo_rand_uniform(ieo_state *st);
/* Potential (src/potential.cpp) */
int ieo_find_bracketing_index(const ieo_state *st, double v, int start, int *err);
double ieo_potential(const ieo_state *st, double v);
double ieo_potential_derivative_value(const ieo_state *st, double v);
/* Initial conditions: initialize(), initializef(), initializeGW()
 * (src/initialize.cpp:40-49,93-242).  Returns 0 or an error code. */
int ieo_initialize(ieo_state *st);
/* The k-space part only (src/initialize.cpp:93-191): afterwards f, fd hold the
 * FFTEASY-layout mode arrays and fields 3, 4 the Nyquist planes. */
int ieo_fill_modes(ieo_state *st);

/* Energies (src/evolution.cpp:130-145, src/potential.cpp:151-171) */
double ieo_gradient_energy(ieo_state *st);
double ieo_kin_energy(ieo_state *st);
double ieo_potential_energy(ieo_state *st);

/* Leapfrog pieces (src/evolution.cpp:150-284) */
void ieo_evolve_derivs(ieo_state *st, double d);
void ieo_evolve_fields(ieo_state *st, double d);
/* n leapfrog steps of the main loop body (src/evolution.cpp:937-945) without
 * outputs; returns the number of steps taken (stops when a > af). */
int ieo_leapfrog_steps(ieo_state *st, int nsteps);
/* one RK4 step / one accepted RK45 step (src/evolution.cpp:760-904) */
void ieo_rk4_step(ieo_state *st, double h);
int ieo_rk45_step(ieo_state *st, double h, double hmax, double *h_next);

/* FFTEASY-layout real FFT in place (src/ffteasy.hpp:127-190). */
void ieo_fftrn_d(double *f, double *fny, int N, int forward);
void ieo_fftrn_f(float *f, float *fny, int N, int forward);

/* Isotropic spectrum of an FFTEASY-layout transform (src/output.cpp:63-151).
 * p, counts, vals have numbins = int(1.73205*(N/2))+1 entries; vals = norm1*<|F|^2>. */
int ieo_spectrum_numbins(int N);
void ieo_bin_spectrum(int N, double L, const double *ffield, const double *fny,
                      double norm1, double *p, int *counts, double *vals);
/* spectraf without the cutoff branch: forward FFT, bin, inverse FFT of f. */
void ieo_spectrum_f(ieo_state *st, double *p, int *counts, double *vals);
/* GW TT spectrum (src/output.cpp:466-625). hdot=0: h_ij, 1: hdot_ij (DoubleThenFloat),
 * 2: hdot_ij with the post-inflation float factor (FloatMultiply, :662-669). */
void ieo_gw_spectrum(ieo_state *st, int hdot, double *p, int *counts, double *vals);
/* The k-space forcing cutoff of spectraf (src/output.cpp:333-404): with f already
 * forward-transformed in place (FFTEASY layout, f and fny), transform fd, zero the
 * modes with |n| > high_cutoff_index or < low_cutoff_index with the reference's
 * mirror rule, transform fd back.  The caller transforms f back. */
void ieo_cutoff_modes(ieo_state *st);
/* bispectraf (src/output.cpp:740-911): equilateral bispectrum of f, numbins =
 * int(sqrt(3)*(N/2))+1 entries of p, counts, re, im (norm1 applied); f is
 * mean-subtracted, transformed, transformed back and the mean added, in place. */
void ieo_bispectrum(ieo_state *st, double *p, int *counts, double *re, double *im);
/* build_deltaN_log (src/output.cpp:167-198) into the deltaN array; returns 0 when
 * the mean velocity factor is not usable (map set to 0). */
int ieo_deltaN_log(ieo_state *st);
/* spectraLOG (src/output.cpp:1280-1308): the map, then its spectrum (in place). */
void ieo_spectrum_log(ieo_state *st, double *p, int *counts, double *vals);
/* histogramsLOG (src/output.cpp:1312-1403): bins of the map over the sites with
 * a positive logarithm argument; freq[nbins] = counts / points, *bmin, *df as
 * written to histogramtimesLOG.dat.  Returns 0 when the factor is not usable. */
int ieo_histogram_log(ieo_state *st, double *bmin, double *df, double *freq);
/* modes.dat (src/output.cpp:674-738) */
void ieo_modes(int N, double L, double rescale_B, double *out);

/* deltaN (src/initialize.cpp:219-232, src/evolution.cpp:1004-1037,1271-1375) */
void ieo_initializeN(ieo_state *st);
void ieo_evolve_derivsN(ieo_state *st, double d);
void ieo_evolve_fieldsN(ieo_state *st, double d);
double ieo_get_phiref(ieo_state *st);
/* run_deltaN_loop with the leapfrog integrator (src/evolution.cpp:1296-1328) and
 * saveN's zeta map without its outputs (src/output.cpp:1523-1553): the float
 * mean over finished sites subtracted, unfinished sites zeroed, then the FP64
 * mean removed as spectraN's build_deltaN_linear does (src/output.cpp:156-164);
 * st->dn holds the map spectraN transforms.  Returns the deltaN steps taken,
 * *nmean the float mean. */
int ieo_deltaN_zeta_map(ieo_state *st, float *nmean);
/* saveN's float mean of a deltaN map over the sites <= cutoff (src/output.cpp:1527-1537). */
float ieo_float_mean(const double *dn, size_t gs, double cutoff);
void ieo_rk4_stepN(ieo_state *st, double h);
int ieo_rk45_stepN(ieo_state *st, double h, double hmax, double *h_next);

/* post-inflation (src/initialize.cpp:246-291 without its save; src/evolution.cpp:
 * 1437-1564; the RK steps with the post-inflation right-hand side :520-592). */
void ieo_post_initialize(ieo_state *st);
void ieo_post_evolve_derivs(ieo_state *st, double d);
void ieo_post_rk4_step(ieo_state *st, double h);
int ieo_post_rk45_step(ieo_state *st, double h, double hmax, double *h_next);

/* Whole program (src/main.cpp:103-146; the post-inflation phase when
 * cfg->post_inflation, into results_dir/post_inflation), writing the
 * results/ files of the reference under results_dir.  Returns 0 or an error. */
int ieo_program(const ieo_config *cfg, const char *results_dir);

/* CPU-baseline helpers: OpenMP thread count used by the per-site loops. */
int ieo_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
