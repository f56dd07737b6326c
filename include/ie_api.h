/*
 * ie_api.h -- C ABI of the gfx950 lattice engine for InflationEasy.
 *
 * The reference (caravangelo/inflation-easy) has no FFI: its hot path is a set of
 * free functions over global state (src/main.h:138-205, src/ffteasy.hpp:195-204,
 * file-static back-ends in src/output.cpp).  Each entry point below is the
 * handle-based replacement of one of those functions; the comment names the
 * reference function (file:line) whose arguments, meaning and error behaviour it
 * keeps.  State lives on the GPU and is owned by the handle; host pointers are
 * borrowed for the duration of a call; every call returns an ie_status and leaves
 * a message retrievable with ie_last_error().  A handle is thread-compatible (one
 * host thread at a time), not thread-safe.
 */
#ifndef IE_API_H
#define IE_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ie_status {
    IE_OK = 0,
    IE_ERR_ARG = 1,          /* invalid argument or unsupported configuration */
    IE_ERR_COURANT = 2,      /* dt > dx/sqrt(3) (src/initialize.cpp:94-97) */
    IE_ERR_TABLE_RANGE = 3,  /* field outside the tabulated potential (src/potential.cpp:19-27) */
    IE_ERR_NONFINITE = 4,    /* <phi> not finite (src/output.cpp:251-258) */
    IE_ERR_RK45 = 5,         /* RK45 failed to converge (src/evolution.cpp:974-979) */
    IE_ERR_OOM = 6,          /* device allocation failed */
    IE_ERR_HIP = 7,          /* HIP runtime error */
    IE_ERR_STATE = 8         /* call not valid in the current phase */
} ie_status;

enum { IE_FP64 = 0, IE_SPLIT = 1 };                       /* field storage */
enum { IE_LEAPFROG = 0, IE_RK4 = 1, IE_RK45 = 2 };         /* src/parameters.h:109-113 */
enum { IE_GENERIC = 0, IE_MONOTONIC = 1, IE_ANTIMONOTONIC = -1 }; /* src/parameters.h:60-66 */
enum { IE_FIELD_F = 0, IE_FIELD_FD = 1, IE_FIELD_DELTAN = 2 };

/* One run: the reference's compile-time toggles (src/parameters.h:22-70) become
 * per-run fields, plus its run-time parameters (src/runtime_parameters.cpp:17-112)
 * that the hot path reads.  N must be a power of two, 8 <= N <= 1024. */
typedef struct ie_config {
    int N;
    int precision;             /* IE_FP64 or IE_SPLIT */
    int numerical_potential;   /* 1 tabulated, 0 analytic hilltop */
    int calculate_sigw;
    int monotonic;             /* deltaN stopping criterion */
    int integrator;            /* inflation integrator */
    int deltaN_integrator;
    double rescale_s, af, V0, rescale_B, ns, L, dt;
    double rk45_abs_tol, rk45_rel_tol, rk45_min_dt, rk45_max_dt, rk45_safety;
    double dN, Nend;
    int use_phiref_manual;
    double phiref_manual_value;
    int tab_n;                 /* tabulated potential, descending field grid */
    const double *tab_x, *tab_v, *tab_dv;
} ie_config;

/* Background scalars (src/main.cpp:324-326) plus loop bookkeeping. */
typedef struct ie_background {
    double t, a, ad, ad2, astep, Ne, phiref, hubble_init;
    double dt_last;            /* step of the last kick or drift (dt astep^{s-1} in the loop) */
    int64_t steps;             /* leapfrog steps taken by ie_leapfrog_steps */
    int32_t done;              /* a > af reached */
    int32_t error;             /* ie_status raised inside a kernel (table range) */
} ie_background;

typedef struct ie_run ie_run;

/* --- lifecycle -------------------------------------------------------------- */
/* Allocate device state for one run (replaces the global arrays of
 * src/main.cpp:15-37 and the checks of initialize(), src/initialize.cpp:93-114). */
ie_status ie_create(const ie_config *cfg, ie_run **out);
ie_status ie_destroy(ie_run *run);
/* The message of the last failed call on run, or with run = NULL of the last
 * failed call on this thread that had no run (ie_create, ie_spectrum_noalloc). */
const char *ie_last_error(const ie_run *run);
/* Bytes of device memory the run holds (admission by HBM footprint). */
size_t ie_footprint_bytes(const ie_config *cfg);

/* --- state transfer --------------------------------------------------------- */
/* Upload real-space f, fd (N^3 doubles each), as initializef() leaves them
 * (src/initialize.cpp:118-214).  In IE_SPLIT mode the FP64 volume averages become
 * the homogeneous mode and the rest is stored in FP32. */
ie_status ie_upload_fields(ie_run *run, const double *f, const double *fd);
/* Generate real-space f, fd from FFTEASY-layout mode arrays (the k-space fill of
 * initializef, src/initialize.cpp:142-191) with the inverse transform on the GPU
 * (src/initialize.cpp:194-200): fk, fdk hold N^3 doubles (kz < N/2, interleaved
 * complex), fny, fdny the N x 2N Nyquist planes; phi0, dphi0 are added after. */
ie_status ie_upload_modes(ie_run *run, const double *fk, const double *fny,
                          const double *fdk, const double *fdny, double phi0, double dphi0);
/* Download a field as FP64 (N^3); IE_FIELD_F reconstitutes phibar + dphi in split mode. */
ie_status ie_download_field(ie_run *run, int which, double *out);
/* One i-plane (N^2 doubles, j-major) of a field as ie_download_field would give
 * it: the 2-D snapshots (box2d, box2dot, box2dN; src/output.cpp:933-981,
 * 1195-1218) without moving the whole lattice. */
ie_status ie_download_plane(ie_run *run, int which, int i, double *out);
/* Upload a field as FP64 (N^3): IE_FIELD_DELTAN the deltaN map; IE_FIELD_F and
 * IE_FIELD_FD the FP64 state of the deltaN and post-inflation phases. */
ie_status ie_upload_field(ie_run *run, int which, const double *src);
/* Tensors: comp 0..5 h_ij, 6..11 hdot_ij (N^3 floats each). */
ie_status ie_upload_tensor(ie_run *run, int comp, const float *h);
ie_status ie_download_tensor(ie_run *run, int comp, float *out);
ie_status ie_get_background(ie_run *run, ie_background *out);
ie_status ie_set_background(ie_run *run, const ie_background *in);
/* The run's configuration as ie_create took it (table pointers: the run's own
 * host copies, valid while the run lives). */
ie_status ie_get_config(const ie_run *run, ie_config *out);

/* --- initial conditions ----------------------------------------------------- */
/* Inputs of initializef() (src/initialize.cpp:118-214) that are not in ie_config. */
typedef struct ie_ic_params {
    int seed;                  /* Park-Miller seed (src/initialize.cpp:18-31) */
    double initial_field, initial_derivative;
    double high_cutoff_index, low_cutoff_index;   /* mode cutoff (src/initialize.cpp:69-72) */
} ie_ic_params;
/* Host-side k-space fill of initializef (src/initialize.cpp:118-191): the
 * Park-Miller draws in the reference's order, Rayleigh amplitudes, Bunch-Davies
 * momenta, Hermitian mirror planes, DC removed.  Writes FFTEASY-layout arrays
 * (fk, fdk: N^3 doubles; fny, fdny: N x 2N) and *hubble_init, the homogeneous
 * H0 the momenta use (src/initialize.cpp:135).  Needs no GPU. */
ie_status ie_fill_modes(const ie_config *cfg, const ie_ic_params *ic, double *fk, double *fny, double *fdk,
                        double *fdny, double *hubble_init);
/* initialize() + initializef() (src/initialize.cpp:93-214) on a run: the
 * k-space fill above, the inverse transform on the GPU, phi0 and dphi0 added,
 * then t = 0, a = 1 and ad = H0 from the lattice energies.  *out receives the
 * background. */
ie_status ie_initialize(ie_run *run, const ie_ic_params *ic, ie_background *out);
/* potential(v) and its derivative in units of B^2 (src/potential.cpp:91-112). */
ie_status ie_potential_value(const ie_config *cfg, double v, double *V, double *dV);

/* --- lattice reductions ----------------------------------------------------- */
/* gradient_energy (src/evolution.cpp:130-136) */
ie_status ie_gradient_energy(ie_run *run, double *out);
/* kin_energy (src/evolution.cpp:139-145) */
ie_status ie_kin_energy(ie_run *run, double *out);
/* potential_energy (src/potential.cpp:151-171) */
ie_status ie_potential_energy(ie_run *run, double *out);
/* meansvars sums (src/output.cpp:235-240): out = {sum f, sum fd, sum f^2} */
ie_status ie_field_sums(ie_run *run, double out[3]);
/* min and max of f or deltaN (src/output.cpp:1095-1101) and the histogram with
 * nbins bins over [min, max] (src/output.cpp:1104-1123): counts has nbins entries. */
ie_status ie_histogram(ie_run *run, int which, int nbins, double *bmin, double *bmax, double *counts);

/* --- time stepping ---------------------------------------------------------- */
/* evolve_derivs(d): Friedmann half-kicks around the field (and tensor) kick
 * (src/evolution.cpp:150-257). */
ie_status ie_evolve_derivs(ie_run *run, double d);
/* evolve_fields(d): drift of f (and h), t and a (src/evolution.cpp:260-284); in
 * the post-inflation phase the drift of Phi and h. */
ie_status ie_evolve_fields(ie_run *run, double d);
/* Up to max_steps iterations of the leapfrog loop body of run_evolution_loop
 * (src/evolution.cpp:937-945) with the loop test a <= af evaluated on the device:
 * no host synchronisation between steps.  Steps are enqueued 1024 at a time and
 * the host reads the loop test between those chunks, so a call with a large
 * max_steps returns soon after af is reached.  *out receives the background after. */
ie_status ie_leapfrog_steps(ie_run *run, int max_steps, ie_background *out);
/* The same loop as ie_leapfrog_steps, with HIP events recorded on the run's
 * stream around every launch; *out receives the mean device time per launch of
 * each kernel of the step and per whole step (milliseconds).  Then nsteps more
 * steps with events around the loop alone (loop_ms): events between launches
 * hold each launch apart, so the per-launch figures run above the loop's own.  No reference
 * counterpart: the reference has wall-clock timing only (src/output.cpp:1451-1461). */
typedef struct ie_profile {
    double step_ms;            /* whole step, first to last event */
    double reduce_ms;          /* Friedmann reduction (gradient/potential sums); 0 for
                                  fused steps (128 <= N <= 256), whose kick kernel sums */
    double kick_drift_ms;      /* scalar kick + drift (fused steps: with the sums) */
    double tensor_ms;          /* tensor kick + drift (0 without SIGW) */
    int64_t steps;
    double loop_ms;            /* nsteps further steps with events around the whole loop
                                  only, per step: for a fused step without SIGW (one
                                  launch per step) the step kernel's mean launch
                                  duration, gaps between launches included */
} ie_profile;
ie_status ie_profile_steps(ie_run *run, int nsteps, ie_profile *out);

/* Shape autotuner of the step kernels (no counterpart in the reference: the
 * reference has no kernel shapes).  The unfused kick kernel of N >= 128 lattices
 * can hold 1, 2 or 4 sites per lane of a row; the tensor step can run as the row
 * kernel (1, 2 or 4 sites per lane) or as the plane-marching kernel (1, 2, or 4
 * for split storage); the
 * fused step kernel can load the next plane before or after the current one's
 * arithmetic (variants 0, 1; the same sums in the same order).
 * None of these changes a result (no cross-site sums), so the tuner times real
 * loop steps: with R = max(candidates) rounds, loop step s takes kick (or fused
 * step) candidate (s mod R) mod ncand and tensor candidate (s mod R) mod ntensor,
 * for R x steps_per_candidate steps of the device-resident loop (the candidates
 * in turn, so each samples the same stretch of the run), with events around the
 * two launches; each candidate's first launch is not counted, and the fastest
 * mean of each kind is kept.  The choice is
 * stored process-wide per (N, storage) and taken by later runs and batches of
 * that kind; ie_leapfrog_steps tunes once per kind by itself on a call of >= 256
 * steps.  The reduce and fused step kernels keep one shape, since theirs fixes the
 * summation order. */
typedef struct {
    int ncand;                 /* kick candidates timed */
    int elems[3];              /* sites per lane of each kick candidate */
    double kick_ms[3];         /* mean kick launch time per candidate */
    int ntensor;               /* tensor candidates timed (0 without SIGW) */
    int tensor_code[8];        /* 1, 2, 4: row kernel with that many sites per lane; 16 + E: march
                                  kernel with E = 1, 2, 4 (4: split storage); 24 + E: march kernel
                                  loading the next plane early (E = 1, 2) */
    double tensor_ms[8];       /* mean tensor launch time per candidate */
    int nmarch;                /* fused step kernel variants timed (128 <= N <= 256) */
    double march_ms[2];        /* mean fused step kernel time: plain, early loads */
    int kick_elems;            /* chosen kick shape */
    int tensor_code_chosen;    /* chosen tensor shape code (0 without SIGW) */
    int march_variant;         /* chosen fused step kernel variant */
    int64_t steps;             /* loop steps taken while tuning */
} ie_tuning;
ie_status ie_tune(ie_run *run, int steps_per_candidate, ie_tuning *out);
/* Pins the step kernels' shapes of one run (no counterpart in the reference):
 * kick_elems 1, 2 or 4 (0: the default), tensor_code one of ie_tuning's codes
 * (0: the default row kernel), march_variant 0 or 1.  The run keeps them and
 * ie_leapfrog_steps no longer tunes it; the process-wide choice is unchanged.  A
 * shape beyond the run's lattice (E above what a row holds, E = 4 march with FP64
 * storage) is IE_ERR_ARG.  In the post-inflation phase a pinned row tensor code
 * (1, 2, 4) keeps that phase's row tensor kernel, which otherwise serves only
 * N >= 512 (N = 64 .. 256 take its plane-marching kernel). */
ie_status ie_set_shapes(ie_run *run, int kick_elems, int tensor_code, int march_variant);
/* rk4_step_inflation / rk45_step_inflation (src/evolution.cpp:760-904), FP64
 * storage only (IE_ERR_STATE for split storage) in the inflation phase; in the
 * post-inflation phase both kinds, with the post-inflation right-hand side.  RK45 returns *accepted and the
 * proposed next step clamped to [rk45_min_dt, hmax]; a rejected step leaves the
 * state as it was.  Stage buffers (4 or 7 stages of the state) are allocated on
 * the first call. */
ie_status ie_rk4_step(ie_run *run, double h);
ie_status ie_rk45_step(ie_run *run, double h, double hmax, int *accepted, double *h_next);

/* --- spectra ---------------------------------------------------------------- */
/* Isotropic power spectrum of f or deltaN in the reference shells
 * (write_isotropic_spectrum_from_fft_r2c, src/output.cpp:63-151): numbins =
 * int(1.73205*(N/2))+1 entries of p, counts and vals = norm1 * <|F|^2>.  The
 * forward FFT is out of place; the field is not modified. */
int ie_spectrum_numbins(int N);
ie_status ie_spectrum(ie_run *run, int which, double norm1, double *p, int *counts, double *vals);
/* ie_spectrum without a run handle and without allocating: the isotropic power
 * spectrum, in the same shells and with the same kernels and numbers as
 * ie_spectrum, of an FP64 field that the caller already holds in device memory
 * (f_dev: N^3 doubles, z-fastest as idx(i,j,k) = (i N + j) N + k; the field is
 * not modified).  N must be a power of two with 8 <= N <= 128: the kernels'
 * workspace (half spectrum, binning partials, shell sums, twiddles: about
 * 17.5 MB) is a fixed device array of the library that each GPU using this
 * entry point holds until the process ends; larger N is refused with IE_ERR_ARG
 * (message in ie_last_error(NULL)).  stream: a hipStream_t, NULL for the null
 * stream; the call returns when p, counts and vals are written.  Calls on one GPU
 * share its workspace, so calls that overlap on one GPU race.
 * No counterpart in the reference: this entry point is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). */
ie_status ie_spectrum_noalloc(const double *f_dev, int N, double L, double norm1, double *p, int *counts,
                              double *vals, void *stream);
/* TT-projected tensor spectrum (write_gw_spectrum_impl, src/output.cpp:466-625):
 * hdot = 0 for spectraGW, 1 for spectraGWdot (to_phys = B a^{s-1}).
 * numbins = int(sqrt(3)*(N/2))+1; bins >= floor(0.8 numbins) are empty. */
ie_status ie_gw_spectrum(ie_run *run, int hdot, double *p, int *counts, double *vals);

/* The forcing cutoff of spectraf / spectraf_post_inflation (src/output.cpp:333-404,
 * 1698-1765): the modes of f and fd with |n| > high or |n| < low (n the integer
 * lattice momentum) are removed and both fields are transformed back to real space
 * (the reference applies it after each spectraf when forcing_cutoff is set and
 * high_cutoff_index > 0).  A low cutoff above 0 removes k = 0, the field mean. */
ie_status ie_forcing_cutoff(ie_run *run, double high, double low);
/* bispectraf (src/output.cpp:740-911): the equilateral bispectrum of f (its mean
 * removed) on numbins = int(sqrt(3)*(N/2))+1 shells: p, counts, re, im as
 * bispectra.dat's columns (norm1 = (L/B)^6 / N^9 applied).  f is not modified. */
ie_status ie_bispectrum(ie_run *run, double *p, int *counts, double *re, double *im);
/* build_deltaN_log and histogramsLOG (src/output.cpp:167-198, 1312-1403): the
 * log-mapped zeta (1/eta) log(1 - eta (f - <f>) / <fd a^{s-1}/(ad a^{s-2})>) of
 * the inflation fields into the deltaN buffer (sites with a non-positive
 * argument: 0), and its histogram over the sites with a positive argument:
 * freq[nbins] = counts / points, *bmin and *df as histogramtimesLOG.dat writes
 * them.  *valid = 0 (map and freq 0, *bmin 0, *df 1) when the mean velocity
 * factor is not finite or below 1e-30.  ie_spectrum(IE_FIELD_DELTAN) then gives
 * spectraLOG. */
ie_status ie_deltaN_log(ie_run *run, double eta_log, int nbins, double *bmin, double *df, double *freq, int *valid);

/* --- curvature and tensor power spectra ------------------------------------ */
/* The analysis notebook's formulas (notebooks/plot.ipynb cells 9, 23) on binned
 * spectra, each row paired with the modes.dat row of the same index (k[b], from
 * ie_modes): P_zeta = P_phi k^3/(2 pi^2) (phidot/H)^-2 (phidot and H as written
 * to velocity.dat and sf.dat); P_zeta = P_dN k^3/(2 pi^2); Delta_h^2 =
 * k^3/(2 pi^2) P_h; Omega_GW = k^3/(2 pi^2) (P_hdot + (k/a)^2 P_h) / (12 H^2).
 * Host functions; out has nbins entries. */
ie_status ie_pzeta_linear(int nbins, const double *k, const double *p_phi, double phidot, double hubble,
                          double *out);
ie_status ie_pzeta_deltaN(int nbins, const double *k, const double *p_dn, double *out);
ie_status ie_delta_h2(int nbins, const double *k, const double *p_h, double *out);
ie_status ie_omega_gw(int nbins, const double *k, const double *p_h, const double *p_hdot, double a, double hubble,
                      double *out);

/* --- deltaN ----------------------------------------------------------------- */
/* initializeN (src/initialize.cpp:219-232): fd -> dphi/dN, deltaN = 0. */
ie_status ie_deltaN_init(ie_run *run);
/* evolve_fieldsN (src/evolution.cpp:1015-1037) with the current phiref. */
ie_status ie_evolve_fieldsN(ie_run *run, double d);
/* get_phiref (src/evolution.cpp:1271-1288): ordered search, first wins. */
ie_status ie_get_phiref(ie_run *run, double *out);
/* nsteps iterations of evolve_derivsN + evolve_fieldsN (src/evolution.cpp:1322-1327)
 * kept in registers on the device: one read and one write of the state. */
ie_status ie_deltaN_steps(ie_run *run, double dN, int64_t nsteps);
/* nsteps classical RK4 steps of the deltaN system (rk4_step_deltaN,
 * src/evolution.cpp:1176-1189), also kept in registers per site. */
ie_status ie_deltaN_rk4_steps(ie_run *run, double h, int64_t nsteps);
/* One Dormand-Prince trial step of the deltaN system (rk45_step_deltaN,
 * src/evolution.cpp:1191-1267): the weighted RMS error over (phi, pi, N) of every
 * site and a; *accepted and the proposed next step as for ie_rk45_step.  The
 * stages stay in registers; three N^3 FP64 buffers for the trial state are
 * allocated on the first call. */
ie_status ie_deltaN_rk45_step(ie_run *run, double h, double hmax, int *accepted, double *h_next);
/* The zeta map of saveN (src/output.cpp:1527-1542): mean of deltaN over sites
 * with deltaN <= cutoff, accumulated in float in lattice order and divided by N^3
 * in float as the reference does, then subtracted from every site.  *nmean
 * receives that mean. */
ie_status ie_deltaN_zeta(ie_run *run, double cutoff, float *nmean);
/* Sites with deltaN > cutoff - nmean (not finished) set to 0, after the
 * histogram (src/output.cpp:1549-1553); uses the mean of the last ie_deltaN_zeta. */
ie_status ie_deltaN_zero_unfinished(ie_run *run, double cutoff);
/* subtract the FP64 mean of deltaN (build_deltaN_linear, src/output.cpp:156-164) */
ie_status ie_deltaN_center(ie_run *run);

/* --- post-inflation --------------------------------------------------------- */
/* The fluid-era phase after inflation and the deltaN map (src/parameters.h:31,
 * 148-155; src/evolution.cpp:1378-1651): the Bardeen potential Phi and its
 * derivative in FP64 (for both storage kinds: split runs allocate two more N^3
 * double buffers), the tensors in FP32, the background (a, a') advanced by the
 * host from its closed equation.  Needs calculate_sigw = 1, as the reference
 * (src/parameters.h:33-35). */
typedef struct ie_post_params {
    double horizon_factor, omega, dt, af;  /* src/runtime_parameters.cpp:48-52 (af <= 0: 2N) */
    int integrator;                        /* post_inflation_integrator: IE_LEAPFROG, IE_RK4, IE_RK45 */
    int from_deltaN;                       /* 1: Phi from the deltaN map (perform_deltaN = 1) */
} ie_post_params;
/* initialize_post_inflation (src/initialize.cpp:246-291) without its save:
 * tensors zeroed; Phi = 3(1+w)/(5+3w) (deltaN - <deltaN>) with the mean summed
 * in lattice order, or without deltaN Phi = 3(1+w)/(5+3w) (f - <f>) / (<fd>
 * a^{s-1} / (ad a^{s-2})); fd kept as the previous phase left it; t = 0, a = 1,
 * ad = horizon_factor N 2 pi / L.  from_deltaN needs the deltaN phase
 * (ie_deltaN_init).  Afterwards ie_evolve_fields drifts the post-inflation state,
 * ie_rk4_step / ie_rk45_step use its right-hand side (src/evolution.cpp:520-592,
 * 1606-1640), the output entry points (ie_field_sums, ie_spectrum, ie_histogram,
 * ie_gw_spectrum with spectraGWdot_post_inflation's float factor,
 * src/output.cpp:662-669) read it, and the inflation and deltaN steps refuse. */
ie_status ie_post_init(ie_run *run, const ie_post_params *params, ie_background *out);
/* evolve_derivs_post_inflation (src/evolution.cpp:1437-1564): Phi kick, then the
 * tensor kick whose source uses the kicked Phi' (no drift). */
ie_status ie_post_evolve_derivs(ie_run *run, double d);
/* Up to max_steps iterations of the leapfrog loop body of run_post_inflation_loop
 * (src/evolution.cpp:1598-1604) while a <= af: two launches per step, no host
 * synchronisation.  out->done is set once a > af. */
ie_status ie_post_leapfrog_steps(ie_run *run, int max_steps, ie_background *out);

/* --- batches ---------------------------------------------------------------- */
/* Many runs advanced together: each step of the leapfrog loop
 * (src/evolution.cpp:937-945) is one launch per kernel for all runs of the batch
 * (any mix of N, potentials, SIGW on or off; one storage kind per batch), with
 * each run's own background, loop test and partial sums, captured in a HIP graph
 * and replayed.  No reference counterpart: the reference runs one lattice per
 * process.  A run in a batch keeps working with every other entry point; the
 * batch only holds pointers to the runs' state. */
typedef struct ie_batch ie_batch;
ie_status ie_batch_create(ie_run *const *runs, int nruns, ie_batch **out);
/* nsteps loop steps of every run (runs with a > af are skipped on the device);
 * graph_steps > 0 captures that many steps per HIP graph launch, 0 launches
 * directly. */
ie_status ie_batch_steps(ie_batch *batch, int nsteps, int graph_steps);
/* ie_batch_steps in two halves, so that several batches (each on its own stream)
 * run at once: ie_batch_launch enqueues the steps and returns; ie_batch_wait
 * waits for them and refreshes every run's background and error flag, then
 * returns IE_ERR_TABLE_RANGE if any run raised one (the flag stays set in that
 * run's background, so the caller can tell which). */
ie_status ie_batch_launch(ie_batch *batch, int nsteps, int graph_steps);
ie_status ie_batch_wait(ie_batch *batch);
ie_status ie_batch_destroy(ie_batch *batch);
const char *ie_batch_last_error(const ie_batch *batch);
/* Path of the 64^3 lattices without tensors in batches created afterwards: one
 * resident workgroup per lattice running every step of a launch (throughput:
 * it fills the GPU with many lattices), or the per-step kernels, which spread
 * each lattice over the whole GPU (latency: a lone lattice steps several times
 * faster).  A batch takes the resident path for the 64^3 lattices of a
 * potential kind when it holds at least min_hilltop (analytic potential) or
 * min_tabulated (tabulated potential) of them; 0 restores the measured
 * defaults.  Process-wide; set it before creating batches, not beside another
 * thread's ie_batch_create.  No reference counterpart. */
ie_status ie_batch_set_resident_min(int min_hilltop, int min_tabulated);

/* --- serving ---------------------------------------------------------------- */
/* A request loop over whole runs (initial conditions, inflation loop to af, then
 * the GW spectra and Omega_GW of SIGW runs and the deltaN map and P_zeta).  Requests wait in FIFO order until their device
 * footprint (ie_footprint_bytes) fits the budget; admitted runs stay resident
 * and are advanced together by batches (continuous batching: runs join and
 * leave between quanta); finished runs are completed on worker threads while
 * the others keep stepping.  No reference counterpart (the reference runs one
 * lattice per process). */
typedef struct ie_server ie_server;
typedef struct ie_request {
    ie_config cfg;
    ie_ic_params ic;
    int deltaN;                /* 1: deltaN map and P_zeta after inflation */
    int fp64_check;            /* 1: also run an FP64 twin and report the P_zeta error */
} ie_request;
typedef struct ie_result {
    int64_t id;
    int32_t status;            /* ie_status of the run */
    int32_t N;
    double submit_s, admit_s, done_s;   /* server clock, seconds */
    int64_t steps;             /* inflation steps taken */
    double a_end;
    int32_t nbins;             /* P_zeta shells written to pzeta (<= 512) */
    double pzeta[512];         /* P_zeta(k) = k^3/(2 pi^2) spectraN on the modes.dat k (ie_pzeta_deltaN) */
    double fp64_max_rel_err;   /* max over shells of |P_split - P_fp64| / P_fp64, -1 if not checked */
    int32_t gw_nbins;          /* SIGW runs: Omega_GW shells written to omega_gw, floor(0.8 numbins); else 0 */
    double omega_gw[512];      /* Omega_GW(k) at the end of inflation (ie_omega_gw of spectraGW, spectraGWdot) */
    double gw_fp64_max_rel_err;/* max over kept shells of the FP64 twin's Omega_GW difference, -1 if not checked */
} ie_result;
/* The end of one run as the reference's main() does after the inflation loop
 * (src/main.cpp:129-146): the final leapfrog synchronisation
 * (src/evolution.cpp:993-995); for a SIGW run spectraGW and spectraGWdot
 * (src/output.cpp:631-648) and Omega_GW on the modes.dat k with that output's a
 * and H (notebooks/plot.ipynb cell 23); with deltaN = 1 the deltaN phase
 * (src/evolution.cpp:1291-1375) and saveN's zeta spectrum (src/output.cpp:
 * 1522-1560) as P_zeta (notebooks/plot.ipynb cell 9).  Fills steps, a_end,
 * nbins, pzeta, gw_nbins and omega_gw of *out and leaves its other fields.  The
 * run must be in its inflation phase with its loop ended; the serving loop
 * finishes every run with it. */
ie_status ie_run_outputs(ie_run *run, int deltaN, ie_result *out);
ie_status ie_server_create(size_t hbm_budget_bytes, int quantum_steps, int finish_threads, ie_server **out);
/* Copy a request in (table pointers are copied too); *id identifies it. */
ie_status ie_server_submit(ie_server *s, const ie_request *req, int64_t *id);
/* One scheduling round: admit what fits, advance the resident runs (one batch
 * per storage kind and lattice size; N^3 runs by quantum_steps * max(1, 256/N)
 * steps), hand finished runs to the finishing threads, collect results.
 * *active and *pending receive the counts after the round. */
ie_status ie_server_pump(ie_server *s, int *active, int *pending);
/* Pop one completed result (returns IE_ERR_STATE when none is ready). */
ie_status ie_server_pop_result(ie_server *s, ie_result *out);
double ie_server_clock(const ie_server *s);
/* Loop steps per HIP graph launch of the batches (default: quantum_steps; 0:
 * every step launched directly, e.g. for tools that cannot follow graph
 * launches). */
ie_status ie_server_set_graph_steps(ie_server *s, int graph_steps);
ie_status ie_server_destroy(ie_server *s);

/* --- modes ------------------------------------------------------------------ */
/* get_modes (src/output.cpp:674-738): B * mean lattice momentum per truncated shell. */
ie_status ie_modes(int N, double L, double rescale_B, double *out);

/* --- library ---------------------------------------------------------------- */
const char *ie_version(void);
int ie_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
