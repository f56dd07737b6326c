// ie_fft.h -- 3-D real FFT, shell binning and TT projection on gfx950 (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace ie {

struct FftPlan;

size_t fft_workspace_bytes(int N);
FftPlan *fft_create(int N);
void fft_destroy(FftPlan *p);
// The twiddle table of a plan of size N (N complex: cos, sin of 2 pi k / N).
void fft_twiddles(int N, double2 *tw);
// Scalar spectrum of an FP64 field over caller memory, with no allocation:
// `work` holds fft_view_bytes(N) bytes of device memory, `tw` the N twiddles in
// device memory.  The same kernels and shell sums as fft_forward_double followed
// by fft_bin_scalar; sums_host receives fft_numbins(N) values.
size_t fft_view_bytes(int N);
hipError_t fft_spectrum_in(int N, void *work, const double2 *tw, const double *in, double *sums_host,
                           hipStream_t s);

// Number of reference shells: int(1.73205*(N/2))+1 (src/output.cpp:69).
int fft_numbins(int N);
// Number of GW shells: int(sqrt(3)*(N/2))+1 (src/output.cpp:473).
int fft_gw_numbins(int N);

// Forward r2c transform of a real field (double, or float plus a homogeneous
// offset added to the k = 0 mode) into the plan's half-spectrum buffer `slot`
// (0..5), then (when bin = 1) the isotropic shell sums of |F|^2 with the
// reference weights into sums[numbins] (host).
hipError_t fft_forward_double(FftPlan *p, const double *in, int slot, hipStream_t s);
hipError_t fft_forward_float(FftPlan *p, const float *in, double dc_add, int slot, hipStream_t s);
hipError_t fft_bin_scalar(FftPlan *p, int slot, double *sums_host, hipStream_t s);
// Forward transform and shell sums in one: the last pass bins its modes instead
// of storing them (for 8 <= N <= 256; larger N as forward + bin).  Slot 0 is
// used as workspace; sums_host receives fft_numbins(N) values.
hipError_t fft_spectrum_double(FftPlan *p, const double *in, double *sums_host, hipStream_t s);
hipError_t fft_spectrum_float(FftPlan *p, const float *in, double dc_add, double *sums_host, hipStream_t s);
// as fft_bin_scalar with dc_add added to the k = 0 amplitude first (split storage:
// N^3 phibar is the transform of the homogeneous mode)
hipError_t fft_bin_scalar_dc(FftPlan *p, int slot, double dc_add, double *sums_host, hipStream_t s);
// TT-projected shell sums over the six tensor spectra in slots 0..5, each
// amplitude scaled by to_phys; dx sets k_eff = (2/dx) sin(pi n/N).
hipError_t fft_bin_tt(FftPlan *p, double dx, double to_phys, double *sums_host, hipStream_t s);
// Inverse (FFTEASY convention: e^{-i}, 1/N per dimension) of a half spectrum
// given in FFTEASY layout (interleaved kz < N/2 in an N^3 array + N x 2N Nyquist
// plane), written as real field out[] = result + add.
hipError_t fft_inverse_ffteasy(FftPlan *p, const double *fk_host, const double *fny_host, double add,
                               double *out_dev, hipStream_t s);

// The half spectrum of a slot (H[(i N + j) K + kz], K = N/2 + 1), for kernels of
// other units that read modes.
const double2 *fft_slot(const FftPlan *p, int slot);
// Zero the modes of a slot with |n| > hi or |n| < lo (the forcing cutoff).
hipError_t fft_kill_modes(FftPlan *p, int slot, double hi, double lo, hipStream_t s);
// Zero the k = 0 mode of a slot (the transform of a mean-subtracted field).
hipError_t fft_zero_dc(FftPlan *p, int slot, hipStream_t s);
// Inverse of a slot's half spectrum (1/N^3, the round trip is the identity) into
// out[] = result + add; the slot's contents are destroyed.
hipError_t fft_inverse_slot(FftPlan *p, int slot, double add, double *out, hipStream_t s);

// Reference mode counts per shell (closed form, host).
void fft_counts(int N, int *counts);
void fft_gw_counts(int N, int *counts);
// The same, computed once per N and kind (gw 0: fft_counts, 1: fft_gw_counts)
// and kept for the process.
const int *fft_counts_cached(int N, int gw);

}  // namespace ie
