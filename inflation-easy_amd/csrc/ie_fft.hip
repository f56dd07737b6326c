// ie_fft.hip -- 3-D real FFT, isotropic shell binning and TT projection on gfx950.
//
// Replaces the serial FFTEASY transform (src/ffteasy.hpp:21-190) and the binning
// back-ends of src/output.cpp:63-151 and :466-625.  Differences of design:
//   - out of place: the field is read once and never modified, so the inverse
//     "restore" transforms of the reference (src/output.cpp:404,616,1167) are gone;
//   - half-spectrum layout H[(i N + j) K + kz], K = N/2 + 1, complex FP64, with
//     the kz = N/2 plane in place rather than in a separate Nyquist array;
//   - each pass loads a tile of L lines into LDS (consecutive rows for z, L
//     consecutive kz columns for y and x, so strided passes stay coalesced) and
//     runs log2(N) radix-2 butterfly stages there; twiddles are exact
//     (host libm), not a rotation recurrence;
//   - all spectra are computed in FP64 (the reference uses FP32 for h_ij, which
//     limits GW parity to the reference's own float rounding);
//   - shell sums are accumulated per wave in LDS, per workgroup in a partials
//     array and summed in fixed order; mode counts per shell are closed-form.
// The forward kernel is e^{+2 pi i k n/N} like FFTEASY's (|F|^2 does not depend
// on the sign); the inverse used for initial conditions is FFTEASY's inverse.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#include "ie_fft.h"

namespace ie {

namespace {

constexpr int kThreads = 256;
constexpr int kLdsComplex = 2048;   // complex doubles of LDS per FFT workgroup (32 KB)

__device__ __forceinline__ int bitrev(int t, int logn) { return (int)(__brev((unsigned)t) >> (32 - logn)); }

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// In-place radix-2 DIT transform of nl lines of n points in LDS (line l at
// x[l n], input in bit-reversed order).  tw[k] = e^{+2 pi i k / n}.
__device__ void lds_fft(double2 *x, int n, int logn, int nl, const double2 *__restrict__ tw, int sign)
{
    const int halfn = n >> 1;
    for (int span = 2, ls = 1; span <= n; span <<= 1, ++ls) {
        const int half = span >> 1, twstride = n >> ls;
        __syncthreads();
        for (int p = threadIdx.x; p < nl * halfn; p += blockDim.x) {
            const int l = p >> (logn - 1), q = p & (halfn - 1);
            const int grp = q >> (ls - 1), pos = q & (half - 1);
            const int a = l * n + grp * span + pos, b = a + half;
            double2 w = tw[pos * twstride];
            if (sign < 0) w.y = -w.y;
            const double2 u = x[a], t = cmul(w, x[b]);
            x[a] = make_double2(u.x + t.x, u.y + t.y);
            x[b] = make_double2(u.x - t.x, u.y - t.y);
        }
    }
    __syncthreads();
}

// z pass, real to half-complex: L rows per workgroup.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fft_z_r2c(const T *__restrict__ in, double2 *__restrict__ out, int N,
                                                         int logN, int L, const double2 *__restrict__ tw)
{
    extern __shared__ double2 lds[];
    const int K = N / 2 + 1;
    const size_t row0 = (size_t)blockIdx.x * L;
    for (int e = threadIdx.x; e < L * N; e += blockDim.x) {
        const int l = e >> logN, t = e & (N - 1);
        lds[l * N + bitrev(t, logN)] = make_double2((double)in[(row0 + l) * N + t], 0.0);
    }
    lds_fft(lds, N, logN, L, tw, +1);
    for (int e = threadIdx.x; e < L * K; e += blockDim.x) {
        const int l = e / K, kz = e - l * K;
        out[(row0 + l) * K + kz] = lds[l * N + kz];
    }
}

// y (dim = 1) or x (dim = 0) pass over the half spectrum, in place: a workgroup
// holds L consecutive kz columns of one outer index.
__global__ __launch_bounds__(kThreads) void k_fft_cols(double2 *__restrict__ H, int N, int logN, int dim, int L,
                                                        const double2 *__restrict__ tw, int sign)
{
    extern __shared__ double2 lds[];
    const int K = N / 2 + 1;
    const int o = blockIdx.y, kz0 = blockIdx.x * L;
    const size_t stride = dim == 1 ? (size_t)K : (size_t)N * K;
    const size_t base = dim == 1 ? (size_t)o * N * K : (size_t)o * K;
    for (int e = threadIdx.x; e < L * N; e += blockDim.x) {
        const int t = e / L, l = e - t * L, kz = kz0 + l;
        lds[l * N + bitrev(t, logN)] = kz < K ? H[base + t * stride + kz] : make_double2(0.0, 0.0);
    }
    lds_fft(lds, N, logN, L, tw, sign);
    for (int e = threadIdx.x; e < L * N; e += blockDim.x) {
        const int t = e / L, l = e - t * L, kz = kz0 + l;
        if (kz < K) H[base + t * stride + kz] = lds[l * N + t];
    }
}

// z pass, half-complex to real with Hermitian extension; out = scale Re + add.
__global__ __launch_bounds__(kThreads) void k_fft_z_c2r(const double2 *__restrict__ H, double *__restrict__ out,
                                                         int N, int logN, int L, const double2 *__restrict__ tw,
                                                         double scale, double add)
{
    extern __shared__ double2 lds[];
    const int K = N / 2 + 1;
    const size_t row0 = (size_t)blockIdx.x * L;
    for (int e = threadIdx.x; e < L * N; e += blockDim.x) {
        const int l = e >> logN, t = e & (N - 1);
        double2 v;
        if (t < K) v = H[(row0 + l) * K + t];
        else {
            v = H[(row0 + l) * K + (N - t)];
            v.y = -v.y;
        }
        lds[l * N + bitrev(t, logN)] = v;
    }
    lds_fft(lds, N, logN, L, tw, -1);
    for (int e = threadIdx.x; e < L * N; e += blockDim.x) {
        const int l = e >> logN, t = e & (N - 1);
        out[(row0 + l) * N + t] = lds[l * N + t].x * scale + add;
    }
}

// The shell of a mode, lround(sqrt(r2)) for the integer r2 = |n|^2, without the
// FP64 square root: a float estimate, then the exact integer test
// (2b - 1)^2 <= 4 r2 < (2b + 1)^2, the lower bound for b > 0 (r2 is never a
// half-integer square, so there is no tie).  The same integer as the reference's shell index (r2 < 2^22).
__device__ __forceinline__ int shell_of(int r2)
{
    int b = (int)(__fsqrt_rn((float)r2) + 0.5f);
    const int q = 4 * r2;
    b += q >= 4 * b * b + 4 * b + 1 ? 1 : 0;
    b -= b > 0 && q < 4 * b * b - 4 * b + 1 ? 1 : 0;
    return b;
}

__device__ __forceinline__ void mode_coords(size_t m, int N, int K, int *px, int *py, int *kz)
{
    const size_t row = m / K;
    *kz = (int)(m - row * K);
    const int j = (int)(row % N), i = (int)(row / N);
    *px = i <= N / 2 ? i : i - N;
    *py = j <= N / 2 ? j : j - N;
}

// Shell sums of w |F|^2, w = 2 inside (0 < kz < N/2), 1 on kz = 0, N/2
// (src/output.cpp:85-141); bins by lround(|n|).
__global__ __launch_bounds__(kThreads) void k_bin_scalar(const double2 *__restrict__ H, int N, int nb,
                                                          double dc_add, double *__restrict__ partials)
{
    extern __shared__ double wbins[];   // [kThreads/64][nb]
    const int K = N / 2 + 1, wave = threadIdx.x >> 6;
    for (int e = threadIdx.x; e < (kThreads / 64) * nb; e += blockDim.x) wbins[e] = 0.0;
    __syncthreads();
    const size_t modes = (size_t)N * N * K;
    for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < modes; m += (size_t)gridDim.x * blockDim.x) {
        int px, py, kz;
        mode_coords(m, N, K, &px, &py, &kz);
        const int bin = shell_of(px * px + py * py + kz * kz);
        if (bin >= nb) continue;
        double2 F = H[m];
        if (m == 0) F.x += dc_add;
        const double p2 = F.x * F.x + F.y * F.y;
        const double w = (kz == 0 || kz == N / 2) ? 1.0 : 2.0;
        atomicAdd(&wbins[wave * nb + bin], w * p2);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        double s = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) s += wbins[w * nb + b];
        partials[(size_t)blockIdx.x * nb + b] = s;
    }
}

// TT projection Lambda_{ab,lm} h_lm = (P h P)_ab - 1/2 P_ab tr(P h) with
// P = 1 - k^ k^, k = (2/dx) sin(pi n/N) (src/output.cpp:493-543); shell sums of
// w sum_ab |X_ab|^2 for shells below floor(0.8 numbins).
struct SixSpectra {
    const double2 *h[6];
};

__global__ __launch_bounds__(kThreads) void k_bin_tt(SixSpectra S, int N, int nb, int imax, double kpre,
                                                      double to_phys, double *__restrict__ partials)
{
    extern __shared__ double wbins[];   // [waves][nb], then kpre sin(pi p / N) for p = -N/2 .. N/2
    const int K = N / 2 + 1, wave = threadIdx.x >> 6;
    double *ksin = wbins + (kThreads / 64) * nb + N / 2;
    const double pin = 3.14159265358979323846 / N;
    for (int e = threadIdx.x; e < (kThreads / 64) * nb; e += blockDim.x) wbins[e] = 0.0;
    for (int p = (int)threadIdx.x - N / 2; p <= N / 2; p += blockDim.x) ksin[p] = kpre * sin(pin * p);
    __syncthreads();
    const size_t modes = (size_t)N * N * K;
    for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < modes; m += (size_t)gridDim.x * blockDim.x) {
        int px, py, kz;
        mode_coords(m, N, K, &px, &py, &kz);
        if (px == 0 && py == 0 && kz == 0) continue;
        const int bin = shell_of(px * px + py * py + kz * kz);
        if (bin >= imax) continue;
        double k3[3] = {ksin[px], ksin[py], ksin[kz]};
        const double inv = 1.0 / sqrt(k3[0] * k3[0] + k3[1] * k3[1] + k3[2] * k3[2]);
        double P[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) P[a][b] = (a == b ? 1.0 : 0.0) - (k3[a] * inv) * (k3[b] * inv);
        double Cr[3][3], Ci[3][3];
        const int comp[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const double2 v = S.h[comp[a][b]][m];
                Cr[a][b] = v.x * to_phys;
                Ci[a][b] = v.y * to_phys;
            }
        double PCr[3][3], PCi[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                double sr = 0.0, si = 0.0;
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    sr += P[a][l] * Cr[l][b];
                    si += P[a][l] * Ci[l][b];
                }
                PCr[a][b] = sr;
                PCi[a][b] = si;
            }
        const double Tr = PCr[0][0] + PCr[1][1] + PCr[2][2], Ti = PCi[0][0] + PCi[1][1] + PCi[2][2];
        double fp2 = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                double ar = 0.0, ai = 0.0;
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    ar += PCr[a][l] * P[l][b];
                    ai += PCi[a][l] * P[l][b];
                }
                const double xr = ar - 0.5 * P[a][b] * Tr, xi = ai - 0.5 * P[a][b] * Ti;
                fp2 += xr * xr + xi * xi;
            }
        const double w = (kz == 0 || kz == N / 2) ? 1.0 : 2.0;
        atomicAdd(&wbins[wave * nb + bin], w * fp2);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        double s = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) s += wbins[w * nb + b];
        partials[(size_t)blockIdx.x * nb + b] = s;
    }
}

// The k-space filter of spectraf's forcing cutoff (src/output.cpp:333-404):
// modes with |n| > hi or |n| < lo set to zero.  The reference zeroes each mode of
// the kz = 0 and kz = N/2 planes together with its Hermitian mirror, which is the
// same set: |n| does not change under n -> -n.
__global__ __launch_bounds__(kThreads) void k_kill_modes(double2 *__restrict__ H, int N, double hi, double lo)
{
    const int K = N / 2 + 1;
    const size_t modes = (size_t)N * N * K;
    for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < modes; m += (size_t)gridDim.x * blockDim.x) {
        int px, py, kz;
        mode_coords(m, N, K, &px, &py, &kz);
        const double pdisc = sqrt((double)px * px + (double)py * py + (double)kz * kz);
        if (pdisc > hi || pdisc < lo) H[m] = make_double2(0.0, 0.0);
    }
}


// ------------------------------------------------ register-resident line FFT
//
// A line of N <= 256 points split as N = N1 x N2 (N2 = N1 or 2 N1).  Thread n1
// of the line holds x[n1 + N1 n2], n2 < N2, in registers:
//   1. the N2-point transform over n2 in registers;
//   2. the twiddles w_N^(n1 k2);
//   3. one exchange through LDS, then the N1-point transforms over n1 in
//      registers (N2 / N1 of them per thread): X[k2 + N2 k1].
// Two register passes and one barrier per line, against log2 N LDS stages with a
// barrier each.  Sub-transform twiddles sit at compile-time offsets of the
// workgroup's LDS copy of the exact table (every lane reads the same entry).
template <int N>
struct LineShape;
template <> struct LineShape<8> { static constexpr int N1 = 2; };
template <> struct LineShape<16> { static constexpr int N1 = 4; };
template <> struct LineShape<32> { static constexpr int N1 = 4; };
template <> struct LineShape<64> { static constexpr int N1 = 8; };
template <> struct LineShape<128> { static constexpr int N1 = 8; };
template <> struct LineShape<256> { static constexpr int N1 = 16; };

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
constexpr int bitrev_c(int x, int bits)
{
    int r = 0;
    for (int b = 0; b < bits; ++b) r |= ((x >> b) & 1) << (bits - 1 - b);
    return r;
}

// In-register radix-2 DIT transform of R points, natural order in and out,
// kernel e^{SIGN 2 pi i k n / R}; tw holds e^{+2 pi i m / N}, tstride = N / R.
template <int R, int SIGN>
__device__ __forceinline__ void reg_fft(double2 (&a)[R], const double2 *tw, int tstride)
{
    constexpr int LR = ilog2c(R);
    double2 b[R];
#pragma unroll
    for (int i = 0; i < R; ++i) b[bitrev_c(i, LR)] = a[i];
#pragma unroll
    for (int ls = 1; ls <= LR; ++ls) {
        const int span = 1 << ls, half = span >> 1;
#pragma unroll
        for (int pos = 0; pos < half; ++pos) {
            double2 w = make_double2(1.0, 0.0);
            if (pos > 0) {
                w = tw[pos * (R / span) * tstride];
                if (SIGN < 0) w.y = -w.y;
            }
#pragma unroll
            for (int g = 0; g < R; g += span) {
                const double2 u = b[g + pos], v = b[g + pos + half];
                const double2 t = pos == 0 ? v : cmul(w, v);
                b[g + pos] = make_double2(u.x + t.x, u.y + t.y);
                b[g + pos + half] = make_double2(u.x - t.x, u.y - t.y);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i) a[i] = b[i];
}

// Exchange area of one line: N2 rows of N1 + 1 points (the pad keeps the reads
// of step 3, a row per thread, off a common bank).
template <int N>
constexpr int kExRow = LineShape<N>::N1 + 1;

// Four-step transform of one line held by its N1 threads: in a[n2] = x[n1 + N1 n2]
// (thread n1), out a[u N1 + k1] = X[k2 + N2 k1] with k2 = n1 + N1 u.  Every
// thread of the workgroup calls it (one barrier inside); ex is the line's
// exchange area, tw the workgroup's twiddle table (written before a barrier).
template <int N, int SIGN>
__device__ __forceinline__ void line_fft(double2 (&a)[N / LineShape<N>::N1], int n1, double2 *ex,
                                         const double2 *tw)
{
    constexpr int N1 = LineShape<N>::N1, N2 = N / N1, Q = N2 / N1;
    reg_fft<N2, SIGN>(a, tw, N / N2);
#pragma unroll
    for (int k2 = 1; k2 < N2; ++k2) {
        double2 w = tw[n1 * k2];
        if (SIGN < 0) w.y = -w.y;
        a[k2] = cmul(w, a[k2]);
    }
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) ex[k2 * kExRow<N> + n1] = a[k2];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < Q; ++u) {
        double2 c[N1];
#pragma unroll
        for (int m = 0; m < N1; ++m) c[m] = ex[(n1 + N1 * u) * kExRow<N> + m];
        reg_fft<N1, SIGN>(c, tw, N / N1);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) a[u * N1 + k1] = c[k1];
    }
}

constexpr int kRegThreads = 256;   // z passes: 256 / N1 rows per workgroup
constexpr int kColLines = 16;      // column passes: 16 consecutive kz per workgroup
constexpr unsigned kColXcd = 8;     // XCDs of the MI355X (workgroup b runs on XCD b mod 8)

// Doubles of a column workgroup's exchange area, which holds the binning pass's
// per-wave shell sums after the transform.
template <int N>
constexpr size_t col_bin_capacity() { return 2u * kColLines * (N / LineShape<N>::N1) * kExRow<N>; }

// z pass, real to half-complex.
template <typename T, int N>
__global__ __launch_bounds__(kRegThreads) void k_z_r2c_reg(const T *__restrict__ in, double2 *__restrict__ out,
                                                            const double2 *__restrict__ twg)
{
    constexpr int N1 = LineShape<N>::N1, N2 = N / N1, L = kRegThreads / N1, K = N / 2 + 1;
    __shared__ double2 tw[N];
    __shared__ double2 ex[L][N2 * kExRow<N>];
    for (int e = threadIdx.x; e < N; e += kRegThreads) tw[e] = twg[e];
    const int l = threadIdx.x / N1, n1 = threadIdx.x % N1;
    const size_t row = (size_t)blockIdx.x * L + l;
    double2 a[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) a[n2] = make_double2((double)in[row * N + n1 + N1 * n2], 0.0);
    __syncthreads();
    line_fft<N, +1>(a, n1, ex[l], tw);
#pragma unroll
    for (int u = 0; u < N2 / N1; ++u)
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) {
            const int k = n1 + N1 * u + N2 * k1;
            if (k < K) out[row * K + k] = a[u * N1 + k1];
        }
}

// z pass, half-complex to real with Hermitian extension; out = scale Re + add.
template <int N>
__global__ __launch_bounds__(kRegThreads) void k_z_c2r_reg(const double2 *__restrict__ H, double *__restrict__ out,
                                                            const double2 *__restrict__ twg, double scale,
                                                            double add)
{
    constexpr int N1 = LineShape<N>::N1, N2 = N / N1, L = kRegThreads / N1, K = N / 2 + 1;
    __shared__ double2 tw[N];
    __shared__ double2 ex[L][N2 * kExRow<N>];
    for (int e = threadIdx.x; e < N; e += kRegThreads) tw[e] = twg[e];
    const int l = threadIdx.x / N1, n1 = threadIdx.x % N1;
    const size_t row = (size_t)blockIdx.x * L + l;
    double2 a[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
        const int n = n1 + N1 * n2;
        double2 v = H[row * K + (n < K ? n : N - n)];
        if (n >= K) v.y = -v.y;
        a[n2] = v;
    }
    __syncthreads();
    line_fft<N, -1>(a, n1, ex[l], tw);
#pragma unroll
    for (int u = 0; u < N2 / N1; ++u)
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) out[row * N + n1 + N1 * u + N2 * k1] = a[u * N1 + k1].x * scale + add;
}

// y (DIM = 1) or x (DIM = 0) pass over the half spectrum, in place: a workgroup
// holds 16 consecutive kz columns of one outer index, thread (l, n1) with l the
// fastest index so that a load instruction reads 16 consecutive modes per row.
// BIN (x pass of a spectrum): instead of storing, the shell sums of w |F|^2 with
// the reference weights (as k_bin_scalar), per wave in LDS and per workgroup
// into partials, dc_add added to the k = 0 amplitude first.  The per-wave sums
// reuse the exchange area once the transform has read it, so the binning pass
// needs no LDS beyond the plain pass (two workgroups per CU at N = 256).
template <int N, int SIGN, int DIM, bool BIN>
__global__ __launch_bounds__(kColLines * LineShape<N>::N1) void k_cols_reg(double2 *__restrict__ H,
                                                                            const double2 *__restrict__ twg,
                                                                            int nb, double dc_add,
                                                                            double *__restrict__ partials)
{
    constexpr int N1 = LineShape<N>::N1, N2 = N / N1, K = N / 2 + 1, NT = kColLines * N1;
    constexpr int NW = NT >= 64 ? NT / 64 : 1;
    __shared__ double2 tw[N];
    __shared__ double2 ex[kColLines][N2 * kExRow<N>];
    for (int e = threadIdx.x; e < N; e += NT) tw[e] = twg[e];
    // Workgroup b runs on XCD b mod 8: XCD x takes the x-th eighth of the
    // (outer index, column group) pairs in order, so the column groups of one
    // outer index share an L2 and the cache line two neighbouring groups
    // straddle (rows of K modes are not line-aligned) is fetched once.
    const unsigned nblk = gridDim.x * gridDim.y, blk = blockIdx.x + gridDim.x * blockIdx.y;
    const unsigned lid = nblk % kColXcd == 0 ? (blk % kColXcd) * (nblk / kColXcd) + blk / kColXcd : blk;
    const int l = threadIdx.x % kColLines, n1 = threadIdx.x / kColLines;
    const int o = (int)(lid / gridDim.x), kz = (int)(lid % gridDim.x) * kColLines + l;
    const size_t stride = DIM == 1 ? (size_t)K : (size_t)N * K;
    const size_t base = (DIM == 1 ? (size_t)o * N * K : (size_t)o * K) + (size_t)(kz < K ? kz : 0);
    const bool live = kz < K;
    double2 a[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2)
        a[n2] = live ? H[base + (size_t)(n1 + N1 * n2) * stride] : make_double2(0.0, 0.0);
    __syncthreads();
    line_fft<N, SIGN>(a, n1, ex[l], tw);
    if constexpr (!BIN) {
        if (live)
#pragma unroll
            for (int u = 0; u < N2 / N1; ++u)
#pragma unroll
                for (int k1 = 0; k1 < N1; ++k1)
                    H[base + (size_t)(n1 + N1 * u + N2 * k1) * stride] = a[u * N1 + k1];
    } else {
        double *wbins = &ex[0][0].x;   // [NW][nb], within col_bin_capacity<N>()
        const int wave = NT >= 64 ? (int)(threadIdx.x >> 6) : 0;
        __syncthreads();   // every thread's last read of ex is done
        for (int e = threadIdx.x; e < NW * nb; e += NT) wbins[e] = 0.0;
        __syncthreads();
        if (live) {
            const int py = o <= N / 2 ? o : o - N;
            const double w = (kz == 0 || kz == N / 2) ? 1.0 : 2.0;
#pragma unroll
            for (int u = 0; u < N2 / N1; ++u)
#pragma unroll
                for (int k1 = 0; k1 < N1; ++k1) {
                    const int i = n1 + N1 * u + N2 * k1;
                    const int px = i <= N / 2 ? i : i - N;
                    const int bin = shell_of(px * px + py * py + kz * kz);
                    double2 F = a[u * N1 + k1];
                    if (i == 0 && o == 0 && kz == 0) F.x += dc_add;
                    if (bin < nb) atomicAdd(&wbins[wave * nb + bin], w * (F.x * F.x + F.y * F.y));
                }
        }
        __syncthreads();
        for (int b = threadIdx.x; b < nb; b += NT) {
            double t = 0.0;
            for (int q = 0; q < NW; ++q) t += wbins[q * nb + b];
            partials[(size_t)lid * nb + b] = t;
        }
    }
}

// Fixed-order sum of per-workgroup shell partials, one workgroup per shell.
__global__ __launch_bounds__(256) void k_sum_bins_tree(const double *__restrict__ partials, int nblocks, int nb,
                                                        double *__restrict__ out)
{
    __shared__ double red[256];
    const int b = blockIdx.x;
    double t = 0.0;
    for (int q = threadIdx.x; q < nblocks; q += 256) t += partials[(size_t)q * nb + b];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = red[0];
}

int ilog2(int n)
{
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

}  // namespace

struct FftPlan {
    int N = 0, logN = 0, K = 0;
    size_t modes = 0;
    double2 *H[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double2 *tw = nullptr;
    double *partials = nullptr;
    double *sums = nullptr;
    int bin_blocks = 0;
};

int fft_numbins(int N) { return (int)(1.73205 * (N / 2)) + 1; }
int fft_gw_numbins(int N) { return (int)(sqrt(3.0) * (N / 2)) + 1; }

static int bin_blocks_for(size_t modes)
{
    size_t b = (modes + kThreads * 8 - 1) / (kThreads * 8);
    if (b > 1024) b = 1024;
    return (int)(b < 1 ? 1 : b);
}

// Lines of the register-resident transforms: 8 <= N <= 256.
static bool reg_path(int N) { return N >= 8 && N <= 256; }
static int col_groups(int N) { return (N / 2 + 1 + kColLines - 1) / kColLines; }

// Rows of shell partials: the binning kernels' workgroups, or the x pass's of a
// fused spectrum, whichever is more.
static int partial_rows(int N)
{
    const size_t modes = (size_t)N * N * (N / 2 + 1);
    return std::max(bin_blocks_for(modes), reg_path(N) ? N * col_groups(N) : 0);
}

size_t fft_workspace_bytes(int N)
{
    const size_t K = (size_t)N / 2 + 1;
    const size_t modes = (size_t)N * N * K;
    const int nb = fft_gw_numbins(N);
    return modes * sizeof(double2) + (size_t)N * sizeof(double2) + (size_t)partial_rows(N) * nb * 8 +
           (size_t)nb * 8;
}

// Calls f with std::integral_constant<int, N> for a register-path N.
template <typename F>
static hipError_t with_line(int N, F &&f)
{
    switch (N) {
    case 8: return f(std::integral_constant<int, 8>{});
    case 16: return f(std::integral_constant<int, 16>{});
    case 32: return f(std::integral_constant<int, 32>{});
    case 64: return f(std::integral_constant<int, 64>{});
    case 128: return f(std::integral_constant<int, 128>{});
    case 256: return f(std::integral_constant<int, 256>{});
    default: return hipErrorInvalidValue;
    }
}

template <int N, int SIGN, int DIM, bool BIN>
static hipError_t launch_cols(double2 *H, const double2 *tw, int nb, double dc_add, double *partials, hipStream_t s)
{
    constexpr int N1 = LineShape<N>::N1, NT = kColLines * N1, NW = NT >= 64 ? NT / 64 : 1;
    if (BIN && (size_t)NW * nb > col_bin_capacity<N>()) return hipErrorInvalidValue;
    const dim3 grid(col_groups(N), N);
    k_cols_reg<N, SIGN, DIM, BIN><<<grid, NT, 0, s>>>(H, tw, nb, dc_add, partials);
    return hipGetLastError();
}

void fft_twiddles(int N, double2 *tw)
{
    const double two_pi = 4.0 * asin(1.0);
    for (int k = 0; k < N; ++k) tw[k] = make_double2(cos(two_pi * k / N), sin(two_pi * k / N));
}

// Parts of a one-slot plan over caller memory, each 256-byte aligned: the half
// spectrum, the binning partials, the shell sums.
static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t fft_view_bytes(int N)
{
    const size_t modes = (size_t)N * N * (N / 2 + 1);
    const int nb = fft_gw_numbins(N);
    return align256(modes * sizeof(double2)) + align256((size_t)partial_rows(N) * nb * 8) +
           align256((size_t)nb * 8);
}

FftPlan *fft_create(int N)
{
    FftPlan *p = new FftPlan();
    p->N = N;
    p->logN = ilog2(N);
    p->K = N / 2 + 1;
    p->modes = (size_t)N * N * p->K;
    p->bin_blocks = bin_blocks_for(p->modes);
    const int nb = fft_gw_numbins(N);
    bool ok = hipMalloc(&p->H[0], p->modes * sizeof(double2)) == hipSuccess &&
              hipMalloc(&p->tw, (size_t)N * sizeof(double2)) == hipSuccess &&
              hipMalloc(&p->partials, (size_t)partial_rows(N) * nb * 8) == hipSuccess &&
              hipMalloc(&p->sums, (size_t)nb * 8) == hipSuccess;
    if (!ok) {
        fft_destroy(p);
        return nullptr;
    }
    std::vector<double2> tw(N);
    fft_twiddles(N, tw.data());
    if (hipMemcpy(p->tw, tw.data(), N * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
        fft_destroy(p);
        return nullptr;
    }
    return p;
}

void fft_destroy(FftPlan *p)
{
    if (!p) return;
    for (int s = 0; s < 6; ++s)
        if (p->H[s]) (void)hipFree(p->H[s]);
    if (p->tw) (void)hipFree(p->tw);
    if (p->partials) (void)hipFree(p->partials);
    if (p->sums) (void)hipFree(p->sums);
    delete p;
}

static hipError_t ensure_slot(FftPlan *p, int slot)
{
    if (slot < 0 || slot > 5) return hipErrorInvalidValue;
    if (p->H[slot]) return hipSuccess;
    return hipMalloc(&p->H[slot], p->modes * sizeof(double2));
}

// Rows of the z pass per workgroup: fills kLdsComplex, never more than the N^2
// rows there are (both powers of two, so L divides N^2).
static int rows_per_block(int N)
{
    int L = kLdsComplex / N;
    if (L > N * N) L = N * N;
    return L < 1 ? 1 : L;
}

static hipError_t columns(FftPlan *p, double2 *H, int sign, hipStream_t s)
{
    const int N = p->N;
    if (reg_path(N))
        return with_line(N, [&](auto n) {
            constexpr int NN = decltype(n)::value;
            hipError_t e = sign > 0 ? launch_cols<NN, +1, 1, false>(H, p->tw, 0, 0.0, nullptr, s)
                                    : launch_cols<NN, -1, 1, false>(H, p->tw, 0, 0.0, nullptr, s);
            if (e != hipSuccess) return e;
            return sign > 0 ? launch_cols<NN, +1, 0, false>(H, p->tw, 0, 0.0, nullptr, s)
                            : launch_cols<NN, -1, 0, false>(H, p->tw, 0, 0.0, nullptr, s);
        });
    int L = kLdsComplex / N;
    if (L > 8) L = 8;
    if (L < 1) L = 1;
    const dim3 grid((p->K + L - 1) / L, N);
    const size_t lds = (size_t)L * N * sizeof(double2);
    k_fft_cols<<<grid, kThreads, lds, s>>>(H, N, p->logN, 1, L, p->tw, sign);
    k_fft_cols<<<grid, kThreads, lds, s>>>(H, N, p->logN, 0, L, p->tw, sign);
    return hipGetLastError();
}

// z pass of a forward transform into H.
template <typename T>
static hipError_t z_forward(FftPlan *p, const T *in, double2 *H, hipStream_t s)
{
    const int N = p->N;
    if (reg_path(N))
        return with_line(N, [&](auto n) {
            constexpr int NN = decltype(n)::value, L = kRegThreads / LineShape<NN>::N1;
            k_z_r2c_reg<T, NN><<<NN * NN / L, kRegThreads, 0, s>>>(in, H, p->tw);
            return hipGetLastError();
        });
    const int L = rows_per_block(N);
    k_fft_z_r2c<T><<<N * N / L, kThreads, (size_t)L * N * sizeof(double2), s>>>(in, H, N, p->logN, L, p->tw);
    return hipGetLastError();
}

// z pass of an inverse transform: out = scale Re + add.
static hipError_t z_inverse(FftPlan *p, const double2 *H, double *out, double scale, double add, hipStream_t s)
{
    const int N = p->N;
    if (reg_path(N))
        return with_line(N, [&](auto n) {
            constexpr int NN = decltype(n)::value, L = kRegThreads / LineShape<NN>::N1;
            k_z_c2r_reg<NN><<<NN * NN / L, kRegThreads, 0, s>>>(H, out, p->tw, scale, add);
            return hipGetLastError();
        });
    const int L = rows_per_block(N);
    k_fft_z_c2r<<<N * N / L, kThreads, (size_t)L * N * sizeof(double2), s>>>(H, out, N, p->logN, L, p->tw, scale,
                                                                             add);
    return hipGetLastError();
}

template <typename T>
static hipError_t forward(FftPlan *p, const T *in, int slot, hipStream_t s)
{
    hipError_t e = ensure_slot(p, slot);
    if (e != hipSuccess) return e;
    if ((e = z_forward<T>(p, in, p->H[slot], s)) != hipSuccess) return e;
    return columns(p, p->H[slot], +1, s);
}

// Forward transform and shell sums of |F|^2 (dc_add on the k = 0 amplitude),
// the x pass binning its modes instead of storing them (slot 0 holds the
// partially transformed field afterwards).
template <typename T>
static hipError_t spectrum(FftPlan *p, const T *in, double dc_add, double *sums_host, hipStream_t s)
{
    const int N = p->N, nb = fft_numbins(N);
    hipError_t e = ensure_slot(p, 0);
    if (e != hipSuccess) return e;
    if (!reg_path(N)) {
        if ((e = forward<T>(p, in, 0, s)) != hipSuccess) return e;
        return fft_bin_scalar_dc(p, 0, dc_add, sums_host, s);
    }
    if ((e = z_forward<T>(p, in, p->H[0], s)) != hipSuccess) return e;
    e = with_line(N, [&](auto n) {
        constexpr int NN = decltype(n)::value;
        hipError_t q = launch_cols<NN, +1, 1, false>(p->H[0], p->tw, 0, 0.0, nullptr, s);
        if (q != hipSuccess) return q;
        return launch_cols<NN, +1, 0, true>(p->H[0], p->tw, nb, dc_add, p->partials, s);
    });
    if (e != hipSuccess) return e;
    k_sum_bins_tree<<<nb, 256, 0, s>>>(p->partials, N * col_groups(N), nb, p->sums);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(sums_host, p->sums, (size_t)nb * 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

hipError_t fft_spectrum_double(FftPlan *p, const double *in, double *sums_host, hipStream_t s)
{
    return spectrum<double>(p, in, 0.0, sums_host, s);
}

hipError_t fft_spectrum_float(FftPlan *p, const float *in, double dc_add, double *sums_host, hipStream_t s)
{
    return spectrum<float>(p, in, dc_add, sums_host, s);
}

hipError_t fft_forward_double(FftPlan *p, const double *in, int slot, hipStream_t s)
{
    return forward<double>(p, in, slot, s);
}

hipError_t fft_forward_float(FftPlan *p, const float *in, double dc_add, int slot, hipStream_t s)
{
    (void)dc_add;  // applied at binning time (only the k = 0 mode changes)
    return forward<float>(p, in, slot, s);
}

static hipError_t finish_bins(FftPlan *p, int nb, double *sums_host, hipStream_t s)
{
    k_sum_bins_tree<<<nb, 256, 0, s>>>(p->partials, p->bin_blocks, nb, p->sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = hipMemcpyAsync(sums_host, p->sums, (size_t)nb * 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

hipError_t fft_bin_scalar_dc(FftPlan *p, int slot, double dc_add, double *sums_host, hipStream_t s)
{
    const int nb = fft_numbins(p->N);
    k_bin_scalar<<<p->bin_blocks, kThreads, (size_t)(kThreads / 64) * nb * 8, s>>>(p->H[slot], p->N, nb, dc_add,
                                                                                   p->partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return finish_bins(p, nb, sums_host, s);
}

hipError_t fft_bin_scalar(FftPlan *p, int slot, double *sums_host, hipStream_t s)
{
    return fft_bin_scalar_dc(p, slot, 0.0, sums_host, s);
}

hipError_t fft_spectrum_in(int N, void *work, const double2 *tw, const double *in, double *sums_host,
                           hipStream_t s)
{
    FftPlan p;
    p.N = N;
    p.logN = ilog2(N);
    p.K = N / 2 + 1;
    p.modes = (size_t)N * N * p.K;
    p.bin_blocks = bin_blocks_for(p.modes);
    char *w = (char *)work;
    p.H[0] = (double2 *)w;
    w += align256(p.modes * sizeof(double2));
    p.partials = (double *)w;
    w += align256((size_t)partial_rows(N) * fft_gw_numbins(N) * 8);
    p.sums = (double *)w;
    p.tw = (double2 *)tw;
    hipError_t e = spectrum<double>(&p, in, 0.0, sums_host, s);
    p.H[0] = nullptr;   // caller memory: nothing for the destructor-free plan to own
    return e;
}

hipError_t fft_bin_tt(FftPlan *p, double dx, double to_phys, double *sums_host, hipStream_t s)
{
    const int nb = fft_gw_numbins(p->N);
    const int imax = (int)floor(0.80 * nb);
    SixSpectra S;
    for (int c = 0; c < 6; ++c) {
        if (!p->H[c]) return hipErrorInvalidValue;
        S.h[c] = p->H[c];
    }
    k_bin_tt<<<p->bin_blocks, kThreads, ((size_t)(kThreads / 64) * nb + p->N + 1) * 8, s>>>(S, p->N, nb, imax, 2.0 / dx,
                                                                                           to_phys, p->partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return finish_bins(p, nb, sums_host, s);
}

hipError_t fft_inverse_ffteasy(FftPlan *p, const double *fk, const double *fny, double add, double *out,
                               hipStream_t s)
{
    const int N = p->N, K = p->K;
    std::vector<double2> h((size_t)N * N * K);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            const size_t row = (size_t)i * N + j;
            for (int kz = 0; kz < N / 2; ++kz)
                h[row * K + kz] = make_double2(fk[row * N + 2 * kz], fk[row * N + 2 * kz + 1]);
            h[row * K + N / 2] = make_double2(fny[(size_t)i * 2 * N + 2 * j], fny[(size_t)i * 2 * N + 2 * j + 1]);
        }
    hipError_t e = hipMemcpyAsync(p->H[0], h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    if ((e = columns(p, p->H[0], -1, s)) != hipSuccess) return e;
    if ((e = z_inverse(p, p->H[0], out, 1.0 / ((double)N * N * N), add, s)) != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

const double2 *fft_slot(const FftPlan *p, int slot) { return slot >= 0 && slot < 6 ? p->H[slot] : nullptr; }

hipError_t fft_kill_modes(FftPlan *p, int slot, double hi, double lo, hipStream_t s)
{
    if (!p->H[slot]) return hipErrorInvalidValue;
    k_kill_modes<<<bin_blocks_for(p->modes) * 8, kThreads, 0, s>>>(p->H[slot], p->N, hi, lo);
    return hipGetLastError();
}

hipError_t fft_zero_dc(FftPlan *p, int slot, hipStream_t s)
{
    if (!p->H[slot]) return hipErrorInvalidValue;
    return hipMemsetAsync(p->H[slot], 0, sizeof(double2), s);
}

hipError_t fft_inverse_slot(FftPlan *p, int slot, double add, double *out, hipStream_t s)
{
    if (!p->H[slot]) return hipErrorInvalidValue;
    const int N = p->N;
    hipError_t e = columns(p, p->H[slot], -1, s);
    if (e != hipSuccess) return e;
    return z_inverse(p, p->H[slot], out, 1.0 / ((double)N * N * N), add, s);
}

void fft_counts(int N, int *counts)
{
    const int nb = fft_numbins(N);
    for (int b = 0; b < nb; ++b) counts[b] = 0;
    for (int i = 0; i < N; ++i) {
        const int px = i <= N / 2 ? i : i - N;
        for (int j = 0; j < N; ++j) {
            const int py = j <= N / 2 ? j : j - N;
            for (int kz = 0; kz <= N / 2; ++kz) {
                const int bin = (int)lround(sqrt((double)px * px + (double)py * py + (double)kz * kz));
                if (bin >= nb) continue;
                counts[bin] += (kz == 0 || kz == N / 2) ? 1 : 2;
            }
        }
    }
}

void fft_gw_counts(int N, int *counts)
{
    const int nb = fft_gw_numbins(N), imax = (int)floor(0.80 * nb);
    for (int b = 0; b < nb; ++b) counts[b] = 0;
    for (int i = 0; i < N; ++i) {
        const int px = i <= N / 2 ? i : i - N;
        for (int j = 0; j < N; ++j) {
            const int py = j <= N / 2 ? j : j - N;
            for (int kz = 0; kz <= N / 2; ++kz) {
                if (px == 0 && py == 0 && kz == 0) continue;
                const int bin = (int)lround(sqrt((double)px * px + (double)py * py + (double)kz * kz));
                if (bin >= imax) continue;
                counts[bin] += (kz == 0 || kz == N / 2) ? 1 : 2;
            }
        }
    }
}

// Shell counts depend on N alone (a triple loop over the N^2 (N/2 + 1) modes, 25 ms
// at 256^3 on the host): computed once per N and kind and kept, under a lock so
// that runs stepped from several threads share them.
const int *fft_counts_cached(int N, int gw)
{
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::vector<int>> cache;
    std::lock_guard<std::mutex> lock(mu);
    std::vector<int> &v = cache[{N, gw ? 1 : 0}];
    if (v.empty()) {
        v.resize((size_t)(gw ? fft_gw_numbins(N) : fft_numbins(N)));
        if (gw) fft_gw_counts(N, v.data());
        else fft_counts(N, v.data());
    }
    return v.data();
}

}  // namespace ie
