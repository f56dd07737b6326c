// ie_diag.hip -- output diagnostics that are off by default in the reference:
// the equilateral bispectrum (bispectraf, src/output.cpp:740-911) and the
// log-mapped zeta (build_deltaN_log, histogramsLOG, src/output.cpp:167-198,
// 1312-1403).
//
// Bispectrum.  For each shell k the reference walks every pair (m1, m2) of modes
// of its shell set S_k (|m| within 1.5 of k, kz from round(sqrt(k^2 - px^2 -
// py^2)) - 1 .. + 1, 0 <= kz <= N/2) and adds F(m1) F(m2) F(m3) for the closing
// modes m3 = m1 + m2 (and, for kz1, kz2 > 0, the mirror with kz3 = kz1 - kz2) that
// fall in the shell.  The host lists S_k (N^2 x 3 tests per shell); one workgroup
// takes one m1 of one shell and its threads stride over the m2 of that shell,
// reading the FP64 half spectrum (L2-resident up to N = 128); the partials are
// summed per shell in a fixed order.  The pair set and each term are the
// reference's; the sum order is not.
#include <math.h>

#include <cmath>
#include <vector>

#include "ie_fft.h"
#include "ie_internal.h"

namespace ie {

namespace {

constexpr int kDiagThreads = 256;

struct BisItem {
    int k;        // shell
    int m1;       // index of m1 in the mode list
    int begin;    // the shell's modes: [begin, begin + len)
    int len;
};

__device__ __forceinline__ double2 mode_at(const double2 *__restrict__ H, int N, int px, int py, int kz)
{
    const int i = px >= 0 ? px : px + N, j = py >= 0 ? py : py + N;
    return H[((size_t)i * N + j) * (N / 2 + 1) + kz];
}

__global__ __launch_bounds__(kDiagThreads) void k_bispectrum(const double2 *__restrict__ H, int N,
                                                              const BisItem *__restrict__ items,
                                                              const int3 *__restrict__ modes,
                                                              double *__restrict__ partials)
{
    __shared__ double smem[3 * (kDiagThreads / 64)];
    const BisItem it = items[blockIdx.x];
    const double kf = (double)it.k;
    const int3 a = modes[it.m1];
    const int px1 = a.x, py1 = a.y, k1 = a.z;
    const double2 F1 = mode_at(H, N, px1, py1, k1);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int b = it.begin + threadIdx.x; b < it.begin + it.len; b += blockDim.x) {
        const int3 m2 = modes[b];
        const int px2 = m2.x, py2 = m2.y, k2 = m2.z;
        const int px3 = px1 + px2, py3 = py1 + py2;
        if (!(px3 <= N / 2 && px3 > -N / 2 && py3 <= N / 2 && py3 > -N / 2)) continue;
        const double q3 = (double)px3 * px3 + (double)py3 * py3;
        int k3 = k1 + k2;
        if (fabs(sqrt(q3 + (double)k3 * k3) - kf) < 1.5 && k3 <= N / 2) {
            const double2 F2 = mode_at(H, N, px2, py2, k2), F3 = mode_at(H, N, px3, py3, k3);
            const double counts = (k1 != N / 2 && k2 != N / 2 && (k1 != 0 || k2 != 0)) ? 2.0 : 1.0;
            const double f1r = F1.x, f1i = F1.y, f2r = F2.x, f2i = F2.y, f3r = F3.x, f3i = F3.y;
            acc[0] += counts * (f1r * f2r * f3r - f1i * f2i * f3r + f1i * f2r * f3i + f2i * f1r * f3i);
            acc[1] += counts * (-f1r * f2r * f3i + f1i * f2i * f3i + f1i * f2r * f3r + f2i * f1r * f3r);
            acc[2] += counts;
        }
        if (k1 != 0 && k2 != 0) {
            k3 = k1 - k2;
            if (k3 >= 0 && fabs(sqrt(q3 + (double)k3 * k3) - kf) < 1.5 && k3 <= N / 2) {
                // F(-px2, -py2, kz2) conjugated
                const double2 F2 = mode_at(H, N, -px2, -py2, k2);
                const double2 F3 = mode_at(H, N, px3, py3, k3);
                const double f1r = F1.x, f1i = F1.y, f2r = F2.x, f2i = -F2.y, f3r = F3.x, f3i = F3.y;
                acc[0] += 2.0 * (f1r * f2r * f3r - f1i * f2i * f3r + f1i * f2r * f3i + f2i * f1r * f3i);
                acc[1] += 2.0 * (-f1r * f2r * f3i + f1i * f2i * f3i + f1i * f2r * f3r + f2i * f1r * f3r);
                acc[2] += 2.0;
            }
        }
    }
    block_sum<3>(acc, smem);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) partials[(size_t)blockIdx.x * 3 + q] = acc[q];
}

// per shell: the items of the shell are contiguous; sum them in order
__global__ void k_bis_shells(const double *__restrict__ partials, const int *__restrict__ first, int nshells,
                             double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nshells) return;
    double s[3] = {0.0, 0.0, 0.0};
    for (int it = first[k]; it < first[k + 1]; ++it)
        for (int q = 0; q < 3; ++q) s[q] += partials[(size_t)it * 3 + q];
    for (int q = 0; q < 3; ++q) out[k * 3 + q] = s[q];
}

// ---------------------------------------------------------------- LOG map

// sum of fd a^{s-1} / (ad a^{s-2}) and of f (build_deltaN_log's factt, fmean)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_log_sums(const T *__restrict__ f, const T *__restrict__ fd, size_t gs,
                                                      double fbar, double pbar, double P1, double Q,
                                                      double *__restrict__ partials)
{
    __shared__ double smem[2 * kWaves];
    double acc[2] = {0.0, 0.0};
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        const double v = pbar + (double)fd[id];
        acc[0] += v * P1 / Q;
        acc[1] += fbar + (double)f[id];
    }
    block_sum<2>(acc, smem);
    if (threadIdx.x == 0) {
        partials[blockIdx.x * 2] = acc[0];
        partials[blockIdx.x * 2 + 1] = acc[1];
    }
}

__global__ void k_sum2_final(const double *__restrict__ partials, int nb, double *__restrict__ out)
{
    __shared__ double smem[2 * kWaves];
    double tot[2];
    sum_partials<2>(partials, nb, tot, smem);
    if (threadIdx.x == 0) {
        out[0] = tot[0];
        out[1] = tot[1];
    }
}

__device__ __forceinline__ double log_arg(double fv, double eta, double fmean, double factt)
{
    return 1.0 - eta * (fv - fmean) / factt;
}

// map into dn; masked min / max (sites with a positive argument) per block
template <typename T>
__global__ __launch_bounds__(kBlock) void k_log_map(const T *__restrict__ f, size_t gs, double fbar, double eta,
                                                     double fmean, double factt, double *__restrict__ dn,
                                                     double *__restrict__ mm)
{
    __shared__ double smn[kWaves], smx[kWaves];
    double lo = INFINITY, hi = -INFINITY;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        const double arg = log_arg(fbar + (double)f[id], eta, fmean, factt);
        double v = 0.0;
        if (arg > 0.0) {
            v = (1.0 / eta) * log(arg);
            lo = fmin(lo, v);
            hi = fmax(hi, v);
        }
        dn[id] = v;
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, 64));
        hi = fmax(hi, __shfl_xor(hi, off, 64));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        smn[wave] = lo;
        smx[wave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kWaves; ++w) {
            lo = fmin(lo, smn[w]);
            hi = fmax(hi, smx[w]);
        }
        mm[blockIdx.x * 2] = smn[0] < lo ? smn[0] : lo;
        mm[blockIdx.x * 2 + 1] = smx[0] > hi ? smx[0] : hi;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_log_hist(const T *__restrict__ f, const double *__restrict__ dn,
                                                      size_t gs, double fbar, double eta, double fmean, double factt,
                                                      double bmin, double bmax, double df, int nbins,
                                                      unsigned long long *__restrict__ counts)
{
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        if (!(log_arg(fbar + (double)f[id], eta, fmean, factt) > 0)) continue;
        const double v = dn[id];
        int b = (int)((v - bmin) / df);
        if (v == bmax) b = nbins - 1;
        if (b >= 0 && b < nbins) atomicAdd(&counts[b], 1ull);
    }
}

__global__ __launch_bounds__(kBlock) void k_store_float(const double *__restrict__ src, float *__restrict__ dst,
                                                         size_t gs)
{
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x)
        dst[id] = (float)src[id];
}

unsigned grid_of(size_t gs)
{
    size_t b = (gs + kBlock - 1) / kBlock;
    if (b > 1024) b = 1024;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

hipError_t store_float(const double *src, float *dst, size_t gs, hipStream_t s)
{
    k_store_float<<<grid_of(gs), kBlock, 0, s>>>(src, dst, gs);
    return hipGetLastError();
}

// Shell sets S_k of bispectraf, in the reference's (i1, j1, k1) order per shell.
void bis_shell_modes(int N, std::vector<int> &first, std::vector<int> &px, std::vector<int> &py,
                     std::vector<int> &kz)
{
    const int numbins = (int)(sqrt(3.0) * (N / 2)) + 1;
    first.assign(numbins + 1, 0);
    px.clear();
    py.clear();
    kz.clear();
    for (int k = 0; k < numbins; ++k) {
        first[k] = (int)px.size();
        const double kf = (double)k;
        for (int i1 = 0; i1 < N; ++i1)
            for (int j1 = 0; j1 < N; ++j1) {
                const double px1 = i1 <= N / 2 ? i1 : i1 - N;
                const double py1 = j1 <= N / 2 ? j1 : j1 - N;
                const double pz = kf * kf - px1 * px1 - py1 * py1;
                if (!(pz >= 0)) continue;
                const int r = (int)round(sqrt(pz));
                for (int k1 = r - 1; k1 < r + 2; ++k1)
                    if (fabs(sqrt(px1 * px1 + py1 * py1 + (double)k1 * k1) - kf) < 1.5 && k1 <= N / 2 && k1 >= 0) {
                        px.push_back((int)px1);
                        py.push_back((int)py1);
                        kz.push_back(k1);
                    }
            }
    }
    first[numbins] = (int)px.size();
}

hipError_t bispectrum_sums(const FftPlan *plan, int slot, int N, double *out_host, hipStream_t s)
{
    std::vector<int> first, px, py, kz;
    bis_shell_modes(N, first, px, py, kz);
    const int nshells = (int)first.size() - 1;
    std::vector<int3> modes(px.size());
    for (size_t q = 0; q < px.size(); ++q) modes[q] = make_int3(px[q], py[q], kz[q]);
    std::vector<BisItem> items;
    std::vector<int> ifirst(nshells + 1, 0);
    for (int k = 0; k < nshells; ++k) {
        ifirst[k] = (int)items.size();
        for (int m = first[k]; m < first[k + 1]; ++m) items.push_back({k, m, first[k], first[k + 1] - first[k]});
    }
    ifirst[nshells] = (int)items.size();
    int3 *d_modes = nullptr;
    BisItem *d_items = nullptr;
    int *d_first = nullptr;
    double *d_part = nullptr, *d_out = nullptr;
    hipError_t e = hipSuccess;
    auto A = [&](void **p, size_t b) {
        if (e == hipSuccess) e = hipMallocAsync(p, b ? b : 8, s);
    };
    A((void **)&d_modes, modes.size() * sizeof(int3));
    A((void **)&d_items, items.size() * sizeof(BisItem));
    A((void **)&d_first, ifirst.size() * sizeof(int));
    A((void **)&d_part, items.size() * 3 * sizeof(double));
    A((void **)&d_out, (size_t)nshells * 3 * sizeof(double));
    if (e == hipSuccess && !modes.empty())
        e = hipMemcpyAsync(d_modes, modes.data(), modes.size() * sizeof(int3), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !items.empty())
        e = hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(BisItem), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_first, ifirst.data(), ifirst.size() * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !items.empty()) {
        k_bispectrum<<<(unsigned)items.size(), kDiagThreads, 0, s>>>(fft_slot(plan, slot), N, d_items, d_modes,
                                                                      d_part);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        k_bis_shells<<<(nshells + 63) / 64, 64, 0, s>>>(d_part, d_first, nshells, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out_host, d_out, (size_t)nshells * 3 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (void *p : {(void *)d_modes, (void *)d_items, (void *)d_first, (void *)d_part, (void *)d_out})
        if (p) (void)hipFreeAsync(p, s);
    (void)hipStreamSynchronize(s);
    return e;
}

template <typename T>
static hipError_t log_map_t(const T *f, const T *fd, size_t gs, double fbar, double pbar, double a, double ad,
                            double s_rescale, double eta, int nbins, double *dn, double *work,
                            unsigned long long *counts_dev, LogMapResult *res, hipStream_t s)
{
    const unsigned G = grid_of(gs);
    const double P1 = pow(a, s_rescale - 1.0), Q = ad * pow(a, s_rescale - 2.0);
    k_log_sums<T><<<G, kBlock, 0, s>>>(f, fd, gs, fbar, pbar, P1, Q, work);
    k_sum2_final<<<1, kBlock, 0, s>>>(work, (int)G, work + 4000);
    double sums[2];
    hipError_t e = hipMemcpyAsync(sums, work + 4000, 16, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    const double factt = sums[0] / (double)gs, fmean = sums[1] / (double)gs;
    res->factt = factt;
    res->fmean = fmean;
    res->valid = !(!isfinite(factt) || fabs(factt) < 1e-30);
    if (!res->valid) {
        e = hipMemsetAsync(dn, 0, gs * 8, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        res->bmin = 0.0;
        res->df = 1.0;
        return e;
    }
    k_log_map<T><<<G, kBlock, 0, s>>>(f, gs, fbar, eta, fmean, factt, dn, work);
    std::vector<double> mm(2 * G);
    double d0 = 0.0;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(mm.data(), work, 2 * G * 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(&d0, dn, 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    // histogramsLOG starts its min and max at site 0's value, masked or not
    double bmin = d0, bmax = d0;
    for (unsigned q = 0; q < G; ++q) {
        bmin = mm[2 * q] < bmin ? mm[2 * q] : bmin;
        bmax = mm[2 * q + 1] > bmax ? mm[2 * q + 1] : bmax;
    }
    double df = (bmax - bmin) / (double)nbins;
    if (!std::isfinite(df) || df <= 0.0) df = 1.0;
    if ((e = hipMemsetAsync(counts_dev, 0, (size_t)nbins * 8, s)) != hipSuccess) return e;
    k_log_hist<T><<<G, kBlock, 0, s>>>(f, dn, gs, fbar, eta, fmean, factt, bmin, bmax, df, nbins, counts_dev);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    res->bmin = bmin;
    res->df = df;
    return hipStreamSynchronize(s);
}

hipError_t log_map(const void *f, const void *fd, int split, size_t gs, double fbar, double pbar, double a,
                   double ad, double s_rescale, double eta, int nbins, double *dn, double *work,
                   unsigned long long *counts_dev, LogMapResult *res, hipStream_t s)
{
    if (split)
        return log_map_t<float>((const float *)f, (const float *)fd, gs, fbar, pbar, a, ad, s_rescale, eta, nbins,
                                dn, work, counts_dev, res, s);
    return log_map_t<double>((const double *)f, (const double *)fd, gs, 0.0, 0.0, a, ad, s_rescale, eta, nbins, dn,
                             work, counts_dev, res, s);
}

}  // namespace ie
