// ie_noalloc.hip -- the entry point that allocates nothing (ie_spectrum_noalloc).
//
// The isotropic power spectrum of an FP64 field the caller already holds in
// device memory, with the kernels of ie_spectrum (forward r2c transform, shell
// binning with the reference's weights, src/output.cpp:63-151), for callers
// that cannot allocate device memory.  The kernels' workspace -- one half
// spectrum, the binning partials and shell sums, and the twiddle tables of
// every supported N -- is the file-scope array `scratch` below: a GPU that runs
// this entry point holds it (kScratchBytes) until the process ends.  Its device
// address is looked up once per device (hipGetSymbolAddress, with that device
// current) and only that address is handed to kernels and copies; the name
// itself is a host-side stand-in.  Every call on a GPU shares that GPU's array:
// the transform and the binning overwrite every element they read, so nothing
// needs clearing between calls, and calls that overlap on one GPU race.
#include <math.h>
#include <string.h>

#include <mutex>
#include <string>

#include "ie_api.h"
#include "ie_fft.h"
#include "ie_internal.h"

namespace {

constexpr int kMinN = 8, kMaxN = 128;
constexpr int kMaxDevices = 64;

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Largest plan workspace (N = kMaxN, as fft_view_bytes computes it: half
// spectrum N^2 (N/2+1) complex doubles, <= 1024 x gw_numbins binning partials,
// gw_numbins sums) and the twiddles of N = 8, 16, ..., kMaxN.
constexpr size_t kGwBinsMax = 111;   // int(sqrt(3) * 64) + 1
constexpr size_t kWorkBytes = align256((size_t)kMaxN * kMaxN * (kMaxN / 2 + 1) * 16) +
                              align256((size_t)1024 * kGwBinsMax * 8) + align256(kGwBinsMax * 8);
constexpr size_t kTwiddles = 2 * kMaxN;   // 8 + 16 + ... + kMaxN < 2 kMaxN
constexpr size_t kScratchBytes = kWorkBytes + align256(kTwiddles * 16);

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// Offset of the twiddles of size N in the twiddle region.
size_t twiddle_offset(int N)
{
    size_t off = 0;
    for (int n = kMinN; n < N; n *= 2) off += (size_t)n;
    return off;
}

// the message ie_last_error(NULL) returns on this thread
ie_status fail(ie_status st, const std::string &m)
{
    ie::set_unbound_error(m);
    return st;
}

// The device address of this device's scratch, looked up (and its twiddle
// tables written) on the first call on the device.
ie_status device_scratch(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(IE_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(IE_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(IE_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        double2 tw[kTwiddles];
        for (int n = kMinN; n <= kMaxN; n *= 2) ie::fft_twiddles(n, tw + twiddle_offset(n));
        e = hipMemcpy((char *)ptr + kWorkBytes, tw, sizeof(tw), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(IE_ERR_HIP, std::string("twiddle upload: ") + hipGetErrorString(e));
        g_scratch[dev] = ptr;
    }
    *out = g_scratch[dev];
    return IE_OK;
}

}  // namespace

extern "C" {

ie_status ie_spectrum_noalloc(const double *f_dev, int N, double L, double norm1, double *p, int *counts,
                              double *vals, void *stream)
{
    if (!f_dev || !p || !counts || !vals) return fail(IE_ERR_ARG, "null argument");
    if (N < kMinN || N > kMaxN || (N & (N - 1)))
        return fail(IE_ERR_ARG, "ie_spectrum_noalloc: N must be a power of two in [" + std::to_string(kMinN) + ", " +
                                    std::to_string(kMaxN) + "] (its workspace is a fixed " +
                                    std::to_string(kScratchBytes) + "-byte device array)");
    if (!(L > 0.0)) return fail(IE_ERR_ARG, "L must be positive");
    if (ie::fft_view_bytes(N) > kWorkBytes) return fail(IE_ERR_ARG, "workspace larger than the scratch array");
    void *base = nullptr;
    ie_status st = device_scratch(&base);
    if (st != IE_OK) return st;
    const int nb = ie::fft_numbins(N);
    double sums[kGwBinsMax + 1];
    const double2 *tw = (const double2 *)((char *)base + kWorkBytes) + twiddle_offset(N);
    const hipError_t e = ie::fft_spectrum_in(N, base, tw, f_dev, sums, (hipStream_t)stream);
    if (e != hipSuccess) return fail(IE_ERR_HIP, std::string("spectrum kernels: ") + hipGetErrorString(e));
    ie::fft_counts(N, counts);
    const double pi = 2.0 * asin(1.0), dp = 2.0 * pi / L;
    for (int b = 0; b < nb; ++b) {
        p[b] = dp * b;
        double v = sums[b];
        if (counts[b] > 0) v /= (double)counts[b];
        vals[b] = norm1 * v;
    }
    return IE_OK;
}

}  // extern "C"
