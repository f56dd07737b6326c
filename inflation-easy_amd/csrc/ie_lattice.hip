// ie_lattice.hip -- leapfrog step of the inflaton lattice on gfx950.
//
// One step of run_evolution_loop (src/evolution.cpp:937-945) is three launches on
// the run's stream, with no host involvement:
//   1. k_friedmann_reduce: sum_x -f lapl f and sum_x V(f) (gradient_energy and
//      potential_energy, src/evolution.cpp:130-136, src/potential.cpp:151-171),
//      and in its last workgroup the implicit Friedmann half-kick of ad
//      (src/evolution.cpp:153-167);
//   2. k_kick_drift: fd += d (laplnorm lapl f - friction fd - potnorm V'(f)) and
//      f_next = f + d fd (src/evolution.cpp:170-181, 265-270), f ping-ponged
//      between two buffers so neighbours read the old f;
//   3. (SIGW) k_tensor: the same for the six h_ij with source 2 d_i phi d_j phi
//      (src/evolution.cpp:185-253, 274-280);
// the last workgroup of the final launch does the second half-kick of ad, the
// drift of t and a, flips the buffer parity and evaluates the loop test a <= af.
// Every launch starts by reading DevBg::done, so steps queued past af are no-ops.
//
// Memory: fields are z-fastest (idx = (i N + j) N + k, src/main.h:34-38); a wave
// covers 64 consecutive k, so every access is a coalesced 256/512-byte row
// segment.  The grid is a fixed number of workgroups that stride over the
// lattice; workgroup b works on the contiguous slab of XCD b % 8, keeping the
// i +- 1 planes of the stencil inside one XCD's L2.
//
// Storage: FP64 (f, fd double) or split (float fluctuation about the FP64
// homogeneous mode phibar, pibar in DevBg).  Every right-hand side is evaluated
// in FP64; only the stored fluctuation is rounded to FP32.
#include <vector>

#include "ie_internal.h"

namespace ie {

namespace {

constexpr int kXcd = 8;

struct Walk {
    size_t begin, stride, end;
};

// Site range of workgroup b of the G working on one lattice: XCD-major
// contiguous slabs, strided inside.
__device__ __forceinline__ Walk site_walk(size_t gs, unsigned G, unsigned b)
{
    Walk w;
    if (G % kXcd == 0 && gs % ((size_t)kXcd * kBlock) == 0) {
        const unsigned xcd = b % kXcd, lb = b / kXcd, per = G / kXcd;
        const size_t slab = gs / kXcd;
        w.begin = (size_t)xcd * slab + (size_t)lb * kBlock + threadIdx.x;
        w.stride = (size_t)per * kBlock;
        w.end = (size_t)(xcd + 1) * slab;
    } else {
        w.begin = (size_t)b * kBlock + threadIdx.x;
        w.stride = (size_t)G * kBlock;
        w.end = gs;
    }
    return w;
}

__device__ __forceinline__ Walk site_walk(size_t gs) { return site_walk(gs, gridDim.x, blockIdx.x); }

// 32-bit form of the walk for the step kernels (N <= 1024, so N^3 < 2^31):
// integer index arithmetic stays in single 32-bit VALU instructions.
struct Walk32 {
    unsigned begin, stride, end;
};

__device__ __forceinline__ Walk32 site_walk32(unsigned gs, unsigned G, unsigned b)
{
    Walk32 w;
    if (G % kXcd == 0 && gs % (unsigned)(kXcd * kBlock) == 0) {
        const unsigned xcd = b % kXcd, lb = b / kXcd, per = G / kXcd;
        const unsigned slab = gs / kXcd;
        w.begin = xcd * slab + lb * kBlock + threadIdx.x;
        w.stride = per * kBlock;
        w.end = (xcd + 1) * slab;
    } else {
        w.begin = b * kBlock + threadIdx.x;
        w.stride = G * kBlock;
        w.end = gs;
    }
    return w;
}

struct Nb32 {
    unsigned ip, im, jp, jm, kp, km;
};

__device__ __forceinline__ Nb32 nb32(unsigned id, unsigned N, unsigned logN)
{
    const unsigned m = N - 1;
    const unsigned k = id & m, j = (id >> logN) & m, i = id >> (2 * logN);
    const unsigned base_k = id - k, base_j = id - (j << logN), base_i = id - (i << (2 * logN));
    Nb32 n;
    n.kp = base_k + ((k + 1) & m);
    n.km = base_k + ((k - 1) & m);
    n.jp = base_j + (((j + 1) & m) << logN);
    n.jm = base_j + (((j - 1) & m) << logN);
    n.ip = base_i + (((i + 1) & m) << (2 * logN));
    n.im = base_i + (((i - 1) & m) << (2 * logN));
    return n;
}

// One lattice of a batched launch (ie_batch_steps): its state, its workgroup
// count G and its first workgroup in the launch.
struct RunDesc {
    Fields F;
    Pot P;
    DevBg *bg;
    double *partials;
    unsigned G;
    int split;
};

struct Nb {
    size_t ip, im, jp, jm, kp, km;
};

__device__ __forceinline__ Nb neighbours(size_t id, int N, int logN)
{
    const size_t m = (size_t)N - 1;
    const size_t k = id & m, j = (id >> logN) & m, i = id >> (2 * logN);
    const size_t rowN = (size_t)N, planeN = (size_t)N * N;
    const size_t base_j = id - j * rowN, base_i = id - i * planeN;
    Nb n;
    n.kp = id - k + ((k + 1) & m);
    n.km = id - k + ((k - 1) & m);
    n.jp = base_j + ((j + 1) & m) * rowN;
    n.jm = base_j + ((j - 1) & m) * rowN;
    n.ip = base_i + ((i + 1) & m) * planeN;
    n.im = base_i + ((i - 1) & m) * planeN;
    return n;
}

// 7-point Laplacian in FP64 in the reference's summation order
// (src/evolution.cpp:34-51): k+, k-, j+, j-, i+, i-, then -6 f.
template <typename T>
__device__ __forceinline__ double lapl64(const T *__restrict__ g, size_t id, const Nb &n)
{
    return (double)g[n.kp] + (double)g[n.km] + (double)g[n.jp] + (double)g[n.jm] + (double)g[n.ip] +
           (double)g[n.im] - 6.0 * (double)g[id];
}

// FP32 Laplacian of a tensor component (lapl<float>, same order).
__device__ __forceinline__ float lapl32(const float *__restrict__ g, size_t id, const Nb &n)
{
    return g[n.kp] + g[n.km] + g[n.jp] + g[n.jm] + g[n.ip] + g[n.im] - 6.0f * g[id];
}

__device__ __forceinline__ void flag_error(DevBg *bg, int err)
{
    if (__any(err) && (threadIdx.x & 63) == 0) atomicOr(&bg->error, 1);
}

// Scalar updates that close a kick, a drift or a whole step (thread 0 of the
// last workgroup).  Mirrors src/evolution.cpp:256,262,283 and :942-944.
__device__ void finalize_scalars(DevBg *bg, int mode, int split)
{
    const double d = bg->d;
    if (mode & KICK) {
        bg->ad += 0.5 * d * bg->ad2;
        if (split) bg->pibar += d * (-bg->friction * bg->pibar - bg->potnorm * bg->vpmean);
    }
    if (mode & DRIFT) {
        bg->t += d;
        if (split) bg->phibar += d * bg->pibar;
        bg->a += d * bg->ad;
    }
    if (mode == KICKDRIFT) {  // only the device-resident loop uses the fused step
        bg->parity ^= 1;
        bg->steps += 1;
        bg->astep = bg->a;
        if (!(bg->a <= bg->af)) bg->done = 1;
    }
}

// ----------------------------------------------------------------- kernels

// The seven stencil values of U sites of a workgroup's walk, loaded before any
// arithmetic so U x 7 loads are in flight per thread.  Sites past the end of the
// walk load a valid site (the last one) and are masked when used.
template <typename T, int U>
struct Stencil {
    T c[U], kp[U], km[U], jp[U], jm[U], ip[U], im[U];
    unsigned id[U];
    bool ok[U];
};

template <typename T, int U>
__device__ __forceinline__ void load_stencil(const T *__restrict__ f, const Walk32 &w, unsigned base, unsigned N,
                                             unsigned logN, Stencil<T, U> &s)
{
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const unsigned raw = base + (unsigned)u * w.stride;
        s.ok[u] = raw < w.end;
        const unsigned id = s.ok[u] ? raw : w.end - 1;
        s.id[u] = id;
        const Nb32 n = nb32(id, N, logN);
        s.c[u] = f[id];
        s.kp[u] = f[n.kp];
        s.km[u] = f[n.km];
        s.jp[u] = f[n.jp];
        s.jm[u] = f[n.jm];
        s.ip[u] = f[n.ip];
        s.im[u] = f[n.im];
    }
}

// 7-point Laplacian of site u in FP64, summed in the reference's order
// (src/evolution.cpp:34-51): k+, k-, j+, j-, i+, i-, then -6 f.
template <typename T, int U>
__device__ __forceinline__ double stencil_lapl(const Stencil<T, U> &s, int u)
{
    return (double)s.kp[u] + (double)s.km[u] + (double)s.jp[u] + (double)s.jm[u] + (double)s.ip[u] +
           (double)s.im[u] - 6.0 * (double)s.c[u];
}

template <typename T>
struct Unroll {
    static constexpr int value = sizeof(T) == 4 ? 8 : 4;
};

template <typename T, bool SPLIT>
__device__ __forceinline__ void friedmann_reduce_body(const Fields &F, const Pot &P, DevBg *bg, double *partials,
                                                      double d_arg, int loop, unsigned G, unsigned blk, double *smem)
{
    constexpr int U = Unroll<T>::value;
    if (loop && bg->done) return;
    const unsigned N = (unsigned)F.N, logN = (unsigned)F.logN;
    const size_t gs = (size_t)F.N * F.N * F.N;
    const T *__restrict__ f = (const T *)F.f[bg->parity];
    const double phibar = SPLIT ? bg->phibar : 0.0;
    double acc[3] = {0.0, 0.0, 0.0};
    int err = 0;
    const Walk32 w = site_walk32((unsigned)gs, G, blk);
    for (unsigned base = w.begin; base < w.end; base += (unsigned)U * w.stride) {
        Stencil<T, U> s;
        load_stencil<T, U>(f, w, base, N, logN, s);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (!s.ok[u]) continue;
            const double c = (double)s.c[u];
            const double lap = stencil_lapl<T, U>(s, u);
            acc[0] -= c * lap;
            if (SPLIT) {
                double V, dV;
                pot_eval_fast(P, phibar + c, &V, &dV, &err);
                acc[1] += V;
                acc[2] += dV;
            } else {
                acc[1] += pot_V_fast(P, c, &err);
            }
        }
    }
    flag_error(bg, err);
    block_sum<3>(acc, smem);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) partials[(size_t)blk * 3 + q] = acc[q];
    if (!last_block(&bg->counter, G)) return;
    double tot[3];
    sum_partials<3>(partials, G, tot, smem);
    if (threadIdx.x != 0) return;
    bg->red[0] = tot[0];
    bg->red[1] = tot[1];
    bg->red[2] = tot[2];
    const double a = bg->a, ad = bg->ad, s = bg->rescale_s, dx = bg->dx;
    const double gsd = (double)gs;
    const double d = loop ? bg->dt * pow(bg->astep, s - 1.0) : d_arg;
    const double Ge = 0.5 * tot[0] * ((1.0 / (a * dx)) * (1.0 / (a * dx))) / gsd;
    const double Pm = tot[1] / gsd;
    const double sfev1 = s + 1.0, sfev2 = -2.0 * s + 2.0;
    bg->d = d;
    bg->ad2 = (-2.0 * ad - 2.0 * a / d / sfev1 *
               (1.0 - sqrt(1.0 + 2.0 * d * sfev1 * ad / a +
                           d * d * sfev1 * pow(a, sfev2) * (2.0 * Ge / 3.0 + Pm)))) / d;
    bg->ad = ad + 0.5 * d * bg->ad2;
    bg->friction = (2.0 + s) * bg->ad / a;
    bg->potnorm = pow(a, 2.0 - 2.0 * s);
    bg->laplnorm = pow(a, -2.0 * s) / (dx * dx);
    bg->srcamp = 2.0 * pow(a, -2.0 * s);
    bg->vpmean = tot[2] / gsd;
}

template <typename T, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_friedmann_reduce(Fields F, Pot P, DevBg *bg,
                                                              double *partials, double d_arg, int loop)
{
    __shared__ double smem[3 * kWaves];
    friedmann_reduce_body<T, SPLIT>(F, P, bg, partials, d_arg, loop, gridDim.x, blockIdx.x, smem);
}

template <typename T, bool SPLIT>
__global__ __launch_bounds__(kBlock) void kb_friedmann_reduce(const RunDesc *__restrict__ runs,
                                                               const uint2 *__restrict__ map)
{
    __shared__ double smem[3 * kWaves];
    const uint2 m = map[blockIdx.x];
    const RunDesc &R = runs[m.x];
    friedmann_reduce_body<T, SPLIT>(R.F, R.P, R.bg, R.partials, 0.0, 1, R.G, m.y, smem);
}

template <typename T, bool SPLIT, int MODE>
__device__ __forceinline__ void kick_drift_body(const Fields &F, const Pot &P, DevBg *bg, double d_arg, int loop,
                                                int finalize, unsigned G, unsigned blk)
{
    if (loop && bg->done) return;
    const size_t gs = (size_t)F.N * F.N * F.N;
    const int par = bg->parity;
    const T *__restrict__ f = (const T *)F.f[par];
    T *__restrict__ fout = (T *)F.f[MODE == KICKDRIFT ? par ^ 1 : par];
    T *__restrict__ fd = (T *)F.fd;
    // explicit drift takes d from the caller; kicks use the step set by the reduce
    const double d = (MODE == DRIFT && !loop) ? d_arg : bg->d;
    const double laplnorm = bg->laplnorm, friction = bg->friction, potnorm = bg->potnorm;
    const double phibar = SPLIT ? bg->phibar : 0.0, vpmean = SPLIT ? bg->vpmean : 0.0;
    int err = 0;
    if (MODE & KICK) {
        constexpr int U = Unroll<T>::value;
        const Walk32 w = site_walk32((unsigned)gs, G, blk);
        for (unsigned base = w.begin; base < w.end; base += (unsigned)U * w.stride) {
            Stencil<T, U> s;
            load_stencil<T, U>(f, w, base, (unsigned)F.N, (unsigned)F.logN, s);
            T vin[U];
#pragma unroll
            for (int u = 0; u < U; ++u) vin[u] = fd[s.id[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!s.ok[u]) continue;
                const double c = (double)s.c[u];
                const double lap = stencil_lapl<T, U>(s, u);
                double dV = pot_dV_fast(P, SPLIT ? phibar + c : c, &err);
                if (SPLIT) dV -= vpmean;
                double v = (double)vin[u];
                v = v + d * (laplnorm * lap - friction * v - potnorm * dV);
                const T vs = (T)v;
                fd[s.id[u]] = vs;
                if (MODE & DRIFT) fout[s.id[u]] = (T)(c + d * (double)vs);
            }
        }
    } else {
        // drift alone, in place (pointwise): through one unrestricted pointer
        T *g = (T *)F.f[par];
        const Walk w = site_walk(gs, G, blk);
        for (size_t id = w.begin; id < w.end; id += w.stride) g[id] = (T)((double)g[id] + d * (double)fd[id]);
    }
    flag_error(bg, err);
    if (!finalize) return;
    if (!last_block(&bg->counter, G)) return;
    if (threadIdx.x == 0) {
        if (MODE == DRIFT) bg->d = d;
        finalize_scalars(bg, MODE, SPLIT);
    }
}

template <typename T, bool SPLIT, int MODE>
__global__ __launch_bounds__(kBlock) void k_kick_drift(Fields F, Pot P, DevBg *bg, double d_arg,
                                                        int loop, int finalize)
{
    kick_drift_body<T, SPLIT, MODE>(F, P, bg, d_arg, loop, finalize, gridDim.x, blockIdx.x);
}

// batched loop step: the scalar kernel closes the step of runs without tensors
template <typename T, bool SPLIT>
__global__ __launch_bounds__(kBlock) void kb_kick_drift(const RunDesc *__restrict__ runs,
                                                         const uint2 *__restrict__ map)
{
    const uint2 m = map[blockIdx.x];
    const RunDesc &R = runs[m.x];
    kick_drift_body<T, SPLIT, KICKDRIFT>(R.F, R.P, R.bg, 0.0, 1, R.F.sigw ? 0 : 1, R.G, m.y);
}

template <typename T, int MODE>
__device__ __forceinline__ void tensor_body(const Fields &F, DevBg *bg, double d_arg, int loop, int split,
                                            unsigned G, unsigned blk)
{
    if (loop && bg->done) return;
    const size_t gs = (size_t)F.N * F.N * F.N;
    const int par = bg->parity;
    const T *__restrict__ f = (const T *)F.f[par];
    const double d = (MODE == DRIFT && !loop) ? d_arg : bg->d;
    const double laplnorm = bg->laplnorm, friction = bg->friction, srcamp = bg->srcamp;
    const double half_over_dx = 0.5 / bg->dx;
    const unsigned N = (unsigned)F.N, logN = (unsigned)F.logN;
    if (MODE == DRIFT) {
        // drift alone, in place (pointwise), through unrestricted pointers
        const Walk w = site_walk(gs, G, blk);
        for (int c = 0; c < 6; ++c) {
            float *h = F.h[par][c];
            const float *hd = F.hd[c];
            for (size_t id = w.begin; id < w.end; id += w.stride) h[id] = h[id] + (float)(d * (double)hd[id]);
        }
    } else {
        const Walk32 w = site_walk32((unsigned)gs, G, blk);
        for (unsigned id = w.begin; id < w.end; id += w.stride) {
            const Nb32 n = nb32(id, N, logN);
            // central gradient of phi (dfdx, src/evolution.cpp:55-77)
            const double gx = ((double)f[n.ip] - (double)f[n.im]) * half_over_dx;
            const double gy = ((double)f[n.jp] - (double)f[n.jm]) * half_over_dx;
            const double gz = ((double)f[n.kp] - (double)f[n.km]) * half_over_dx;
            const double src[6] = {gx * gx, gy * gy, gz * gz, gx * gy, gx * gz, gy * gz};
            float hc[6], hl[6], hv[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) {   // all loads of the site first
                const float *__restrict__ h = F.h[par][c];
                hc[c] = h[id];
                hl[c] = h[n.kp] + h[n.km] + h[n.jp] + h[n.jm] + h[n.ip] + h[n.im] - 6.0f * hc[c];
                hv[c] = F.hd[c][id];
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const double rhs = d * (laplnorm * (double)hl[c] - friction * (double)hv[c] + srcamp * src[c]);
                const float hn = hv[c] + (float)rhs;
                F.hd[c][id] = hn;
                if (MODE & DRIFT) F.h[par ^ 1][c][id] = hc[c] + (float)(d * (double)hn);
            }
        }
    }
    if (!last_block(&bg->counter, G)) return;
    if (threadIdx.x == 0) {
        if (MODE == DRIFT) bg->d = d;
        finalize_scalars(bg, MODE, split);
    }
}

template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_tensor(Fields F, DevBg *bg, double d_arg, int loop, int split)
{
    tensor_body<T, MODE>(F, bg, d_arg, loop, split, gridDim.x, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void kb_tensor(const RunDesc *__restrict__ runs, const uint2 *__restrict__ map)
{
    const uint2 m = map[blockIdx.x];
    const RunDesc &R = runs[m.x];
    tensor_body<T, KICKDRIFT>(R.F, R.bg, 0.0, 1, R.split, R.G, m.y);
}

__global__ void kb_loop_begin(const RunDesc *__restrict__ runs, int nruns)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nruns) {
        DevBg *bg = runs[r].bg;
        bg->done = !(bg->a <= bg->af) ? 1 : 0;
    }
}

// Volume sums for outputs and for the split upload.
//   red[0] sum f (split: sum dphi), red[1] sum fd (split: sum dpi),
//   red[2] sum f^2 (split: sum dphi^2), red[3] sum fd^2 (split: sum dpi^2),
//   red[4] sum -f lapl f (split: -dphi lapl dphi), red[5] sum V(f).
template <typename T, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_field_sums(Fields F, Pot P, DevBg *bg, double *partials)
{
    __shared__ double smem[6 * kWaves];
    const size_t gs = (size_t)F.N * F.N * F.N;
    const T *__restrict__ f = (const T *)F.f[bg->parity];
    const T *__restrict__ fd = (const T *)F.fd;
    const double phibar = SPLIT ? bg->phibar : 0.0;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    int err = 0;
    const Walk w = site_walk(gs);
    for (size_t id = w.begin; id < w.end; id += w.stride) {
        const Nb n = neighbours(id, F.N, F.logN);
        const double c = (double)f[id], v = (double)fd[id];
        acc[0] += c;
        acc[1] += v;
        acc[2] += c * c;
        acc[3] += v * v;
        acc[4] -= c * lapl64(f, id, n);
        double V, dV;
        pot_eval(P, SPLIT ? phibar + c : c, &V, &dV, &err);
        acc[5] += V;
    }
    flag_error(bg, err);
    block_sum<6>(acc, smem);
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q) partials[(size_t)blockIdx.x * 6 + q] = acc[q];
    if (!last_block(&bg->counter, gridDim.x)) return;
    double tot[6];
    sum_partials<6>(partials, gridDim.x, tot, smem);
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q) bg->red[q] = tot[q];
}

// Split upload, pass 1: FP64 means of f and fd (the homogeneous mode).
__global__ __launch_bounds__(kBlock) void k_mean2(const double *f, const double *fd, size_t gs, DevBg *bg,
                                                   double *partials)
{
    __shared__ double smem[2 * kWaves];
    double acc[2] = {0.0, 0.0};
    const Walk w = site_walk(gs);
    for (size_t id = w.begin; id < w.end; id += w.stride) {
        acc[0] += f[id];
        acc[1] += fd[id];
    }
    block_sum<2>(acc, smem);
    if (threadIdx.x == 0) {
        partials[(size_t)blockIdx.x * 2] = acc[0];
        partials[(size_t)blockIdx.x * 2 + 1] = acc[1];
    }
    if (!last_block(&bg->counter, gridDim.x)) return;
    double tot[2];
    sum_partials<2>(partials, gridDim.x, tot, smem);
    if (threadIdx.x == 0) {
        bg->phibar = tot[0] / (double)gs;
        bg->pibar = tot[1] / (double)gs;
    }
}

// Split upload, pass 2: store the fluctuations in FP32.
__global__ __launch_bounds__(kBlock) void k_split_store(const double *f, const double *fd, size_t gs,
                                                        const DevBg *bg, float *df, float *dfd)
{
    const double pb = bg->phibar, qb = bg->pibar;
    const Walk w = site_walk(gs);
    for (size_t id = w.begin; id < w.end; id += w.stride) {
        df[id] = (float)(f[id] - pb);
        dfd[id] = (float)(fd[id] - qb);
    }
}

template <typename T, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_reconstitute(Fields F, const DevBg *bg, int which, double *out)
{
    const size_t gs = (size_t)F.N * F.N * F.N;
    const T *__restrict__ src = which == 0 ? (const T *)F.f[bg->parity] : (const T *)F.fd;
    const double base = SPLIT ? (which == 0 ? bg->phibar : bg->pibar) : 0.0;
    const Walk w = site_walk(gs);
    for (size_t id = w.begin; id < w.end; id += w.stride) out[id] = SPLIT ? base + (double)src[id] : (double)src[id];
}

}  // namespace

int reduce_blocks(size_t gs)
{
    size_t b = gs / kBlock;
    if (b > 2048) b = 2048;
    if (b >= kXcd) b -= b % kXcd;
    return (int)(b < 1 ? 1 : b);
}

hipError_t launch_friedmann_reduce(const Fields &F, int split, const Pot &P, DevBg *bg, double *partials,
                                   double d_arg, int loop, hipStream_t s)
{
    const size_t gs = (size_t)F.N * F.N * F.N;
    const int G = reduce_blocks(gs);
    if (split) k_friedmann_reduce<float, true><<<G, kBlock, 0, s>>>(F, P, bg, partials, d_arg, loop);
    else k_friedmann_reduce<double, false><<<G, kBlock, 0, s>>>(F, P, bg, partials, d_arg, loop);
    return hipGetLastError();
}

template <typename T, bool SPLIT>
static void kd_dispatch(const Fields &F, const Pot &P, DevBg *bg, int mode, double d_arg, int loop, int fin,
                        int G, hipStream_t s)
{
    if (mode == KICK) k_kick_drift<T, SPLIT, KICK><<<G, kBlock, 0, s>>>(F, P, bg, d_arg, loop, fin);
    else if (mode == DRIFT) k_kick_drift<T, SPLIT, DRIFT><<<G, kBlock, 0, s>>>(F, P, bg, d_arg, loop, fin);
    else k_kick_drift<T, SPLIT, KICKDRIFT><<<G, kBlock, 0, s>>>(F, P, bg, d_arg, loop, fin);
}

template <typename T>
static void tensor_dispatch(const Fields &F, DevBg *bg, int mode, double d_arg, int loop, int split, int G,
                            hipStream_t s)
{
    if (mode == KICK) k_tensor<T, KICK><<<G, kBlock, 0, s>>>(F, bg, d_arg, loop, split);
    else if (mode == DRIFT) k_tensor<T, DRIFT><<<G, kBlock, 0, s>>>(F, bg, d_arg, loop, split);
    else k_tensor<T, KICKDRIFT><<<G, kBlock, 0, s>>>(F, bg, d_arg, loop, split);
}

hipError_t launch_kick_drift(const Fields &F, int split, const Pot &P, DevBg *bg, int mode, double d_arg,
                             int loop, hipStream_t s, int part)
{
    const size_t gs = (size_t)F.N * F.N * F.N;
    const int G = reduce_blocks(gs);
    const int fin_scalar = F.sigw ? 0 : 1;
    if (part != 1) {
        if (split) kd_dispatch<float, true>(F, P, bg, mode, d_arg, loop, fin_scalar, G, s);
        else kd_dispatch<double, false>(F, P, bg, mode, d_arg, loop, fin_scalar, G, s);
    }
    if (F.sigw && part != 0) {
        if (split) tensor_dispatch<float>(F, bg, mode, d_arg, loop, split, G, s);
        else tensor_dispatch<double>(F, bg, mode, d_arg, loop, split, G, s);
    }
    return hipGetLastError();
}

hipError_t launch_field_sums(const Fields &F, int split, const Pot &P, DevBg *bg, double *partials, hipStream_t s)
{
    const size_t gs = (size_t)F.N * F.N * F.N;
    const int G = reduce_blocks(gs);
    if (split) k_field_sums<float, true><<<G, kBlock, 0, s>>>(F, P, bg, partials);
    else k_field_sums<double, false><<<G, kBlock, 0, s>>>(F, P, bg, partials);
    return hipGetLastError();
}

hipError_t launch_split_upload(const double *f, const double *fd, Fields &F, DevBg *bg, double *partials,
                               hipStream_t s)
{
    const size_t gs = (size_t)F.N * F.N * F.N;
    const int G = reduce_blocks(gs);
    k_mean2<<<G, kBlock, 0, s>>>(f, fd, gs, bg, partials);
    k_split_store<<<G, kBlock, 0, s>>>(f, fd, gs, bg, (float *)F.f[0], (float *)F.fd);
    return hipGetLastError();
}

hipError_t launch_reconstitute(const Fields &F, int split, int which, const DevBg *bg, double *out, hipStream_t s)
{
    const size_t gs = (size_t)F.N * F.N * F.N;
    const int G = reduce_blocks(gs);
    if (split) k_reconstitute<float, true><<<G, kBlock, 0, s>>>(F, bg, which, out);
    else k_reconstitute<double, false><<<G, kBlock, 0, s>>>(F, bg, which, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ batches

struct BatchDev {
    RunDesc *runs = nullptr;
    uint2 *map = nullptr, *tmap = nullptr;
    int nruns = 0, nblocks = 0, ntblocks = 0, split = 0;
};

// Workgroups per lattice in a batched launch: a multiple of 8 (whole XCD slabs),
// at most 256 so the last workgroup's fixed-order sum stays short.
static unsigned batch_blocks(size_t gs)
{
    size_t b = gs / kBlock;
    if (b > 256) b = 256;
    b = (b + kXcd - 1) / kXcd * kXcd;
    return (unsigned)(b < kXcd ? kXcd : b);
}

hipError_t batch_build(const BatchRun *runs, int n, int split, BatchDev **out)
{
    *out = nullptr;
    std::vector<RunDesc> d(n);
    std::vector<uint2> map, tmap;
    for (int r = 0; r < n; ++r) {
        const size_t gs = (size_t)runs[r].F.N * runs[r].F.N * runs[r].F.N;
        d[r].F = runs[r].F;
        d[r].P = runs[r].P;
        d[r].bg = runs[r].bg;
        d[r].partials = runs[r].partials;
        d[r].G = batch_blocks(gs);
        d[r].split = split;
        for (unsigned b = 0; b < d[r].G; ++b) {
            map.push_back(make_uint2((unsigned)r, b));
            if (runs[r].F.sigw) tmap.push_back(make_uint2((unsigned)r, b));
        }
    }
    BatchDev *B = new BatchDev();
    B->nruns = n;
    B->nblocks = (int)map.size();
    B->ntblocks = (int)tmap.size();
    B->split = split;
    hipError_t e = hipMalloc(&B->runs, sizeof(RunDesc) * n);
    if (e == hipSuccess) e = hipMalloc(&B->map, sizeof(uint2) * map.size());
    if (e == hipSuccess && !tmap.empty()) e = hipMalloc(&B->tmap, sizeof(uint2) * tmap.size());
    if (e == hipSuccess) e = hipMemcpy(B->runs, d.data(), sizeof(RunDesc) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B->map, map.data(), sizeof(uint2) * map.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !tmap.empty())
        e = hipMemcpy(B->tmap, tmap.data(), sizeof(uint2) * tmap.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        batch_free(B);
        return e;
    }
    *out = B;
    return hipSuccess;
}

void batch_free(BatchDev *B)
{
    if (!B) return;
    if (B->runs) (void)hipFree(B->runs);
    if (B->map) (void)hipFree(B->map);
    if (B->tmap) (void)hipFree(B->tmap);
    delete B;
}

hipError_t batch_loop_begin(const BatchDev *B, hipStream_t s)
{
    kb_loop_begin<<<(B->nruns + 63) / 64, 64, 0, s>>>(B->runs, B->nruns);
    return hipGetLastError();
}

hipError_t batch_step(const BatchDev *B, hipStream_t s)
{
    if (B->split) {
        kb_friedmann_reduce<float, true><<<B->nblocks, kBlock, 0, s>>>(B->runs, B->map);
        kb_kick_drift<float, true><<<B->nblocks, kBlock, 0, s>>>(B->runs, B->map);
        if (B->ntblocks) kb_tensor<float><<<B->ntblocks, kBlock, 0, s>>>(B->runs, B->tmap);
    } else {
        kb_friedmann_reduce<double, false><<<B->nblocks, kBlock, 0, s>>>(B->runs, B->map);
        kb_kick_drift<double, false><<<B->nblocks, kBlock, 0, s>>>(B->runs, B->map);
        if (B->ntblocks) kb_tensor<double><<<B->ntblocks, kBlock, 0, s>>>(B->runs, B->tmap);
    }
    return hipGetLastError();
}

}  // namespace ie
