// ie_lattice.hip -- leapfrog step of the inflaton lattice on gfx950.
//
// One step of run_evolution_loop (src/evolution.cpp:937-945) is a short chain of
// launches on the run's stream, with no host involvement:
//   1. reduce: per-workgroup partial sums of -f lapl f and V(f) (gradient_energy
//      and potential_energy, src/evolution.cpp:130-136, src/potential.cpp:151-171);
//   2. friedmann_finish (one workgroup): the fixed-order total and the implicit
//      Friedmann half-kick of ad (src/evolution.cpp:153-167), the step's
//      coefficients into DevBg;
//   3. scalar step: fd += d (laplnorm lapl f - friction fd - potnorm V'(f)) and
//      f_next = f + d fd (src/evolution.cpp:170-181, 265-270), f ping-ponged
//      between two buffers so neighbours read the old f;
//   4. (SIGW) tensor step: the same for the six h_ij with source 2 d_i phi d_j phi
//      (src/evolution.cpp:185-253, 274-280);
//   5. step_finish (one thread): second half-kick of ad, drift of t and a, buffer
//      parity flip, loop test a <= af (src/evolution.cpp:256-283, 942-944).
// Every launch starts by reading DevBg::done, so steps queued past af are no-ops.
// Cross-workgroup results travel only through kernel boundaries: the streaming
// kernels end with no device-scope fence (on a multi-XCD part a release fence
// writes back the workgroup's L2, which stalls a streaming kernel).
//
// Memory: fields are z-fastest (idx = (i N + j) N + k, src/main.h:34-38).
// For N >= 64 the step kernels work on rows: a wave owns R consecutive j rows of
// one (i, k-segment) column, each lane holds E = min(N/64, 4) consecutive k of a
// 64 E-site segment (one vector load per row), the k +- 1 neighbours come from the
// lane's own registers and two lane shuffles, the j +- 1 rows from a sliding
// window in registers, and the i +- 1 rows from the planes that neighbouring
// workgroups of the same XCD are streaming (L2 hits).  Row and plane offsets are
// wave-uniform (scalar registers).  Workgroup b works in the contiguous slab of
// planes of XCD b % 8.  N < 64 uses a per-site walk with the same arithmetic.
//
// Storage: FP64 (f, fd double) or split (float fluctuation about the FP64
// homogeneous mode phibar, pibar in DevBg).  Every right-hand side is evaluated
// in FP64; only the stored fluctuation is rounded to FP32.
#include <vector>

#include <type_traits>

#include "ie_internal.h"

namespace ie {

namespace {

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

// 32-bit form of the walk for the small-lattice step kernels.
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

// One lattice of a batched launch (ie_batch_steps): its state and its workgroup
// count G (the same as for a launch of the lattice alone).
struct RunDesc {
    Fields F;
    Pot P;
    DevBg *bg;
    double *partials;
    unsigned G;        // workgroups of the row / site kernels
    unsigned Gm;       // workgroups of the fused march kernel (0: not fused)
    int split;
    int march;         // loop steps fused (march kernel + close)
    int res;           // batched 64^3 lattice stepped by one resident workgroup (kb_resident)
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

__device__ __forceinline__ double lapl_sites(double kp, double km, double jp, double jm, double ip, double im,
                                             double c)
{
    return kp + km + jp + jm + ip + im - 6.0 * c;
}

__device__ __forceinline__ void flag_error(DevBg *bg, int err)
{
    if (__any(err) && (threadIdx.x & 63) == 0) atomicOr(&bg->error, 1);
}

// ------------------------------------------------------------ finish kernels

// Scalar updates that close a kick, a drift or a whole step.  Mirrors
// src/evolution.cpp:256,262,283 and :942-944.
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

// std::pow of the scale factor in the step's coefficients.  The exponents of the
// default rescale_s = 0 (-1, 0, 2) are evaluated exactly as a correctly rounded
// pow gives them (1/x, 1, x*x: one IEEE operation each), skipping the FP64 pow
// routine in the one-thread chain between two steps.
__device__ __forceinline__ double scale_pow(double x, double y)
{
    if (y == 0.0) return 1.0;
    if (y == 2.0) return x * x;
    if (y == -1.0) return 1.0 / x;
    return pow(x, y);
}

// Fixed-order total of the G partials and the implicit Friedmann half-kick
// (evolve_derivs, src/evolution.cpp:153-167).  One workgroup.
__device__ void friedmann_finish_body(DevBg *bg, const double *partials, unsigned G, size_t gs, double d_arg,
                                      int loop, double *smem)
{
    if (loop && bg->done) return;
    double tot[3];
    sum_partials<3>(partials, (int)G, tot, smem);
    if (threadIdx.x != 0) return;
    bg->red[0] = tot[0];
    bg->red[1] = tot[1];
    bg->red[2] = tot[2];
    const double a = bg->a, ad = bg->ad, s = bg->rescale_s, dx = bg->dx;
    const double gsd = (double)gs;
    const double d = loop ? bg->dt * scale_pow(bg->astep, s - 1.0) : d_arg;
    const double Ge = 0.5 * tot[0] * ((1.0 / (a * dx)) * (1.0 / (a * dx))) / gsd;
    const double Pm = tot[1] / gsd;
    const double sfev1 = s + 1.0, sfev2 = -2.0 * s + 2.0;
    bg->d = d;
    bg->ad2 = (-2.0 * ad - 2.0 * a / d / sfev1 *
               (1.0 - sqrt(1.0 + 2.0 * d * sfev1 * ad / a +
                           d * d * sfev1 * scale_pow(a, sfev2) * (2.0 * Ge / 3.0 + Pm)))) / d;
    bg->ad = ad + 0.5 * d * bg->ad2;
    bg->friction = (2.0 + s) * bg->ad / a;
    bg->potnorm = scale_pow(a, 2.0 - 2.0 * s);
    bg->laplnorm = scale_pow(a, -2.0 * s) / (dx * dx);
    bg->srcamp = 2.0 * scale_pow(a, -2.0 * s);
    bg->vpmean = tot[2] / gsd;
}

// The StepScalars of a fused loop step, evaluated once per step (no contraction)
// by the one-workgroup kernel before it and stored in bg->nx, so the step's
// kernels start loading at once instead of evaluating four FP64 pows each.
#pragma clang fp contract(off)
__device__ __forceinline__ StepScalars step_scalars(const DevBg *bg, double gsd)
{
    StepScalars S;
    const double a = bg->a, ad = bg->ad, s = bg->rescale_s, dx = bg->dx;
    const double d = bg->dt * scale_pow(bg->astep, s - 1.0);
    const double Ge = 0.5 * bg->red[0] * ((1.0 / (a * dx)) * (1.0 / (a * dx))) / gsd;
    const double Pm = bg->red[1] / gsd;
    const double sfev1 = s + 1.0, sfev2 = -2.0 * s + 2.0;
    S.d = d;
    S.ad2 = (-2.0 * ad - 2.0 * a / d / sfev1 *
             (1.0 - sqrt(1.0 + 2.0 * d * sfev1 * ad / a +
                         d * d * sfev1 * scale_pow(a, sfev2) * (2.0 * Ge / 3.0 + Pm)))) /
            d;
    S.ad_half = ad + 0.5 * d * S.ad2;
    S.ad_next = S.ad_half + 0.5 * d * S.ad2;
    S.friction = (2.0 + s) * S.ad_half / a;
    S.potnorm = scale_pow(a, 2.0 - 2.0 * s);
    S.laplnorm = scale_pow(a, -2.0 * s) / (dx * dx);
    S.srcamp = 2.0 * scale_pow(a, -2.0 * s);
    S.vpmean = bg->red[2] / gsd;
    S.pibar_next = bg->pibar + d * (-S.friction * bg->pibar - S.potnorm * S.vpmean);
    S.phibar_next = bg->phibar + d * S.pibar_next;
    return S;
}
#pragma clang fp contract(fast)

// The fused loop's sums of a hilltop lattice are those of f and f^2 (f the
// stored value: the fluctuation about phibar in split storage); with
// V = hV - hV4 (phibar + f)^2 and V' = -hV2 (phibar + f) they give sum V and
// sum V' over the gsd sites.
__device__ __forceinline__ void hill_sums(const DevBg &b, double gsd, double (&tot)[3])
{
    if (!b.hill_sums) return;
    const double pb = b.hill_sums == 2 ? b.phibar : 0.0;
    const double s1 = tot[1], s2 = tot[2];
    tot[1] = gsd * (b.hV - b.hV4 * (pb * pb)) - b.hV4 * (2.0 * pb * s1 + s2);
    tot[2] = -b.hV2 * (gsd * pb + s1);
}

// Closes a fused loop step on a background record: applies the step's scalar
// updates (b.nx, the values its kernels used), stores the fixed-order sums of the
// new f for the next step and evaluates that step's scalars.
__device__ void close_state(DevBg &b, const double (&tot_in)[3], double gsd, int split)
{
    double tot[3] = {tot_in[0], tot_in[1], tot_in[2]};
    const StepScalars S = b.nx;
    b.d = S.d;
    b.ad2 = S.ad2;
    b.friction = S.friction;
    b.potnorm = S.potnorm;
    b.laplnorm = S.laplnorm;
    b.srcamp = S.srcamp;
    b.vpmean = S.vpmean;
    b.ad = S.ad_next;
    b.t += S.d;
    if (split) {
        b.pibar = S.pibar_next;
        b.phibar = S.phibar_next;
    }
    b.a += S.d * b.ad;
    b.parity ^= 1;
    b.steps += 1;
    b.astep = b.a;
    if (!(b.a <= b.af)) b.done = 1;
    hill_sums(b, gsd, tot);   // of the new f, about the new phibar
    b.red[0] = tot[0];
    b.red[1] = tot[1];
    b.red[2] = tot[2];
    b.nx = step_scalars(&b, gsd);
}

// A fused loop keeps the background of its current step in two records
// (Fields::rec).  Step n of a loop (counted by the host from the loop's start)
// finds the previous step's record in slot n & 1: its kernel closes it with the
// partial sums that step wrote to partials slot n & 1 (every workgroup alike, the
// same bits), marches with the result, writes that record to slot (n + 1) & 1
// (workgroup 0) and its own partials to slot (n + 1) & 1.  A step whose record
// says the loop has ended writes that record on to slot (n + 1) & 1 unchanged, so
// both slots say so and no tensor launch after the end finds an older record.
// No launch writes what it reads; bg itself is brought up to date once, after
// the loop.
// Word-wise copy of a background record by the lanes of one wave.
static_assert(sizeof(DevBg) % 8 == 0 && sizeof(DevBg) / 8 <= 64, "DevBg is copied by one wave");
__device__ __forceinline__ void copy_bg_wave(DevBg *dst, const DevBg *src, unsigned lane)
{
    if (lane < sizeof(DevBg) / 8)
        reinterpret_cast<unsigned long long *>(dst)[lane] = reinterpret_cast<const unsigned long long *>(src)[lane];
}

// After a fused loop: closes its last step (if one ran and the loop had not
// ended) into bg, keeping the error flag that the step kernels raised in bg.
// One workgroup.
__device__ void loop_end_body(DevBg *bg, const DevBg *rec, unsigned sp, const double *partials, unsigned G, size_t gs,
                              int split, double *smem, DevBg &sb)
{
    const DevBg *prev = rec + sp;
    const bool close = prev->marched && !prev->done;
    double tot[3] = {0.0, 0.0, 0.0};
    if (close) sum_partials<3>(partials + (size_t)sp * G * 3, (int)G, tot, smem);
    if (threadIdx.x < 64) copy_bg_wave(&sb, prev, threadIdx.x);
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (close) close_state(sb, tot, (double)gs, split);
    const int err = bg->error;
    const unsigned cnt = bg->counter;
    *bg = sb;
    bg->error = err;
    bg->counter = cnt;
}

// Sums of the current f into bg->red at the start of a fused loop (from a reduce),
// the scalars of the loop's first step, and the loop's first record.
__device__ void sums_finish_body(DevBg *bg, const double *partials, unsigned G, size_t gs, DevBg *rec, double *smem)
{
    if (!bg->done) {
        double tot[3];
        sum_partials<3>(partials, (int)G, tot, smem);
        if (threadIdx.x == 0) {
            hill_sums(*bg, (double)gs, tot);
            bg->red[0] = tot[0];
            bg->red[1] = tot[1];
            bg->red[2] = tot[2];
            bg->nx = step_scalars(bg, (double)gs);
        }
    }
    if (threadIdx.x != 0) return;
    rec[0] = *bg;
    rec[0].marched = 0;
}

__global__ __launch_bounds__(kBlock) void k_friedmann_finish(DevBg *bg, const double *partials, unsigned G,
                                                              unsigned long long gs, double d_arg, int loop)
{
    __shared__ double smem[3 * kWaves];
    friedmann_finish_body(bg, partials, G, (size_t)gs, d_arg, loop, smem);
}

__global__ __launch_bounds__(kBlock) void kb_friedmann_finish(const RunDesc *__restrict__ runs)
{
    __shared__ double smem[3 * kWaves];
    const RunDesc &R = runs[blockIdx.x];
    if (R.march || R.res) return;
    const size_t gs = (size_t)R.F.N * R.F.N * R.F.N;
    friedmann_finish_body(R.bg, R.partials, R.G, gs, 0.0, 1, smem);
}

__global__ void k_step_finish(DevBg *bg, int mode, int split, double d_arg, int loop)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (loop && bg->done) return;
    if (mode == DRIFT && !loop) bg->d = d_arg;
    finalize_scalars(bg, mode, split);
}

__global__ void kb_step_finish(const RunDesc *__restrict__ runs, int nruns)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns || runs[r].march || runs[r].res) return;
    DevBg *bg = runs[r].bg;
    if (bg->done) return;
    finalize_scalars(bg, KICKDRIFT, runs[r].split);
}

__global__ void kb_loop_begin(const RunDesc *__restrict__ runs, int nruns)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nruns) {
        DevBg *bg = runs[r].bg;
        bg->done = !(bg->a <= bg->af) ? 1 : 0;
    }
}

// ------------------------------------------------------------ row kernels

template <typename T, bool SPLIT, int E>
__device__ __forceinline__ void row_reduce_body(const Fields &F, const Pot &P0, DevBg *bg, double *partials, int loop,
                                                unsigned G, unsigned blk, double *smem)
{
    if (loop && bg->done) return;
    const PotView Pv = pot_view(P0);
    const unsigned N = (unsigned)F.N, logN = (unsigned)F.logN, m = N - 1, lane = threadIdx.x & 63;
    const T *__restrict__ f = (const T *)F.f[bg->parity];
    const double phibar = SPLIT ? bg->phibar : 0.0;
    const RowWalk w = row_walk(N, G, blk, E);
    const T *plane = f + (w.i << (2 * logN)) + w.k0;
    const T *plane_p = f + (((w.i + 1) & m) << (2 * logN)) + w.k0;
    const T *plane_m = f + (((w.i - 1) & m) << (2 * logN)) + w.k0;
    T prev[E], cur[E], nxt[E], up[E], dn[E];
    ld_row<T, E>(plane + (((w.j0 - 1) & m) << logN), lane, prev);
    ld_row<T, E>(plane + (w.j0 << logN), lane, cur);
    double acc[3] = {0.0, 0.0, 0.0};
    int err = 0, hint[E];
#pragma unroll
    for (int e = 0; e < E; ++e) hint[e] = Pv.start(phibar + (double)cur[e]);
    for (unsigned r = 0; r < w.R; ++r) {
        const unsigned j = w.j0 + r;
        ld_row<T, E>(plane + (((j + 1) & m) << logN), lane, nxt);
        ld_row<T, E>(plane_p + (j << logN), lane, up);
        ld_row<T, E>(plane_m + (j << logN), lane, dn);
        T kp_last, km_first;
        k_neighbours<T, E>(cur, plane + (j << logN) - w.k0, w, N, lane, kp_last, km_first);
        double tv[E], tV[E], tdV[E];
#pragma unroll
        for (int e = 0; e < E; ++e) tv[e] = phibar + (double)cur[e];
        Pv.eval<E>(tv, hint, tV, tdV, &err);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const double c = (double)cur[e];
            const double kp = (double)(e < E - 1 ? cur[e + 1 < E ? e + 1 : e] : kp_last);
            const double km = (double)(e > 0 ? cur[e > 0 ? e - 1 : 0] : km_first);
            const double lap = lapl_sites(kp, km, (double)nxt[e], (double)prev[e], (double)up[e], (double)dn[e], c);
            acc[0] -= c * lap;
            acc[1] += tV[e];
            if (SPLIT) acc[2] += tdV[e];
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            prev[e] = cur[e];
            cur[e] = nxt[e];
        }
    }
    flag_error(bg, err);
    block_sum<3>(acc, smem);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) partials[(size_t)blk * 3 + q] = acc[q];
}

template <typename T, bool SPLIT, int E, int MODE>
__device__ __forceinline__ void row_kick_body(const Fields &F, const Pot &P0, DevBg *bg, int loop, unsigned G,
                                              unsigned blk)
{
    if (loop && bg->done) return;
    const PotView Pv = pot_view(P0);
    const unsigned N = (unsigned)F.N, logN = (unsigned)F.logN, m = N - 1, lane = threadIdx.x & 63;
    const int par = bg->parity;
    const T *__restrict__ f = (const T *)F.f[par];
    T *__restrict__ fout = (T *)F.f[par ^ 1];
    const T *fd = (const T *)F.fd[par];
    T *fdout = (T *)F.fd[MODE == KICKDRIFT ? par ^ 1 : par];
    const double d = bg->d;
    const double laplnorm = bg->laplnorm, friction = bg->friction, potnorm = bg->potnorm;
    const double phibar = SPLIT ? bg->phibar : 0.0, vpmean = SPLIT ? bg->vpmean : 0.0;
    const RowWalk w = row_walk(N, G, blk, E);
    const unsigned poff = (w.i << (2 * logN)) + w.k0;
    const T *plane = f + poff;
    const T *plane_p = f + (((w.i + 1) & m) << (2 * logN)) + w.k0;
    const T *plane_m = f + (((w.i - 1) & m) << (2 * logN)) + w.k0;
    T prev[E], cur[E], nxt[E], up[E], dn[E], vin[E];
    ld_row<T, E>(plane + (((w.j0 - 1) & m) << logN), lane, prev);
    ld_row<T, E>(plane + (w.j0 << logN), lane, cur);
    int err = 0, hint[E];
#pragma unroll
    for (int e = 0; e < E; ++e) hint[e] = Pv.start((SPLIT ? phibar : 0.0) + (double)cur[e]);
    for (unsigned r = 0; r < w.R; ++r) {
        const unsigned j = w.j0 + r;
        const unsigned roff = poff + (j << logN);
        ld_row<T, E>(plane + (((j + 1) & m) << logN), lane, nxt);
        ld_row<T, E>(plane_p + (j << logN), lane, up);
        ld_row<T, E>(plane_m + (j << logN), lane, dn);
        ld_row<T, E>(fd + roff, lane, vin);
        T kp_last, km_first;
        k_neighbours<T, E>(cur, f + roff - w.k0, w, N, lane, kp_last, km_first);
        double tv[E], tdV[E];
#pragma unroll
        for (int e = 0; e < E; ++e) tv[e] = (SPLIT ? phibar : 0.0) + (double)cur[e];
        Pv.dV<E>(tv, hint, tdV, &err);
        T vout[E], fnew[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const double c = (double)cur[e];
            const double kp = (double)(e < E - 1 ? cur[e + 1 < E ? e + 1 : e] : kp_last);
            const double km = (double)(e > 0 ? cur[e > 0 ? e - 1 : 0] : km_first);
            const double lap = lapl_sites(kp, km, (double)nxt[e], (double)prev[e], (double)up[e], (double)dn[e], c);
            double dV = tdV[e];
            if (SPLIT) dV -= vpmean;
            double v = (double)vin[e];
            v = v + d * (laplnorm * lap - friction * v - potnorm * dV);
            vout[e] = (T)v;
            fnew[e] = (T)(c + d * (double)vout[e]);
        }
        st_row<T, E>(fdout + roff, lane, vout);
        if (MODE & DRIFT) st_row<T, E>(fout + roff, lane, fnew);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            prev[e] = cur[e];
            cur[e] = nxt[e];
        }
    }
    flag_error(bg, err);
}

// Tensor kick (+ drift) of the six h_ij on rows (src/evolution.cpp:185-253,
// 274-280): source 2 d_i phi d_j phi from the central gradient of the old f (a
// j-window of f rows in registers), the FP32 Laplacian of each component from its
// five rows and two lane shuffles, the right-hand side in FP64.
template <typename T, int E, int MODE>
__device__ __forceinline__ void row_tensor_body(const Fields &F, DevBg *bg, int loop, int fresh, unsigned G,
                                                unsigned blk)
{
    // fused loop steps: the step's record (its scalars, as the march kernel's)
    if (fresh) bg = F.rec + (fresh - 1);
    if (loop && bg->done) return;
    const unsigned N = (unsigned)F.N, logN = (unsigned)F.logN, m = N - 1, lane = threadIdx.x & 63;
    const int par = bg->parity;
    const T *__restrict__ f = (const T *)F.f[par];
    double d = bg->d, laplnorm = bg->laplnorm, friction = bg->friction, srcamp = bg->srcamp;
    if (fresh) {
        const StepScalars S = bg->nx;
        d = S.d;
        laplnorm = S.laplnorm;
        friction = S.friction;
        srcamp = S.srcamp;
    }
    const double hx = 0.5 / bg->dx;
    const RowWalk w = row_walk(N, G, blk, E);
    const unsigned poff = (w.i << (2 * logN)) + w.k0;
    const unsigned pp = (((w.i + 1) & m) << (2 * logN)) + w.k0, pm = (((w.i - 1) & m) << (2 * logN)) + w.k0;
    T prev[E], cur[E], nxt[E];
    ld_row<T, E>(f + poff + (((w.j0 - 1) & m) << logN), lane, prev);
    ld_row<T, E>(f + poff + (w.j0 << logN), lane, cur);
    for (unsigned r = 0; r < w.R; ++r) {
        const unsigned j = w.j0 + r;
        const unsigned rj = j << logN, rjp = ((j + 1) & m) << logN, rjm = ((j - 1) & m) << logN;
        T up[E], dn[E];
        ld_row<T, E>(f + poff + rjp, lane, nxt);
        ld_row<T, E>(f + pp + rj, lane, up);
        ld_row<T, E>(f + pm + rj, lane, dn);
        T kp_last, km_first;
        k_neighbours<T, E>(cur, f + poff + rj - w.k0, w, N, lane, kp_last, km_first);
        double gx[E], gy[E], gz[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const double kp = (double)(e < E - 1 ? cur[e + 1 < E ? e + 1 : e] : kp_last);
            const double km = (double)(e > 0 ? cur[e > 0 ? e - 1 : 0] : km_first);
            gx[e] = ((double)up[e] - (double)dn[e]) * hx;
            gy[e] = ((double)nxt[e] - (double)prev[e]) * hx;
            gz[e] = (kp - km) * hx;
        }
#pragma unroll 1
        for (int c = 0; c < 6; ++c) {
            const float *__restrict__ h = F.h[par][c];
            float hc[E], hjp[E], hjm[E], hip[E], him[E], hv[E];
            ld_row<float, E>(h + poff + rj, lane, hc);
            ld_row<float, E>(h + poff + rjp, lane, hjp);
            ld_row<float, E>(h + poff + rjm, lane, hjm);
            ld_row<float, E>(h + pp + rj, lane, hip);
            ld_row<float, E>(h + pm + rj, lane, him);
            ld_row<float, E>(F.hd[c] + poff + rj, lane, hv);
            float hkp, hkm;
            k_neighbours<float, E>(hc, h + poff + rj - w.k0, w, N, lane, hkp, hkm);
            float hdn[E], hn[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const float kp = e < E - 1 ? hc[e + 1 < E ? e + 1 : e] : hkp;
                const float km = e > 0 ? hc[e > 0 ? e - 1 : 0] : hkm;
                const float hl = kp + km + hjp[e] + hjm[e] + hip[e] + him[e] - 6.0f * hc[e];
                double src;
                switch (c) {
                case 0: src = gx[e] * gx[e]; break;
                case 1: src = gy[e] * gy[e]; break;
                case 2: src = gz[e] * gz[e]; break;
                case 3: src = gx[e] * gy[e]; break;
                case 4: src = gx[e] * gz[e]; break;
                default: src = gy[e] * gz[e]; break;
                }
                const double rhs = d * (laplnorm * (double)hl - friction * (double)hv[e] + srcamp * src);
                hdn[e] = hv[e] + (float)rhs;
                hn[e] = hc[e] + (float)(d * (double)hdn[e]);
            }
            st_row<float, E>(F.hd[c] + poff + rj, lane, hdn);
            if (MODE & DRIFT) st_row<float, E>(F.h[par ^ 1][c] + poff + rj, lane, hn);
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            prev[e] = cur[e];
            cur[e] = nxt[e];
        }
    }
}

template <typename T, int E, int MODE>
__global__ __launch_bounds__(kBlock) void k_row_tensor(Fields F, DevBg *bg, int loop, int fresh)
{
    row_tensor_body<T, E, MODE>(F, bg, loop, fresh, gridDim.x, blockIdx.x);
}

template <typename T, int E>
__global__ __launch_bounds__(kBlock) void kb_row_tensor(const RunDesc *__restrict__ runs,
                                                         const uint2 *__restrict__ map, unsigned rslot)
{
    const uint2 mp = map[blockIdx.x];
    const RunDesc &R = runs[mp.x];
    row_tensor_body<T, E, KICKDRIFT>(R.F, R.bg, 1, R.march ? 1 + (int)rslot : 0, R.G, mp.y);
}


template <typename T, bool SPLIT, int E>
__global__ __launch_bounds__(kBlock) void k_row_reduce(Fields F, Pot P, DevBg *bg, double *partials, int loop)
{
    __shared__ double smem[3 * kWaves];
    row_reduce_body<T, SPLIT, E>(F, P, bg, partials, loop, gridDim.x, blockIdx.x, smem);
}

template <typename T, bool SPLIT, int E, int MODE>
__global__ __launch_bounds__(kBlock) void k_row_kick(Fields F, Pot P, DevBg *bg, int loop)
{
    row_kick_body<T, SPLIT, E, MODE>(F, P, bg, loop, gridDim.x, blockIdx.x);
}

template <typename T, bool SPLIT, int E>
__global__ __launch_bounds__(kBlock) void kb_row_reduce(const RunDesc *__restrict__ runs,
                                                         const uint2 *__restrict__ map)
{
    __shared__ double smem[3 * kWaves];
    const uint2 mp = map[blockIdx.x];
    const RunDesc &R = runs[mp.x];
    row_reduce_body<T, SPLIT, E>(R.F, R.P, R.bg, R.partials, 1, R.G, mp.y, smem);
}

template <typename T, bool SPLIT, int E>
__global__ __launch_bounds__(kBlock) void kb_row_kick(const RunDesc *__restrict__ runs, const uint2 *__restrict__ map)
{
    const uint2 mp = map[blockIdx.x];
    const RunDesc &R = runs[mp.x];
    row_kick_body<T, SPLIT, E, KICKDRIFT>(R.F, R.P, R.bg, 1, R.G, mp.y);
}

// ------------------------------------------------------ fused march kernel
//
// One loop step of the scalar sector for 128 <= N <= 256 in one pass that also
// produces the next step's Friedmann sums.  A workgroup of 9 waves owns a tile of
// 8 j rows x the whole k row and marches along i over TI planes.  At each plane
// every wave kicks and drifts its row (the row kernel's arithmetic), stores it
// (waves 0-7 only), and puts the new row in LDS.  The sums of the new f follow
// from  -sum_x f lapl f = sum_x sum_d (f(x + e_d) - f(x))^2  on the periodic
// lattice: the k pair from a lane shuffle, the j pair from the next wave's LDS
// row, the i pair from the wave's previous plane.  Wave 8 recomputes the first
// row of the next tile (the partner of row 7's j pair) and one extra plane step
// gives the last i pair; both are stored by their owners, not here.  V and V' of
// the new f (about the new homogeneous mode for split storage) are summed per
// owned site.  Partials per workgroup, in a fixed order.
constexpr int kTileJ = 8;
constexpr int kMarchWaves = kTileJ + 1;
constexpr int kMarchThreads = 64 * kMarchWaves;

// Planes per tile: a tile spans the whole slab of its XCD at N = 256 (32 planes,
// one workgroup per CU; measured 0.059 ms per step against 0.068 with 16 and
// 0.081 with 8), 8 planes at N = 128.
__device__ __host__ __forceinline__ unsigned march_ti(unsigned N) { return N >= 256 ? N / kXcd : 8u; }

// A value equal in every lane, moved to scalar registers.
__device__ __forceinline__ double uniform(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// SUMS: the sums alone, of the current f taken as the new f (loop start), with
// the same partition and operation order as a step's, so that a loop split into
// several calls gives the same bits as one call.
// The rows a plane step streams from HBM (f of plane i + 1, fd of plane i) are
// loaded one plane step ahead; the plane loop is unrolled by 3 over three row
// sets, so a row set in flight is never moved between registers (a move waits
// for its load).
// JL: where a kick finds its j +- 1 rows of old f.  false: loaded from L2 with
// the plane's other rows (the neighbouring waves streamed them from HBM one step
// earlier); true: the neighbouring waves' own rows, passed through LDS in FP64
// (the tile's halo rows j0 - 1 and j0 + 9 by waves 0 and 8) and read right after
// the previous plane's barrier, so that the sums' arithmetic covers their
// latency.  The same values either way, so the same results.
// POT: the hilltop potential (kPotHill: V' linear in f, so its Friedmann sums are
// those of f and f^2, turned into sum V and sum V' by the step's close) or the
// tabulated one (kPotTab: V and V' interpolated per site).
// Loads and stores go through buffer resources with the row offset in a scalar
// register, so a row costs no per-lane address arithmetic.
// A step launch first closes the previous step from its record in slot sp (its
// first rows already in flight) and writes the result to the other slot.
constexpr int kPotHill = 0, kPotTab = 1;

// LDS rows of E values per lane stored as two halves of E/2 (8 or 16 bytes per
// lane in each), so that each LDS access of a wave is free of bank conflicts.
template <typename U, int E>
struct LdsRow {
    U v[2][64][E / 2];
    __device__ __forceinline__ U &at(unsigned lane, int e) { return v[e / (E / 2)][lane][e % (E / 2)]; }
};

template <typename T, int E, bool JL>
struct MarchLds {
    LdsRow<double, E> oldf[JL ? 2 : 1][JL ? kTileJ + 3 : 1];   // old f of plane t, rows j0 - 1 .. j0 + 9
    LdsRow<T, E> newf[2][kMarchWaves];                          // new f of plane t, rows j0 .. j0 + 8
    double smem[3 * kMarchWaves];
};

// Dynamic LDS of a march launch: the index of the tabulated potential's table
// (and its records with four sites per lane, for the variant whose old rows stay
// out of LDS: with them, 64-82 KB of its own, the records would pass the 160 KB
// a workgroup can hold).
template <int POT, bool JL, int E>
__host__ __device__ constexpr unsigned march_lds()
{
    return POT == kPotTab ? ((E == 4 && !JL) ? kTabRecordsLds : kTabIndexLds) : 0u;
}
constexpr unsigned kLdsPerWorkgroup = 160 * 1024;

template <typename T, bool SPLIT, int E, int POT, bool JL, bool SUMS = false>
__device__ __forceinline__ void march_body(const Fields &F, const Pot &P0, DevBg *bg, double *partials, unsigned blk,
                                           MarchLds<T, E, JL> &L, unsigned sp = 0)
{
    static_assert(E == 2 || E == 4, "fused steps hold 2 or 4 sites per lane");
    // static LDS (the rows, this step's record, slack for the sums) and dynamic
    static_assert(sizeof(MarchLds<T, E, JL>) + sizeof(DevBg) + 1024 + march_lds<POT, JL, E>() <= kLdsPerWorkgroup,
                  "a march launch's LDS must fit one workgroup");
    const unsigned N = (unsigned)F.N, logN = (unsigned)F.logN, m = N - 1, lane = threadIdx.x & 63;
    const unsigned w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned G = (N / kTileJ) * (N / march_ti(N));
    constexpr bool kLds = JL && !SUMS;
    int par;
    bool pmarched = false;
    __shared__ DevBg sb;        // step launches: this step's record
    double tot[3] = {0.0, 0.0, 0.0};
    if constexpr (SUMS) {
        if (bg->done) return;
        par = bg->parity;
    } else {
        // one round of loads: the previous record (flags, and whole into LDS) and
        // its step's partials, in sum_partials' order
        const DevBg *prev = F.rec + sp;
        if (w == 0) copy_bg_wave(&sb, prev, lane);
        if (threadIdx.x < kBlock)
            for (unsigned b = threadIdx.x; b < G; b += kBlock)
#pragma unroll
                for (int q = 0; q < 3; ++q) tot[q] += partials[((size_t)sp * G + b) * 3 + q];
        if (prev->done) {
            // the ended loop's record goes on into this step's slot, which this
            // step's tensor launch and the next step read: left alone, it would
            // hold the record of two steps back, not done
            if (blk == 0 && w == 0) copy_bg_wave(F.rec + (sp ^ 1), prev, lane);
            return;
        }
        pmarched = prev->marched != 0;
        par = prev->parity ^ (pmarched ? 1 : 0);
    }
    PotView Pv = pot_view(P0);
    if constexpr (POT == kPotTab) {
        // the table's index levels into this launch's dynamic LDS (published by
        // the barriers below: the close's, or the one here for the sums pass)
        extern __shared__ __align__(16) char tab_lds[];
        stage_tab_index(Pv, tab_lds, kMarchThreads, march_lds<POT, JL, E>());
        if constexpr (SUMS) __syncthreads();
    }
    // fd is read from [par] and written to [par ^ 1]: the halo row and the extra
    // plane read fd rows that their owning workgroups update in the same launch
    const unsigned bytes = N * N * N * (unsigned)sizeof(T);
    const __amdgpu_buffer_rsrc_t rf = field_rsrc(F.f[par], bytes), rfo = field_rsrc(F.f[par ^ 1], bytes);
    const __amdgpu_buffer_rsrc_t rv = field_rsrc(F.fd[par], bytes), rvo = field_rsrc(F.fd[par ^ 1], bytes);
    // tile of this workgroup: XCD blk % 8 owns planes [x N/8, (x+1) N/8)
    const unsigned TI = march_ti(N), jblocks = N / kTileJ;
    const unsigned xcd = blk % kXcd, lb = blk / kXcd;
    const unsigned i0 = xcd * (N / kXcd) + (lb / jblocks) * TI;
    const unsigned j0 = (lb % jblocks) * kTileJ;
    const unsigned j = (j0 + w) & m, jp = (j + 1) & m, jm = (j - 1) & m;
    const bool owner = w < (unsigned)kTileJ;
    // LDS rows: the tile's halo rows of old f are j0 - 1 (loaded by wave 0) and
    // j0 + 9 (wave 8, whose own row j0 + 8 is the next tile's first, recomputed)
    const bool halo = kLds && (w == 0 || w == (unsigned)kTileJ);
    const unsigned jh = w == 0 ? jm : jp;
    const unsigned hslot = w == 0 ? 0u : (unsigned)kTileJ + 2u;
    auto roff = [&](unsigned i, unsigned jj) -> unsigned {
        return ((((i & m) << logN) + jj) << logN) * (unsigned)sizeof(T);
    };
    // the rolling rows: f of planes i - 1 and i, and the new f of plane i - 1, in
    // FP64 (each value converted once)
    T fdn[E], fcur[E];
    bld_row<T, E>(rf, roff(i0 - 1, j), lane, fdn);
    bld_row<T, E>(rf, roff(i0, j), lane, fcur);
    double ddn[E], dcur[E], dprev[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        ddn[e] = (double)fdn[e];
        dcur[e] = (double)fcur[e];
        dprev[e] = 0.0;
    }
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    int err = 0;
    // rows plane step t streams: f(i0+t+1) row j, fd(i0+t) row j, and f(i0+t)
    // rows j +- 1 (JL false) or the halo row f(i0+t+1) (JL true)
    struct Rows {
        T up[E], vin[E], a[E], b[E];
    };
    Rows S0, S1, S2;
    auto load_rows = [&](unsigned t, Rows &r) __attribute__((always_inline)) {
        bld_row<T, E>(rf, roff(i0 + t + 1, j), lane, r.up);
        if constexpr (!SUMS) {
            bld_row<T, E>(rv, roff(i0 + t, j), lane, r.vin);
            if constexpr (!JL) {
                bld_row<T, E>(rf, roff(i0 + t, jp), lane, r.a);
                bld_row<T, E>(rf, roff(i0 + t, jm), lane, r.b);
            } else if (halo) {
                bld_row<T, E>(rf, roff(i0 + t + 1, jh), lane, r.a);
            }
        }
    };
    if constexpr (!SUMS) load_rows(0, S0);
    if constexpr (kLds) {
        // plane i0's old f rows for the j +- 1 neighbours of the first kick
#pragma unroll
        for (int e = 0; e < E; ++e) L.oldf[0][w + 1].at(lane, e) = dcur[e];
        if (halo) {
            T hr[E];
            bld_row<T, E>(rf, roff(i0, jh), lane, hr);
#pragma unroll
            for (int e = 0; e < E; ++e) L.oldf[0][hslot].at(lane, e) = (double)hr[e];
        }
    }
    // the step's scalars: the loop's first record as it stands, later ones closed
    // here with the previous step's partials (every workgroup, the same bits)
    double d = 0.0, dl = 0.0, cf = 0.0, dpot = 0.0, ch = 0.0, ck = 0.0, vpmean = 0.0, phibar = 0.0,
           phibar_next = 0.0;
    double *pout = partials;
    if constexpr (SUMS) {
        phibar = SPLIT ? bg->phibar : 0.0;
        phibar_next = phibar;
    } else {
        if (pmarched) block_sum<3>(tot, L.smem);
        __syncthreads();
        if (threadIdx.x == 0) {
            if (pmarched) close_state(sb, tot, (double)N * N * N, SPLIT);
            sb.marched = 1;
        }
        __syncthreads();
        if (blk == 0 && w == 0) copy_bg_wave(F.rec + (sp ^ 1), &sb, lane);
        if (sb.done) return;
        d = uniform(sb.nx.d);
        const double laplnorm = uniform(sb.nx.laplnorm), friction = uniform(sb.nx.friction);
        const double potnorm = uniform(sb.nx.potnorm);
        vpmean = SPLIT ? uniform(sb.nx.vpmean) : 0.0;
        phibar = SPLIT ? uniform(sb.phibar) : 0.0;
        phibar_next = SPLIT ? uniform(sb.nx.phibar_next) : 0.0;
        // fd += d (laplnorm lapl f - friction fd - potnorm (V'(f) - <V'>)) as
        // fd (1 - d friction) + d laplnorm lapl f + pt, with pt = -d potnorm (V' - <V'>):
        // for the hilltop V' = -hV2 (phibar + f), pt = ch f + ck
        dl = d * laplnorm;
        cf = 1.0 - d * friction;
        dpot = d * potnorm;
        if constexpr (POT == kPotHill) {
            ch = dpot * P0.hV2;
            ck = SPLIT ? dpot * (P0.hV2 * phibar + vpmean) : 0.0;
        }
        pout = partials + (size_t)(sp ^ 1) * G * 3;
    }
    // the first kick's j +- 1 rows (JL: from LDS, after the barriers above)
    double njp[E], njm[E];
    if constexpr (kLds) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            njp[e] = L.oldf[0][w + 2].at(lane, e);
            njm[e] = L.oldf[0][w].at(lane, e);
        }
    }
    // intervals of the tabulated potential: each lane's sites start from the
    // intervals of the previous plane's sites (the kick's) and of the same site
    // (the sums')
    int hint[E];
#pragma unroll
    for (int e = 0; e < E; ++e) hint[e] = POT == kPotTab ? Pv.start((SPLIT ? phibar : 0.0) + dcur[e]) : 0;
    // plane step t with its rows in `cur`; the rows of step t + 1 go to `fill`
    // (the last step reloads plane TI's rows, unused, so that every step loads)
    auto plane_step = [&](unsigned t, Rows &cur, Rows &fill) __attribute__((always_inline)) {
        if constexpr (SUMS) load_rows(t, cur);
        else load_rows(t + 1 <= TI ? t + 1 : TI, fill);
        T fnew[E];
        double dnew[E], dup[E];
#pragma unroll
        for (int e = 0; e < E; ++e) dup[e] = (double)cur.up[e];
        if constexpr (SUMS) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                fnew[e] = fcur[e];
                dnew[e] = dcur[e];
            }
        } else {
            const double kp_last = __shfl(dcur[0], (int)((lane + 1) & 63));
            const double km_first = __shfl(dcur[E - 1], (int)((lane + 63) & 63));
            double tdv[E];
            if constexpr (POT == kPotTab) {
                double tv[E];
#pragma unroll
                for (int e = 0; e < E; ++e) tv[e] = SPLIT ? phibar + dcur[e] : dcur[e];
                Pv.dV<E>(tv, hint, tdv, &err);
            }
            T vout[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const double c = dcur[e];
                const double kp = e < E - 1 ? dcur[e + 1 < E ? e + 1 : e] : kp_last;
                const double km = e > 0 ? dcur[e > 0 ? e - 1 : 0] : km_first;
                const double yp = JL ? njp[e] : (double)cur.a[e], ym = JL ? njm[e] : (double)cur.b[e];
                const double lap = lapl_sites(kp, km, yp, ym, dup[e], ddn[e], c);
                double pt;
                if constexpr (POT == kPotHill) {
                    pt = SPLIT ? fma(ch, c, ck) : ch * c;
                } else {
                    pt = -dpot * (SPLIT ? tdv[e] - vpmean : tdv[e]);
                }
                const double v = fma(dl, lap, fma(cf, (double)cur.vin[e], pt));
                vout[e] = (T)v;
                fnew[e] = (T)fma(d, v, c);
                dnew[e] = (double)fnew[e];
            }
            if (owner && t < TI) {
                bst_row<T, E>(rvo, roff(i0 + t, j), lane, vout);
                bst_row<T, E>(rfo, roff(i0 + t, j), lane, fnew);
            }
            if constexpr (JL) {
                // plane t + 1's old f rows for the next kick's j +- 1 neighbours
#pragma unroll
                for (int e = 0; e < E; ++e) L.oldf[(t + 1) & 1][w + 1].at(lane, e) = dup[e];
                if (halo)
#pragma unroll
                    for (int e = 0; e < E; ++e) L.oldf[(t + 1) & 1][hslot].at(lane, e) = (double)cur.a[e];
            }
        }
#pragma unroll
        for (int e = 0; e < E; ++e) L.newf[t & 1][w].at(lane, e) = fnew[e];
        __syncthreads();
        if constexpr (kLds) {
            // the next kick's j +- 1 rows, in flight while the sums below are formed
#pragma unroll
            for (int e = 0; e < E; ++e) {
                njp[e] = L.oldf[(t + 1) & 1][w + 2].at(lane, e);
                njm[e] = L.oldf[(t + 1) & 1][w].at(lane, e);
            }
        }
        if (owner) {
            if (t < TI) {
                const double knext = __shfl(dnew[0], (int)((lane + 1) & 63));
                // sum V and sum V' of the lane's sites (the window's hinge form)
                if constexpr (POT == kPotTab) {
                    double tv[E];
#pragma unroll
                    for (int e = 0; e < E; ++e) tv[e] = SPLIT ? phibar_next + dnew[e] : dnew[e];
                    Pv.sums<E>(tv, hint, acc1, acc2, &err);
                }
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const double c = dnew[e];
                    const double dk = (e < E - 1 ? dnew[e + 1 < E ? e + 1 : e] : knext) - c;
                    const double dj = (double)L.newf[t & 1][w + 1].at(lane, e) - c;
                    acc0 += dk * dk + dj * dj;
                    if constexpr (POT == kPotHill) {
                        acc1 += c;
                        acc2 += c * c;
                    }
                }
            }
            if (t >= 1) {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const double di = dnew[e] - dprev[e];
                    acc0 += di * di;
                }
            }
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            dprev[e] = dnew[e];
            ddn[e] = dcur[e];
            fcur[e] = cur.up[e];
            dcur[e] = dup[e];
        }
    };
    // set of step t: S[t mod 3]; step t fills S[(t + 1) mod 3], last read by
    // step t - 2.  TI + 1 steps, a multiple of 3 (march_ok).
    for (unsigned t = 0; t <= TI; t += 3) {
        plane_step(t, S0, S1);
        plane_step(t + 1, S1, S2);
        plane_step(t + 2, S2, S0);
    }
    flag_error(bg, err);
    // fixed-order workgroup sum over the 9 waves
    double acc[3] = {acc0, acc1, acc2};
#pragma unroll
    for (int q = 0; q < 3; ++q) acc[q] = wave_sum(acc[q]);
    if (lane == 0)
        for (int q = 0; q < 3; ++q) L.smem[q * kMarchWaves + w] = acc[q];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) {
            double s = 0.0;
            for (int ww = 0; ww < kMarchWaves; ++ww) s += L.smem[q * kMarchWaves + ww];
            pout[(size_t)blk * 3 + q] = s;
        }
}

// Two variants with the same results, picked by the tuner: the j +- 1 rows of
// old f from L2 (k_march_pf) or through LDS (k_march).

template <typename T, bool SPLIT, int E, int POT>
__global__ __launch_bounds__(kMarchThreads) void k_march(Fields F, Pot P, DevBg *bg, double *partials, unsigned sp)
{
    __shared__ MarchLds<T, E, true> L;
    march_body<T, SPLIT, E, POT, true>(F, P, bg, partials, blockIdx.x, L, sp);
}

template <typename T, bool SPLIT, int E, int POT>
__global__ __launch_bounds__(kMarchThreads) void k_march_pf(Fields F, Pot P, DevBg *bg, double *partials,
                                                            unsigned sp)
{
    __shared__ MarchLds<T, E, false> L;
    march_body<T, SPLIT, E, POT, false>(F, P, bg, partials, blockIdx.x, L, sp);
}

template <typename T, bool SPLIT, int E, int POT>
__global__ __launch_bounds__(kMarchThreads) void kb_march_pf(const RunDesc *__restrict__ runs,
                                                              const uint2 *__restrict__ map, unsigned sp)
{
    __shared__ MarchLds<T, E, false> L;
    const uint2 mp = map[blockIdx.x];
    const RunDesc &R = runs[mp.x];
    march_body<T, SPLIT, E, POT, false>(R.F, R.P, R.bg, R.partials, mp.y, L, sp);
}

template <typename T, bool SPLIT, int E, int POT>
__global__ __launch_bounds__(kMarchThreads) void kb_march(const RunDesc *__restrict__ runs,
                                                           const uint2 *__restrict__ map, unsigned sp)
{
    __shared__ MarchLds<T, E, true> L;
    const uint2 mp = map[blockIdx.x];
    const RunDesc &R = runs[mp.x];
    march_body<T, SPLIT, E, POT, true>(R.F, R.P, R.bg, R.partials, mp.y, L, sp);
}

template <typename T, bool SPLIT, int E, int POT>
__global__ __launch_bounds__(kMarchThreads) void k_march_sums(Fields F, Pot P, DevBg *bg, double *partials)
{
    __shared__ MarchLds<T, E, false> L;
    march_body<T, SPLIT, E, POT, false, true>(F, P, bg, partials, blockIdx.x, L);
}

template <typename T, bool SPLIT, int E, int POT>
__global__ __launch_bounds__(kMarchThreads) void kb_march_sums(const RunDesc *__restrict__ runs,
                                                                const uint2 *__restrict__ map)
{
    __shared__ MarchLds<T, E, false> L;
    const uint2 mp = map[blockIdx.x];
    const RunDesc &R = runs[mp.x];
    march_body<T, SPLIT, E, POT, false, true>(R.F, R.P, R.bg, R.partials, mp.y, L);
}

__global__ __launch_bounds__(kBlock) void k_loop_end(DevBg *bg, const DevBg *rec, unsigned sp, const double *partials,
                                                      unsigned G, unsigned long long gs, int split)
{
    __shared__ double smem[3 * kWaves];
    __shared__ DevBg sb;
    loop_end_body(bg, rec, sp, partials, G, (size_t)gs, split, smem, sb);
}

__global__ __launch_bounds__(kBlock) void k_sums_finish(DevBg *bg, const double *partials, unsigned G,
                                                         unsigned long long gs, DevBg *rec)
{
    __shared__ double smem[3 * kWaves];
    sums_finish_body(bg, partials, G, (size_t)gs, rec, smem);
}

__global__ __launch_bounds__(kBlock) void kb_loop_end(const RunDesc *__restrict__ runs, unsigned sp)
{
    __shared__ double smem[3 * kWaves];
    __shared__ DevBg sb;
    const RunDesc &R = runs[blockIdx.x];
    if (!R.march) return;
    const size_t gs = (size_t)R.F.N * R.F.N * R.F.N;
    loop_end_body(R.bg, R.F.rec, sp, R.partials, R.Gm, gs, R.split, smem, sb);
}

__global__ __launch_bounds__(kBlock) void kb_sums_finish(const RunDesc *__restrict__ runs)
{
    __shared__ double smem[3 * kWaves];
    const RunDesc &R = runs[blockIdx.x];
    if (!R.march) return;
    const size_t gs = (size_t)R.F.N * R.F.N * R.F.N;
    sums_finish_body(R.bg, R.partials, R.Gm, gs, R.F.rec, smem);
}

// ------------------------------------------------ resident small lattices
//
// Batched 64^3 lattices without SIGW: one workgroup of 8 waves owns a whole
// lattice and runs every loop step of a launch itself.  A plane is its 64 rows,
// 8 consecutive rows per wave with one site per lane, so a kick finds its k +- 1
// neighbours by lane shuffles, its j +- 1 rows in the wave's own registers (the
// wave's two edge rows through LDS), its i +- 1 rows in the registers of the
// planes the wave marches through.  Each site is kicked once, with no halo, and
// the next step's Friedmann sums (the same pairs as the fused march kernel's,
// including the i pair of planes N - 1 and 0 from plane 0's new rows kept in
// registers) never leave the workgroup: after the last plane the workgroup
// reduces them and closes the step on its own copy of the background
// (close_state), then starts the next step.  A launch first forms the sums of
// the current f (the same partition and order), so a loop split into several
// launches gives the same bits as one.  Reads f, fd and writes them once per
// step (16 B/site in split storage).  8 waves of 8 rows hold every row in
// registers without spills (243 VGPRs, 2 waves per SIMD); 16 waves of 4 rows
// spill at the 128 VGPRs that 4 waves per SIMD allow and step 1 % slower
// (256 x 64^3, split storage: 0.2255 against 0.2293 ms per step).
constexpr int kResN = 64;
constexpr int kResWaves = 8;
constexpr int kResRows = kResN / kResWaves;   // rows per wave
constexpr int kResThreads = 64 * kResWaves;

template <typename T>
struct ResLds {
    double oldf[2][kResN][64];   // old f of plane t: the waves' edge rows
    T newf[2][kResN][64];        // new f of plane t: the waves' first rows
    double smem[3 * kResWaves];
};

template <typename T, bool SPLIT, int POT>
__device__ __forceinline__ void resident_body(const RunDesc &R, int nsteps, ResLds<T> &L, DevBg &sb)
{
    constexpr unsigned N = kResN, m = N - 1;
    const unsigned lane = threadIdx.x & 63;
    const unsigned w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned j0 = w * kResRows;
    const Fields &F = R.F;
    DevBg *bg = R.bg;
    if (w == 0) copy_bg_wave(&sb, bg, lane);
    __syncthreads();
    if (sb.done) return;
    const PotView Pv = pot_view(R.P);
    const unsigned bytes = N * N * N * (unsigned)sizeof(T);
    const double gsd = (double)N * N * N;
    int err = 0;
    auto roff = [&](unsigned i, unsigned j) -> unsigned { return ((((i & m) << 6) + (j & m)) << 6) * (unsigned)sizeof(T); };
    // one pass over the planes: SUMS the sums of the current f (launch start),
    // otherwise a loop step (kick, drift, stores, the new f's sums)
    auto pass = [&](auto sums_tag) __attribute__((always_inline)) {
        constexpr bool SUMS = decltype(sums_tag)::value;
        const int par = sb.parity;
        const __amdgpu_buffer_rsrc_t rf = field_rsrc(F.f[par], bytes), rfo = field_rsrc(F.f[par ^ 1], bytes);
        const __amdgpu_buffer_rsrc_t rv = field_rsrc(F.fd[par], bytes), rvo = field_rsrc(F.fd[par ^ 1], bytes);
        double d = 0.0, dl = 0.0, cf = 0.0, dpot = 0.0, ch = 0.0, ck = 0.0, vpmean = 0.0, phibar = 0.0, phibar_next;
        if constexpr (SUMS) {
            phibar = SPLIT ? sb.phibar : 0.0;
            phibar_next = phibar;
        } else {
            d = uniform(sb.nx.d);
            dl = d * uniform(sb.nx.laplnorm);
            cf = 1.0 - d * uniform(sb.nx.friction);
            dpot = d * uniform(sb.nx.potnorm);
            vpmean = SPLIT ? uniform(sb.nx.vpmean) : 0.0;
            phibar = SPLIT ? uniform(sb.phibar) : 0.0;
            phibar_next = SPLIT ? uniform(sb.nx.phibar_next) : 0.0;
            if constexpr (POT == kPotHill) {
                ch = dpot * R.P.hV2;
                ck = SPLIT ? dpot * (R.P.hV2 * phibar + vpmean) : 0.0;
            }
        }
        // rolling rows of the wave: old f of planes t - 1 and t, new f of plane
        // t - 1 and of plane 0 (the last i pair); only plane t in FP64, the others
        // in their storage type (each converts exactly)
        double dcur[kResRows];
        T ddn[kResRows], dprev[kResRows], dfirst[kResRows];
        int hint[kResRows];   // intervals of the tabulated potential, per row
#pragma unroll
        for (int r = 0; r < kResRows; ++r) {
            T a[1], b[1];
            bld_row<T, 1>(rf, roff(m, j0 + r), lane, a);
            bld_row<T, 1>(rf, roff(0, j0 + r), lane, b);
            ddn[r] = a[0];
            dcur[r] = (double)b[0];
            dprev[r] = (T)0;
            dfirst[r] = (T)0;
            hint[r] = POT == kPotTab ? Pv.start((SPLIT ? phibar : 0.0) + dcur[r]) : 0;
        }
        double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
        struct Rows {
            T up[kResRows], vin[kResRows];
        };
        Rows S0, S1;
        auto load_rows = [&](unsigned t, Rows &q) __attribute__((always_inline)) {
#pragma unroll
            for (int r = 0; r < kResRows; ++r) {
                T a[1];
                bld_row<T, 1>(rf, roff(t + 1, j0 + r), lane, a);
                q.up[r] = a[0];
                if constexpr (!SUMS) {
                    bld_row<T, 1>(rv, roff(t, j0 + r), lane, a);
                    q.vin[r] = a[0];
                }
            }
        };
        load_rows(0, S0);
        if constexpr (!SUMS) {
            L.oldf[0][j0][lane] = dcur[0];
            L.oldf[0][j0 + kResRows - 1][lane] = dcur[kResRows - 1];
            __syncthreads();
        }
        // plane step t with its rows in `cur`; plane t + 1's go to `fill` (the set
        // step t - 1 used)
        auto plane_step = [&](unsigned t, Rows &cur, Rows &fill) __attribute__((always_inline)) {
            load_rows(t + 1 < N ? t + 1 : t, fill);
            double dnew[kResRows];
            T fnew[kResRows];
            if constexpr (SUMS) {
#pragma unroll
                for (int r = 0; r < kResRows; ++r) {
                    fnew[r] = (T)dcur[r];
                    dnew[r] = dcur[r];
                }
            } else {
                const double below = L.oldf[t & 1][(j0 - 1) & m][lane];
                const double above = L.oldf[t & 1][(j0 + kResRows) & m][lane];
#pragma unroll
                for (int r = 0; r < kResRows; ++r) {
                    const double c = dcur[r];
                    const double kp = __shfl(c, (int)((lane + 1) & 63)), km = __shfl(c, (int)((lane + 63) & 63));
                    const double yp = r < kResRows - 1 ? dcur[r + 1 < kResRows ? r + 1 : r] : above;
                    const double ym = r > 0 ? dcur[r > 0 ? r - 1 : 0] : below;
                    const double lap = lapl_sites(kp, km, yp, ym, (double)cur.up[r], (double)ddn[r], c);
                    double pt;
                    if constexpr (POT == kPotHill) {
                        pt = SPLIT ? fma(ch, c, ck) : ch * c;
                    } else {
                        double dV = Pv.dV1(SPLIT ? phibar + c : c, hint[r], &err);
                        if (SPLIT) dV -= vpmean;
                        pt = -dpot * dV;
                    }
                    const double v = fma(dl, lap, fma(cf, (double)cur.vin[r], pt));
                    T a[1] = {(T)v};
                    fnew[r] = (T)fma(d, v, c);
                    dnew[r] = (double)fnew[r];
                    T b[1] = {fnew[r]};
                    bst_row<T, 1>(rvo, roff(t, j0 + r), lane, a);
                    bst_row<T, 1>(rfo, roff(t, j0 + r), lane, b);
                }
                // plane t + 1's edge rows for the neighbouring waves' next kick
                L.oldf[(t + 1) & 1][j0][lane] = (double)cur.up[0];
                L.oldf[(t + 1) & 1][j0 + kResRows - 1][lane] = (double)cur.up[kResRows - 1];
            }
            L.newf[t & 1][j0][lane] = fnew[0];
            __syncthreads();
            const T above_new = L.newf[t & 1][(j0 + kResRows) & m][lane];
            if (t == 0) {
#pragma unroll
                for (int r = 0; r < kResRows; ++r) dfirst[r] = fnew[r];
            }
#pragma unroll
            for (int r = 0; r < kResRows; ++r) {
                const double c = dnew[r];
                const double dk = __shfl(c, (int)((lane + 1) & 63)) - c;
                const double dj = (r < kResRows - 1 ? dnew[r + 1 < kResRows ? r + 1 : r] : (double)above_new) - c;
                acc0 += dk * dk + dj * dj;
                if (t >= 1) {
                    const double di = c - (double)dprev[r];
                    acc0 += di * di;
                }
                if (t == N - 1) {
                    const double di = (double)dfirst[r] - c;   // the periodic pair (N - 1, 0)
                    acc0 += di * di;
                }
                if constexpr (POT == kPotHill) {
                    acc1 += c;
                    acc2 += c * c;
                } else {
                    double V, dV;
                    Pv.eval1(SPLIT ? phibar_next + c : c, &V, &dV, hint[r], &err);
                    acc1 += V;
                    acc2 += dV;
                }
            }
#pragma unroll
            for (int r = 0; r < kResRows; ++r) {
                dprev[r] = fnew[r];
                ddn[r] = (T)dcur[r];
                dcur[r] = (double)cur.up[r];
            }
        };
        // N = 64 planes in steps of two over the two row sets
        for (unsigned t = 0; t < N; t += 2) {
            plane_step(t, S0, S1);
            plane_step(t + 1, S1, S0);
        }
        // fixed-order workgroup sums, then the close on the workgroup's record
        double acc[3] = {acc0, acc1, acc2};
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[q] = wave_sum(acc[q]);
        if (lane == 0)
            for (int q = 0; q < 3; ++q) L.smem[q * kResWaves + w] = acc[q];
        __threadfence_block();   // this step's stores before the next pass reads them
        __syncthreads();
        if (threadIdx.x == 0) {
            double tot[3];
            for (int q = 0; q < 3; ++q) {
                double s = 0.0;
                for (int ww = 0; ww < kResWaves; ++ww) s += L.smem[q * kResWaves + ww];
                tot[q] = s;
            }
            if constexpr (SUMS) {
                hill_sums(sb, gsd, tot);
                sb.red[0] = tot[0];
                sb.red[1] = tot[1];
                sb.red[2] = tot[2];
                sb.nx = step_scalars(&sb, gsd);
            } else {
                close_state(sb, tot, gsd, SPLIT);
            }
        }
        __syncthreads();
    };
    pass(std::true_type{});
    for (int s = 0; s < nsteps && !sb.done; ++s) pass(std::false_type{});
    // the closed record back to the run, with the error flag any thread raised
    flag_error(bg, err);
    __threadfence_block();
    __syncthreads();
    if (threadIdx.x == 0 && bg->error) sb.error = bg->error;
    __syncthreads();
    if (w == 0) copy_bg_wave(bg, &sb, lane);
}

template <typename T, bool SPLIT, int POT>
__global__ __launch_bounds__(kResThreads) void kb_resident(const RunDesc *__restrict__ runs,
                                                           const uint2 *__restrict__ map, int nsteps)
{
    __shared__ ResLds<T> L;
    __shared__ DevBg sb;
    resident_body<T, SPLIT, POT>(runs[map[blockIdx.x].x], nsteps, L, sb);
}

// ----------------------------------------------------- tensor march kernel
//
// Tensor kick + drift of a loop step (src/evolution.cpp:185-253, 274-280) for
// N >= 64: a workgroup of 8 waves owns 8 j rows x one k segment and marches along
// i over TI planes.  Each wave keeps the i-1, i, i+1 rows of f and of the six h
// components in registers and puts its plane-i rows in LDS (double-buffered, with
// the tile's two halo rows), where the neighbouring waves read their j +- 1 rows:
// every h row comes from HBM about once (the j +- 1 rows streamed from L2 instead
// were evicted between the two uses, doubling the reads).  hdot is updated in
// place (pointwise), h goes to the other buffer.  Same arithmetic as the row
// tensor kernel; no cross-site sums, so the shape (E sites per lane, PF) is
// tunable.  PF: the rows of plane i+2, the hdot rows and the halo rows of plane
// i+1 and the segment-edge values of the plane-(i+1) rows are loaded before the
// plane-i barrier, so they arrive while plane i is computed.
constexpr int kTensorWaves = 8;

// Planes per tile: with E = 4 one workgroup per CU (LDS of 143 KB) marching the
// XCD's whole slab at N = 256; otherwise 16 (8 below N = 128).
__device__ __host__ __forceinline__ unsigned tmarch_ti(unsigned N, int E)
{
    return (E == 4 && N >= 256) ? N / kXcd : (N >= 128 ? 16u : 8u);
}

// LDS tile of one plane: rows j0-1 .. j0+8 of the six h components and of f.
template <typename T, int E>
struct TensorTile {
    float h[2][6][kTensorWaves + 2][64 * E];
    T f[2][kTensorWaves + 2][64 * E];
};

// The two segment-edge values of a row (sites k0 - 1 and k0 + 64 E), wave-uniform.
template <typename T, int E>
__device__ __forceinline__ void row_edges(const T *__restrict__ row, unsigned k0, unsigned m, T &lo, T &hi)
{
    lo = first_lane(row[(k0 - 1) & m]);
    hi = first_lane(row[(k0 + 64u * E) & m]);
}

template <typename T, int E, bool PF>
__device__ __forceinline__ void tensor_march_body(const Fields &F, DevBg *bg, int fresh, unsigned blk,
                                                  TensorTile<T, E> &L)
{
    if (fresh) bg = F.rec + (fresh - 1);
    if (bg->done) return;
    const unsigned N = (unsigned)F.N, logN = (unsigned)F.logN, m = N - 1, lane = threadIdx.x & 63;
    const unsigned w = threadIdx.x >> 6;
    const int par = bg->parity;
    const T *__restrict__ f = (const T *)F.f[par];
    double d = bg->d, laplnorm = bg->laplnorm, friction = bg->friction, srcamp = bg->srcamp;
    if (fresh) {
        const StepScalars S = bg->nx;
        d = S.d;
        laplnorm = S.laplnorm;
        friction = S.friction;
        srcamp = S.srcamp;
    }
    d = uniform(d);
    laplnorm = uniform(laplnorm);
    friction = uniform(friction);
    srcamp = uniform(srcamp);
    const double hx = 0.5 / bg->dx;
    // tile: XCD blk % 8 owns planes [x N/8, (x+1) N/8)
    const unsigned TI = tmarch_ti(N, E), jblocks = N / kTensorWaves, segs = N / (64u * E);
    const unsigned xcd = blk % kXcd, lb = blk / kXcd;
    const unsigned jb = lb % jblocks, t2 = lb / jblocks;
    RowWalk wk{};
    wk.segs = segs;
    wk.k0 = (t2 % segs) * (64u * E);
    const unsigned i0 = xcd * (N / kXcd) + (t2 / segs) * TI;
    const unsigned j0 = jb * kTensorWaves, j = j0 + w;
    const unsigned rj = (j << logN) + wk.k0;
    const unsigned row0 = j << logN;   // start of the whole row (segment edges)
    // waves 0 and 7 also bring the tile's halo rows j0-1 and j0+8 into LDS
    const bool halo = w == 0 || w == kTensorWaves - 1;
    const unsigned rh = ((w == 0 ? j0 - 1 : j0 + kTensorWaves) & m) << logN;
    const unsigned sh = w == 0 ? 0u : (unsigned)kTensorWaves + 1;
    const unsigned P2 = 2 * logN;
    T fdn[E], fcur[E], fup[E], fh[E];
    float hdn[6][E], hcur[6][E], hup[6][E], hv[6][E], hh[6][E];
    // PF: edge values of the plane-i rows (f and the six h), loaded a plane ahead
    T felo = (T)0, fehi = (T)0;
    float helo[6], hehi[6];
    {
        const unsigned pm = ((i0 - 1) & m) << P2, p0 = (i0 & m) << P2;
        ld_row<T, E>(f + pm + rj, lane, fdn);
        ld_row<T, E>(f + p0 + rj, lane, fcur);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            ld_row<float, E>(F.h[par][c] + pm + rj, lane, hdn[c]);
            ld_row<float, E>(F.h[par][c] + p0 + rj, lane, hcur[c]);
        }
        if constexpr (PF) {
            const unsigned p1 = ((i0 + 1) & m) << P2;
            ld_row<T, E>(f + p1 + rj, lane, fup);
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                ld_row<float, E>(F.h[par][c] + p1 + rj, lane, hup[c]);
                ld_row<float, E>(F.hd[c] + p0 + rj, lane, hv[c]);
            }
            if (halo) {
                ld_row<T, E>(f + p0 + rh + wk.k0, lane, fh);
#pragma unroll
                for (int c = 0; c < 6; ++c) ld_row<float, E>(F.h[par][c] + p0 + rh + wk.k0, lane, hh[c]);
            }
            if (segs > 1) {
                row_edges<T, E>(f + p0 + row0, wk.k0, m, felo, fehi);
#pragma unroll
                for (int c = 0; c < 6; ++c) row_edges<float, E>(F.h[par][c] + p0 + row0, wk.k0, m, helo[c], hehi[c]);
            }
        }
    }
    for (unsigned t = 0; t < TI; ++t) {
        const unsigned b = t & 1;
        const unsigned pi = ((i0 + t) & m) << P2, pu = ((i0 + t + 1) & m) << P2;
        if constexpr (!PF) {
            if (halo) {
                ld_row<T, E>(f + pi + rh + wk.k0, lane, fh);
#pragma unroll
                for (int c = 0; c < 6; ++c) ld_row<float, E>(F.h[par][c] + pi + rh + wk.k0, lane, hh[c]);
            }
            ld_row<T, E>(f + pu + rj, lane, fup);
        }
        // plane i rows into LDS: the wave's own, and the halo rows
#pragma unroll
        for (int e = 0; e < E; ++e) L.f[b][w + 1][lane * E + e] = fcur[e];
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int e = 0; e < E; ++e) L.h[b][c][w + 1][lane * E + e] = hcur[c][e];
        if (halo) {
#pragma unroll
            for (int e = 0; e < E; ++e) L.f[b][sh][lane * E + e] = fh[e];
#pragma unroll
            for (int c = 0; c < 6; ++c)
#pragma unroll
                for (int e = 0; e < E; ++e) L.h[b][c][sh][lane * E + e] = hh[c][e];
        }
        // PF: the next plane's loads go out before the barrier
        T fnx[E], fhn[E];
        float hnx[6][E], hvnx[6][E], hhn[6][E];
        T felon = (T)0, fehin = (T)0;
        float helon[6], hehin[6];
        if constexpr (PF) {
            if (t + 1 < TI) {
                const unsigned pn = ((i0 + t + 2) & m) << P2;
                ld_row<T, E>(f + pn + rj, lane, fnx);
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    ld_row<float, E>(F.h[par][c] + pn + rj, lane, hnx[c]);
                    ld_row<float, E>(F.hd[c] + pu + rj, lane, hvnx[c]);
                }
                if (halo) {
                    ld_row<T, E>(f + pu + rh + wk.k0, lane, fhn);
#pragma unroll
                    for (int c = 0; c < 6; ++c) ld_row<float, E>(F.h[par][c] + pu + rh + wk.k0, lane, hhn[c]);
                }
                if (segs > 1) {
                    row_edges<T, E>(f + pu + row0, wk.k0, m, felon, fehin);
#pragma unroll
                    for (int c = 0; c < 6; ++c)
                        row_edges<float, E>(F.h[par][c] + pu + row0, wk.k0, m, helon[c], hehin[c]);
                }
            }
        }
        __syncthreads();
        T fkp, fkm;
        if constexpr (PF) {
            fkp = __shfl(fcur[0], (int)((lane + 1) & 63));
            fkm = __shfl(fcur[E - 1], (int)((lane + 63) & 63));
            if (segs > 1) {
                fkm = lane == 0 ? felo : fkm;
                fkp = lane == 63 ? fehi : fkp;
            }
        } else {
            k_neighbours<T, E>(fcur, f + pi + row0, wk, N, lane, fkp, fkm);
        }
        double gx[E], gy[E], gz[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            // central gradient of phi (dfdx, src/evolution.cpp:55-77)
            const double kp = (double)(e < E - 1 ? fcur[e + 1 < E ? e + 1 : e] : fkp);
            const double km = (double)(e > 0 ? fcur[e > 0 ? e - 1 : 0] : fkm);
            gx[e] = ((double)fup[e] - (double)fdn[e]) * hx;
            gy[e] = ((double)L.f[b][w + 2][lane * E + e] - (double)L.f[b][w][lane * E + e]) * hx;
            gz[e] = (kp - km) * hx;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const float *__restrict__ h = F.h[par][c];
            if constexpr (!PF) {
                ld_row<float, E>(h + pu + rj, lane, hup[c]);
                ld_row<float, E>(F.hd[c] + pi + rj, lane, hv[c]);
            }
            float hkp, hkm;
            if constexpr (PF) {
                hkp = __shfl(hcur[c][0], (int)((lane + 1) & 63));
                hkm = __shfl(hcur[c][E - 1], (int)((lane + 63) & 63));
                if (segs > 1) {
                    hkm = lane == 0 ? helo[c] : hkm;
                    hkp = lane == 63 ? hehi[c] : hkp;
                }
            } else {
                k_neighbours<float, E>(hcur[c], h + pi + row0, wk, N, lane, hkp, hkm);
            }
            float hdn_new[E], hn[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const float kp = e < E - 1 ? hcur[c][e + 1 < E ? e + 1 : e] : hkp;
                const float km = e > 0 ? hcur[c][e > 0 ? e - 1 : 0] : hkm;
                const float jp = L.h[b][c][w + 2][lane * E + e], jm = L.h[b][c][w][lane * E + e];
                const float hl = kp + km + jp + jm + hup[c][e] + hdn[c][e] - 6.0f * hcur[c][e];
                const double src = c == 0 ? gx[e] * gx[e]
                                 : c == 1 ? gy[e] * gy[e]
                                 : c == 2 ? gz[e] * gz[e]
                                 : c == 3 ? gx[e] * gy[e]
                                 : c == 4 ? gx[e] * gz[e]
                                          : gy[e] * gz[e];
                const double rhs = d * (laplnorm * (double)hl - friction * (double)hv[c][e] + srcamp * src);
                hdn_new[e] = hv[c][e] + (float)rhs;
                hn[e] = hcur[c][e] + (float)(d * (double)hdn_new[e]);
            }
            st_row<float, E>(F.hd[c] + pi + rj, lane, hdn_new);
            st_row<float, E>(F.h[par ^ 1][c] + pi + rj, lane, hn);
        }
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                hdn[c][e] = hcur[c][e];
                hcur[c][e] = hup[c][e];
            }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            fdn[e] = fcur[e];
            fcur[e] = fup[e];
        }
        if constexpr (PF) {
            if (t + 1 < TI) {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    fup[e] = fnx[e];
                    fh[e] = fhn[e];
                }
#pragma unroll
                for (int c = 0; c < 6; ++c) {
#pragma unroll
                    for (int e = 0; e < E; ++e) {
                        hup[c][e] = hnx[c][e];
                        hv[c][e] = hvnx[c][e];
                        hh[c][e] = hhn[c][e];
                    }
                    helo[c] = helon[c];
                    hehi[c] = hehin[c];
                }
                felo = felon;
                fehi = fehin;
            }
        }
    }
}

template <typename T, int E, bool PF>
__global__ __launch_bounds__(64 * kTensorWaves)
void k_tensor_march(Fields F, DevBg *bg, int fresh)
{
    __shared__ TensorTile<T, E> L;
    tensor_march_body<T, E, PF>(F, bg, fresh, blockIdx.x, L);
}

template <typename T, int E, bool PF>
__global__ __launch_bounds__(64 * kTensorWaves)
void kb_tensor_march(const RunDesc *__restrict__ runs, const uint2 *__restrict__ map, unsigned rslot)
{
    __shared__ TensorTile<T, E> L;
    const uint2 mp = map[blockIdx.x];
    const RunDesc &R = runs[mp.x];
    tensor_march_body<T, E, PF>(R.F, R.bg, R.march ? 1 + (int)rslot : 0, mp.y, L);
}

// ------------------------------------------------- small-lattice site kernels

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

template <typename T, int U>
__device__ __forceinline__ double stencil_lapl(const Stencil<T, U> &s, int u)
{
    return lapl_sites((double)s.kp[u], (double)s.km[u], (double)s.jp[u], (double)s.jm[u], (double)s.ip[u],
                      (double)s.im[u], (double)s.c[u]);
}

constexpr int kUnroll = 4;

template <typename T, bool SPLIT>
__device__ __forceinline__ void site_reduce_body(const Fields &F, const Pot &P0, DevBg *bg, double *partials, int loop,
                                                 unsigned G, unsigned blk, double *smem)
{
    if (loop && bg->done) return;
    const PotView Pv = pot_view(P0);
    const unsigned N = (unsigned)F.N, logN = (unsigned)F.logN;
    const unsigned gs = N * N * N;
    const T *__restrict__ f = (const T *)F.f[bg->parity];
    const double phibar = SPLIT ? bg->phibar : 0.0;
    double acc[3] = {0.0, 0.0, 0.0};
    int err = 0;
    const Walk32 w = site_walk32(gs, G, blk);
    for (unsigned base = w.begin; base < w.end; base += (unsigned)kUnroll * w.stride) {
        Stencil<T, kUnroll> s;
        load_stencil<T, kUnroll>(f, w, base, N, logN, s);
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if (!s.ok[u]) continue;
            const double c = (double)s.c[u];
            const double lap = stencil_lapl<T, kUnroll>(s, u);
            acc[0] -= c * lap;
            double V, dV;
            int h = Pv.start(phibar + c);
            Pv.eval1(phibar + c, &V, &dV, h, &err);
            acc[1] += V;
            if (SPLIT) acc[2] += dV;
        }
    }
    flag_error(bg, err);
    block_sum<3>(acc, smem);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) partials[(size_t)blk * 3 + q] = acc[q];
}

template <typename T, bool SPLIT, int MODE>
__device__ __forceinline__ void site_kick_body(const Fields &F, const Pot &P0, DevBg *bg, double d_arg, int loop,
                                               unsigned G, unsigned blk)
{
    if (loop && bg->done) return;
    const PotView Pv = pot_view(P0);
    const unsigned N = (unsigned)F.N, logN = (unsigned)F.logN;
    const unsigned gs = N * N * N;
    const int par = bg->parity;
    const T *__restrict__ f = (const T *)F.f[par];
    T *__restrict__ fout = (T *)F.f[par ^ 1];
    const T *fd = (const T *)F.fd[par];
    T *fdout = (T *)F.fd[MODE == KICKDRIFT ? par ^ 1 : par];
    const double d = (MODE == DRIFT && !loop) ? d_arg : bg->d;
    const double laplnorm = bg->laplnorm, friction = bg->friction, potnorm = bg->potnorm;
    const double phibar = SPLIT ? bg->phibar : 0.0, vpmean = SPLIT ? bg->vpmean : 0.0;
    int err = 0;
    if (MODE & KICK) {
        const Walk32 w = site_walk32(gs, G, blk);
        for (unsigned base = w.begin; base < w.end; base += (unsigned)kUnroll * w.stride) {
            Stencil<T, kUnroll> s;
            load_stencil<T, kUnroll>(f, w, base, N, logN, s);
            T vin[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) vin[u] = fd[s.id[u]];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                if (!s.ok[u]) continue;
                const double c = (double)s.c[u];
                const double lap = stencil_lapl<T, kUnroll>(s, u);
                int h = Pv.start(SPLIT ? phibar + c : c);
                double dV = Pv.dV1(SPLIT ? phibar + c : c, h, &err);
                if (SPLIT) dV -= vpmean;
                double v = (double)vin[u];
                v = v + d * (laplnorm * lap - friction * v - potnorm * dV);
                const T vs = (T)v;
                fdout[s.id[u]] = vs;
                if (MODE & DRIFT) fout[s.id[u]] = (T)(c + d * (double)vs);
            }
        }
    } else {
        // drift alone, in place (pointwise): through one unrestricted pointer
        T *g = (T *)F.f[par];
        const Walk w = site_walk((size_t)gs, G, blk);
        for (size_t id = w.begin; id < w.end; id += w.stride) g[id] = (T)((double)g[id] + d * (double)fd[id]);
    }
    flag_error(bg, err);
}

template <typename T, bool SPLIT>
__global__ __launch_bounds__(kBlock) void k_site_reduce(Fields F, Pot P, DevBg *bg, double *partials, int loop)
{
    __shared__ double smem[3 * kWaves];
    site_reduce_body<T, SPLIT>(F, P, bg, partials, loop, gridDim.x, blockIdx.x, smem);
}

template <typename T, bool SPLIT, int MODE>
__global__ __launch_bounds__(kBlock) void k_site_kick(Fields F, Pot P, DevBg *bg, double d_arg, int loop)
{
    site_kick_body<T, SPLIT, MODE>(F, P, bg, d_arg, loop, gridDim.x, blockIdx.x);
}

template <typename T, bool SPLIT>
__global__ __launch_bounds__(kBlock) void kb_site_reduce(const RunDesc *__restrict__ runs,
                                                          const uint2 *__restrict__ map)
{
    __shared__ double smem[3 * kWaves];
    const uint2 mp = map[blockIdx.x];
    const RunDesc &R = runs[mp.x];
    site_reduce_body<T, SPLIT>(R.F, R.P, R.bg, R.partials, 1, R.G, mp.y, smem);
}

template <typename T, bool SPLIT>
__global__ __launch_bounds__(kBlock) void kb_site_kick(const RunDesc *__restrict__ runs,
                                                        const uint2 *__restrict__ map)
{
    const uint2 mp = map[blockIdx.x];
    const RunDesc &R = runs[mp.x];
    site_kick_body<T, SPLIT, KICKDRIFT>(R.F, R.P, R.bg, 0.0, 1, R.G, mp.y);
}

// ------------------------------------------------------------ tensor kernel

template <typename T, int MODE>
__device__ __forceinline__ void tensor_body(const Fields &F, DevBg *bg, double d_arg, int loop, int fresh, unsigned G,
                                            unsigned blk)
{
    if (fresh) bg = F.rec + (fresh - 1);
    if (loop && bg->done) return;
    const size_t gs = (size_t)F.N * F.N * F.N;
    const int par = bg->parity;
    const T *__restrict__ f = (const T *)F.f[par];
    double d = (MODE == DRIFT && !loop) ? d_arg : bg->d;
    double laplnorm = bg->laplnorm, friction = bg->friction, srcamp = bg->srcamp;
    if (fresh) {
        const StepScalars S = bg->nx;
        d = S.d;
        laplnorm = S.laplnorm;
        friction = S.friction;
        srcamp = S.srcamp;
    }
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
}

template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_tensor(Fields F, DevBg *bg, double d_arg, int loop, int fresh)
{
    tensor_body<T, MODE>(F, bg, d_arg, loop, fresh, gridDim.x, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void kb_tensor(const RunDesc *__restrict__ runs, const uint2 *__restrict__ map,
                                                     unsigned rslot)
{
    const uint2 mp = map[blockIdx.x];
    const RunDesc &R = runs[mp.x];
    tensor_body<T, KICKDRIFT>(R.F, R.bg, 0.0, 1, R.march ? 1 + (int)rslot : 0, R.G, mp.y);
}

// ------------------------------------------------------- output utilities

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
    const T *__restrict__ fd = (const T *)F.fd[bg->parity];
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
    const T *__restrict__ src = which == 0 ? (const T *)F.f[bg->parity] : (const T *)F.fd[bg->parity];
    const double base = SPLIT ? (which == 0 ? bg->phibar : bg->pibar) : 0.0;
    const Walk w = site_walk(gs);
    for (size_t id = w.begin; id < w.end; id += w.stride) out[id] = SPLIT ? base + (double)src[id] : (double)src[id];
}

template <typename T, bool SPLIT>
void reduce_dispatch(const Fields &F, const Pot &P, DevBg *bg, double *partials, int loop, unsigned G, hipStream_t s)
{
    switch (row_elems(F.N)) {
    case 1: k_row_reduce<T, SPLIT, 1><<<G, kBlock, 0, s>>>(F, P, bg, partials, loop); break;
    case 2: k_row_reduce<T, SPLIT, 2><<<G, kBlock, 0, s>>>(F, P, bg, partials, loop); break;
    case 4: k_row_reduce<T, SPLIT, 4><<<G, kBlock, 0, s>>>(F, P, bg, partials, loop); break;
    default: k_site_reduce<T, SPLIT><<<G, kBlock, 0, s>>>(F, P, bg, partials, loop); break;
    }
}

template <typename T, bool SPLIT, int MODE>
void row_kick_dispatch(const Fields &F, const Pot &P, DevBg *bg, int loop, unsigned G, hipStream_t s)
{
    switch (shape_elems(F.N, F.kick_e)) {
    case 1: k_row_kick<T, SPLIT, 1, MODE><<<G, kBlock, 0, s>>>(F, P, bg, loop); break;
    case 2: k_row_kick<T, SPLIT, 2, MODE><<<G, kBlock, 0, s>>>(F, P, bg, loop); break;
    default: k_row_kick<T, SPLIT, 4, MODE><<<G, kBlock, 0, s>>>(F, P, bg, loop); break;
    }
}

template <typename T, bool SPLIT>
void kick_dispatch(const Fields &F, const Pot &P, DevBg *bg, int mode, double d_arg, int loop, unsigned G,
                   hipStream_t s)
{
    if (mode == DRIFT) {
        k_site_kick<T, SPLIT, DRIFT><<<G, kBlock, 0, s>>>(F, P, bg, d_arg, loop);
    } else if (row_elems(F.N) == 0) {
        if (mode == KICK) k_site_kick<T, SPLIT, KICK><<<G, kBlock, 0, s>>>(F, P, bg, d_arg, loop);
        else k_site_kick<T, SPLIT, KICKDRIFT><<<G, kBlock, 0, s>>>(F, P, bg, d_arg, loop);
    } else {
        if (mode == KICK) row_kick_dispatch<T, SPLIT, KICK>(F, P, bg, loop, G, s);
        else row_kick_dispatch<T, SPLIT, KICKDRIFT>(F, P, bg, loop, G, s);
    }
}

template <typename T, int MODE>
void row_tensor_dispatch(const Fields &F, DevBg *bg, int loop, int fresh, unsigned G, hipStream_t s)
{
    switch (shape_elems(F.N, F.tensor_e & 7)) {
    case 1: k_row_tensor<T, 1, MODE><<<G, kBlock, 0, s>>>(F, bg, loop, fresh); break;
    case 2: k_row_tensor<T, 2, MODE><<<G, kBlock, 0, s>>>(F, bg, loop, fresh); break;
    default: k_row_tensor<T, 4, MODE><<<G, kBlock, 0, s>>>(F, bg, loop, fresh); break;
    }
}

template <typename T>
void tensor_dispatch(const Fields &F, DevBg *bg, int mode, double d_arg, int loop, int fresh, unsigned G,
                     hipStream_t s)
{
    if (mode == DRIFT) k_tensor<T, DRIFT><<<G, kBlock, 0, s>>>(F, bg, d_arg, loop, 0);
    else if (mode == KICKDRIFT && loop && row_elems(F.N) > 0 && tensor_is_march(F.tensor_e)) {
        const int E = tensor_march_elems(F.N, F.tensor_e, sizeof(T) == 4);
        const unsigned Gt = tensor_march_blocks(F.N, E);
        const unsigned nt = 64 * kTensorWaves;
        if (F.tensor_e & kTensorPf) {
            if (E == 1) k_tensor_march<T, 1, true><<<Gt, nt, 0, s>>>(F, bg, fresh);
            else if (E == 2) k_tensor_march<T, 2, true><<<Gt, nt, 0, s>>>(F, bg, fresh);
            else if constexpr (sizeof(T) == 4) k_tensor_march<T, 4, true><<<Gt, nt, 0, s>>>(F, bg, fresh);
        } else {
            if (E == 1) k_tensor_march<T, 1, false><<<Gt, nt, 0, s>>>(F, bg, fresh);
            else if (E == 2) k_tensor_march<T, 2, false><<<Gt, nt, 0, s>>>(F, bg, fresh);
            else if constexpr (sizeof(T) == 4) k_tensor_march<T, 4, false><<<Gt, nt, 0, s>>>(F, bg, fresh);
        }
    } else if (row_elems(F.N) == 0) {
        if (mode == KICK) k_tensor<T, KICK><<<G, kBlock, 0, s>>>(F, bg, d_arg, loop, fresh);
        else k_tensor<T, KICKDRIFT><<<G, kBlock, 0, s>>>(F, bg, d_arg, loop, fresh);
    } else {
        if (mode == KICK) row_tensor_dispatch<T, KICK>(F, bg, loop, fresh, G, s);
        else row_tensor_dispatch<T, KICKDRIFT>(F, bg, loop, fresh, G, s);
    }
}

template <typename T, bool SPLIT, int POT>
void march_dispatch_pot(const Fields &F, const Pot &P, DevBg *bg, double *partials, unsigned G, unsigned sp,
                        hipStream_t s)
{
    const int E = row_elems(F.N);
    if (F.march_pf) {
        if (E == 2) k_march_pf<T, SPLIT, 2, POT><<<G, kMarchThreads, march_lds<POT, false, 2>(), s>>>(F, P, bg, partials, sp);
        else k_march_pf<T, SPLIT, 4, POT><<<G, kMarchThreads, march_lds<POT, false, 4>(), s>>>(F, P, bg, partials, sp);
    } else {
        if (E == 2) k_march<T, SPLIT, 2, POT><<<G, kMarchThreads, march_lds<POT, true, 2>(), s>>>(F, P, bg, partials, sp);
        else k_march<T, SPLIT, 4, POT><<<G, kMarchThreads, march_lds<POT, true, 4>(), s>>>(F, P, bg, partials, sp);
    }
}

template <typename T, bool SPLIT>
void march_dispatch(const Fields &F, const Pot &P, DevBg *bg, double *partials, unsigned G, unsigned sp, hipStream_t s)
{
    if (P.numerical) march_dispatch_pot<T, SPLIT, kPotTab>(F, P, bg, partials, G, sp, s);
    else march_dispatch_pot<T, SPLIT, kPotHill>(F, P, bg, partials, G, sp, s);
}

template <typename T, bool SPLIT, int POT>
void march_sums_dispatch(const Fields &F, const Pot &P, DevBg *bg, double *partials, unsigned G, hipStream_t s)
{
    if (row_elems(F.N) == 2) k_march_sums<T, SPLIT, 2, POT><<<G, kMarchThreads, march_lds<POT, false, 2>(), s>>>(F, P, bg, partials);
    else k_march_sums<T, SPLIT, 4, POT><<<G, kMarchThreads, march_lds<POT, false, 4>(), s>>>(F, P, bg, partials);
}

}  // namespace

// Elements per lane of the row kernels, 0 for the site kernels (N < 64).
int row_elems(int N) { return N < 64 ? 0 : (N >= 256 ? 4 : N / 64); }

// Elements per lane of a kick or tensor launch: `want` (1, 2 or 4; 0 = the
// reduce's) within what one row holds.  These kernels have no cross-site sums,
// so their shape never changes a result.
int shape_elems(int N, int want)
{
    const int def = row_elems(N);
    if (def == 0 || want <= 0) return def;
    const int e = want >= 4 ? 4 : (want >= 2 ? 2 : 1);
    return e <= def ? e : def;
}

// Tensor step shape codes: 0 default (row kernel), 1/2/4 row kernel with that many
// sites per lane, kTensorMarch | [kTensorPf] | E the march kernel with E sites per
// lane (E = 4 with FP32 f only: the LDS tile of FP64 rows would not fit).
bool tensor_is_march(int code) { return (code & kTensorMarch) != 0; }

int tensor_march_elems(int N, int code, int split)
{
    const int e = shape_elems(N, code & 7);
    return e > 2 && !split ? 2 : e;
}

unsigned tensor_march_blocks(int N, int E)
{
    const unsigned segs = (unsigned)N / (64u * (unsigned)E);
    return (unsigned)(N / kTensorWaves) * segs * ((unsigned)N / tmarch_ti((unsigned)N, E));
}

int reduce_blocks(size_t gs)
{
    size_t b = gs / kBlock;
    if (b > 2048) b = 2048;
    if (b >= kXcd) b -= b % kXcd;
    return (int)(b < 1 ? 1 : b);
}

unsigned step_blocks(int N)
{
    const int E = row_elems(N);
    if (E == 0) {
        size_t b = (size_t)N * N * N / kBlock;
        b = (b + kXcd - 1) / kXcd * kXcd;
        return (unsigned)(b < kXcd ? kXcd : (b > 2048 ? 2048 : b));
    }
    // row kernels: N^2 * segments / 32 workgroups (8 rows per wave), at most 2048
    const size_t segs = (size_t)N / (64 * (size_t)E);
    const size_t b = (size_t)N * N * segs / 32;
    return (unsigned)(b > 2048 ? 2048 : b);
}

hipError_t launch_reduce(const Fields &F, int split, const Pot &P, DevBg *bg, double *partials, int loop,
                         hipStream_t s)
{
    const unsigned G = step_blocks(F.N);
    if (split) reduce_dispatch<float, true>(F, P, bg, partials, loop, G, s);
    else reduce_dispatch<double, false>(F, P, bg, partials, loop, G, s);
    return hipGetLastError();
}

hipError_t launch_friedmann_finish(const Fields &F, DevBg *bg, const double *partials, double d_arg, int loop,
                                   hipStream_t s)
{
    const size_t gs = (size_t)F.N * F.N * F.N;
    k_friedmann_finish<<<1, kBlock, 0, s>>>(bg, partials, step_blocks(F.N), (unsigned long long)gs, d_arg, loop);
    return hipGetLastError();
}

hipError_t launch_scalar_step(const Fields &F, int split, const Pot &P, DevBg *bg, int mode, double d_arg, int loop,
                              hipStream_t s)
{
    const unsigned G = step_blocks(F.N);
    if (split) kick_dispatch<float, true>(F, P, bg, mode, d_arg, loop, G, s);
    else kick_dispatch<double, false>(F, P, bg, mode, d_arg, loop, G, s);
    return hipGetLastError();
}

hipError_t launch_tensor_step(const Fields &F, int split, DevBg *bg, int mode, double d_arg, int loop, hipStream_t s,
                              int fresh)
{
    const unsigned G = step_blocks(F.N);
    if (split) tensor_dispatch<float>(F, bg, mode, d_arg, loop, fresh, G, s);
    else tensor_dispatch<double>(F, bg, mode, d_arg, loop, fresh, G, s);
    return hipGetLastError();
}

// Fused steps for 128 <= N <= 256: at N = 64 a row is one site per lane and the
// tile's halo row and per-plane barrier cost more than the reduce they save (a
// batch of 256 lattices of 64^3 steps faster unfused); above 256 a row spans
// several segments, whose k pairs would cross waves.
// (the march kernels step TI + 1 planes three at a time)
bool march_ok(int N) { return N >= 128 && N <= 256 && (march_ti((unsigned)N) + 1) % 3 == 0; }

bool resident_ok(int N, int sigw, int split, int numerical)
{
    return N == kResN && !sigw && (split || !numerical);
}

// Minimum number of resident-capable lattices of a potential kind for a batch
// to take the resident path (below it the per-step kernels spread each lattice
// over the GPU and step faster).
static int g_res_min[2] = {kResMinHill, kResMinTab};

void set_resident_min(int hill, int tab)
{
    g_res_min[0] = hill > 0 ? hill : kResMinHill;
    g_res_min[1] = tab > 0 ? tab : kResMinTab;
}

unsigned march_blocks(int N) { return (unsigned)((N / kTileJ) * (N / (int)march_ti((unsigned)N))); }

hipError_t launch_march_begin(const Fields &F, int split, const Pot &P, DevBg *bg, double *partials, hipStream_t s)
{
    const unsigned G = march_blocks(F.N);
    if (split && P.numerical) march_sums_dispatch<float, true, kPotTab>(F, P, bg, partials, G, s);
    else if (split) march_sums_dispatch<float, true, kPotHill>(F, P, bg, partials, G, s);
    else if (P.numerical) march_sums_dispatch<double, false, kPotTab>(F, P, bg, partials, G, s);
    else march_sums_dispatch<double, false, kPotHill>(F, P, bg, partials, G, s);
    k_sums_finish<<<1, kBlock, 0, s>>>(bg, partials, G, (unsigned long long)F.N * F.N * F.N, F.rec);
    return hipGetLastError();
}

hipError_t launch_march(const Fields &F, int split, const Pot &P, DevBg *bg, double *partials, long long n,
                        hipStream_t s)
{
    const unsigned G = march_blocks(F.N);
    if (split) march_dispatch<float, true>(F, P, bg, partials, G, (unsigned)(n & 1), s);
    else march_dispatch<double, false>(F, P, bg, partials, G, (unsigned)(n & 1), s);
    return hipGetLastError();
}

int step_record(long long n) { return 1 + (int)((n + 1) & 1); }

hipError_t launch_loop_end(const Fields &F, int split, DevBg *bg, const double *partials, long long steps,
                           hipStream_t s)
{
    const size_t gs = (size_t)F.N * F.N * F.N;
    k_loop_end<<<1, kBlock, 0, s>>>(bg, F.rec, (unsigned)(steps & 1), partials, march_blocks(F.N),
                                    (unsigned long long)gs, split);
    return hipGetLastError();
}

hipError_t launch_fused_step(const Fields &F, int split, const Pot &P, DevBg *bg, double *partials, long long n,
                             hipStream_t s)
{
    hipError_t e = launch_march(F, split, P, bg, partials, n, s);
    if (e == hipSuccess && F.sigw) e = launch_tensor_step(F, split, bg, KICKDRIFT, 0.0, 1, s, step_record(n));
    return e;
}

hipError_t launch_step_finish(DevBg *bg, int split, int mode, double d_arg, int loop, hipStream_t s)
{
    k_step_finish<<<1, 64, 0, s>>>(bg, mode, split, d_arg, loop);
    return hipGetLastError();
}

hipError_t launch_friedmann_reduce(const Fields &F, int split, const Pot &P, DevBg *bg, double *partials,
                                   double d_arg, int loop, hipStream_t s)
{
    hipError_t e = launch_reduce(F, split, P, bg, partials, loop, s);
    if (e != hipSuccess) return e;
    return launch_friedmann_finish(F, bg, partials, d_arg, loop, s);
}

hipError_t launch_kick_drift(const Fields &F, int split, const Pot &P, DevBg *bg, int mode, double d_arg,
                             int loop, hipStream_t s, int part)
{
    hipError_t e = hipSuccess;
    if (part != 1) e = launch_scalar_step(F, split, P, bg, mode, d_arg, loop, s);
    if (e == hipSuccess && F.sigw && part != 0) e = launch_tensor_step(F, split, bg, mode, d_arg, loop, s);
    if (e == hipSuccess && (part == -1 || (part == 0 && !F.sigw) || part == 1))
        e = launch_step_finish(bg, split, mode, d_arg, loop, s);
    return e;
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
    k_split_store<<<G, kBlock, 0, s>>>(f, fd, gs, bg, (float *)F.f[0], (float *)F.fd[0]);
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

// Runs of one storage kind, grouped by kernel shape (row elements per lane); one
// launch per group and kernel, one finish launch over all runs.
struct BatchGroup {
    int E = 0, kE = 0, tE = 0;   // elements per lane: reduce, kick, tensor
    int march = 0;               // fused steps (march kernel) for this group
    int pf = 0;                  // the march kernel variant loading one plane step ahead
    int pot = 0;                 // fused groups: 1 tabulated potential, 0 hilltop
    int res = 0;                 // resident 64^3 lattices (kb_resident: every step of a launch at once)
    uint2 *map = nullptr;        // kick / march workgroups
    uint2 *rmap = nullptr;       // reduce workgroups (march groups: loop start only)
    uint2 *tmap = nullptr;       // tensor workgroups
    int nblocks = 0, nrblocks = 0, ntblocks = 0;
};

struct BatchDev {
    RunDesc *runs = nullptr;
    std::vector<BatchGroup> groups;
    int nruns = 0, split = 0, has_march = 0, has_plain = 0, has_res = 0;
};

hipError_t batch_build(const BatchRun *runs, int n, int split, BatchDev **out)
{
    *out = nullptr;
    std::vector<RunDesc> d(n);
    BatchDev *B = new BatchDev();
    // one resident workgroup per 64^3 lattice without tensors (FP64 tabulated
    // lattices excepted: their rows and table exceed the workgroup's LDS)
    B->nruns = n;
    B->split = split;
    // resident only with enough lattices of the kind to fill the GPU
    int res_count[2] = {0, 0};
    for (int r = 0; r < n; ++r)
        if (resident_ok(runs[r].F.N, runs[r].F.sigw, split, runs[r].P.numerical))
            ++res_count[runs[r].P.numerical ? 1 : 0];
    // groups of runs launched together: same step structure and kernel shapes
    std::vector<int> keys;
    std::vector<std::vector<uint2>> maps, rmaps, tmaps;
    for (int r = 0; r < n; ++r) {
        const int N = runs[r].F.N;
        d[r].F = runs[r].F;
        d[r].P = runs[r].P;
        d[r].bg = runs[r].bg;
        d[r].partials = runs[r].partials;
        d[r].G = step_blocks(N);
        d[r].march = march_ok(N) ? 1 : 0;
        d[r].Gm = d[r].march ? march_blocks(N) : 0;
        d[r].split = split;
        const int kind = runs[r].P.numerical ? 1 : 0;
        d[r].res = resident_ok(N, runs[r].F.sigw, split, runs[r].P.numerical) && res_count[kind] >= g_res_min[kind]
                       ? 1 : 0;
        if (d[r].res) {
            B->has_res = 1;
            const int key = 100000000 + runs[r].P.numerical * 10000000;
            size_t g = 0;
            while (g < keys.size() && keys[g] != key) ++g;
            if (g == keys.size()) {
                keys.push_back(key);
                maps.emplace_back();
                rmaps.emplace_back();
                tmaps.emplace_back();
            }
            maps[g].push_back(make_uint2((unsigned)r, 0u));
            continue;
        }
        (d[r].march ? B->has_march : B->has_plain) = 1;
        const int kE = d[r].march ? row_elems(N) : shape_elems(N, runs[r].F.kick_e);
        const int tc = runs[r].F.tensor_e;
        const int tE = row_elems(N) == 0 ? 0
                       : tensor_is_march(tc) ? kTensorMarch | (tc & kTensorPf) | tensor_march_elems(N, tc, split)
                                             : shape_elems(N, tc);
        const int pf = d[r].march && runs[r].F.march_pf ? 1 : 0;
        const int pot = d[r].march && runs[r].P.numerical ? 1 : 0;   // fused kernels are per potential kind
        const int key = pot * 10000000 + pf * 1000000 + d[r].march * 100000 + row_elems(N) * 10000 + kE * 100 + tE;
        size_t g = 0;
        while (g < keys.size() && keys[g] != key) ++g;
        if (g == keys.size()) {
            keys.push_back(key);
            maps.emplace_back();
            rmaps.emplace_back();
            tmaps.emplace_back();
        }
        for (unsigned b = 0; b < d[r].G; ++b) rmaps[g].push_back(make_uint2((unsigned)r, b));
        if (runs[r].F.sigw) {
            // tensor step: march kernel (N >= 64) or site kernel
            const unsigned Gt = (row_elems(N) > 0 && tensor_is_march(tc))
                                    ? tensor_march_blocks(N, tensor_march_elems(N, tc, split))
                                    : d[r].G;
            for (unsigned b = 0; b < Gt; ++b) tmaps[g].push_back(make_uint2((unsigned)r, b));
        }
        const unsigned Gk = d[r].march ? d[r].Gm : d[r].G;
        for (unsigned b = 0; b < Gk; ++b) maps[g].push_back(make_uint2((unsigned)r, b));
    }
    hipError_t e = hipMalloc(&B->runs, sizeof(RunDesc) * n);
    if (e == hipSuccess) e = hipMemcpy(B->runs, d.data(), sizeof(RunDesc) * n, hipMemcpyHostToDevice);
    auto upload = [&](const std::vector<uint2> &v, uint2 **p) {
        if (e != hipSuccess || v.empty()) return;
        e = hipMalloc(p, sizeof(uint2) * v.size());
        if (e == hipSuccess) e = hipMemcpy(*p, v.data(), sizeof(uint2) * v.size(), hipMemcpyHostToDevice);
    };
    for (size_t g = 0; g < keys.size() && e == hipSuccess; ++g) {
        BatchGroup G;
        G.res = keys[g] / 100000000;
        G.pot = keys[g] / 10000000 % 10;
        G.pf = keys[g] / 1000000 % 10;
        G.march = keys[g] / 100000 % 10;
        G.E = keys[g] / 10000 % 10;
        G.kE = keys[g] / 100 % 100;
        G.tE = keys[g] % 100;
        G.nblocks = (int)maps[g].size();
        G.nrblocks = (int)rmaps[g].size();
        G.ntblocks = (int)tmaps[g].size();
        upload(maps[g], &G.map);
        upload(rmaps[g], &G.rmap);
        upload(tmaps[g], &G.tmap);
        B->groups.push_back(G);
    }
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
    for (BatchGroup &G : B->groups) {
        if (G.map) (void)hipFree(G.map);
        if (G.rmap) (void)hipFree(G.rmap);
        if (G.tmap) (void)hipFree(G.tmap);
    }
    delete B;
}

template <typename T, bool SPLIT>
static void batch_group_reduce(const BatchDev *B, const BatchGroup &G, hipStream_t s);

template <typename T, bool SPLIT, int POT>
static void batch_march_sums(const BatchDev *B, const BatchGroup &G, hipStream_t s)
{
    if (G.E == 2) kb_march_sums<T, SPLIT, 2, POT><<<G.nblocks, kMarchThreads, march_lds<POT, false, 2>(), s>>>(B->runs, G.map);
    else kb_march_sums<T, SPLIT, 4, POT><<<G.nblocks, kMarchThreads, march_lds<POT, false, 4>(), s>>>(B->runs, G.map);
}

template <typename T, bool SPLIT, int POT>
static void batch_march(const BatchDev *B, const BatchGroup &G, unsigned sp, hipStream_t s)
{
    // (march groups have N = 128 or 256: E = 2 or 4)
    if (G.pf) {
        if (G.E == 2) kb_march_pf<T, SPLIT, 2, POT><<<G.nblocks, kMarchThreads, march_lds<POT, false, 2>(), s>>>(B->runs, G.map, sp);
        else kb_march_pf<T, SPLIT, 4, POT><<<G.nblocks, kMarchThreads, march_lds<POT, false, 4>(), s>>>(B->runs, G.map, sp);
    } else {
        if (G.E == 2) kb_march<T, SPLIT, 2, POT><<<G.nblocks, kMarchThreads, march_lds<POT, true, 2>(), s>>>(B->runs, G.map, sp);
        else kb_march<T, SPLIT, 4, POT><<<G.nblocks, kMarchThreads, march_lds<POT, true, 4>(), s>>>(B->runs, G.map, sp);
    }
}

hipError_t batch_loop_begin(const BatchDev *B, hipStream_t s)
{
    kb_loop_begin<<<(B->nruns + 63) / 64, 64, 0, s>>>(B->runs, B->nruns);
    if (B->has_march) {
        // fused groups start from the sums of the current f, in a step's order
        for (const BatchGroup &G : B->groups) {
            if (!G.march) continue;
            if (B->split && G.pot) batch_march_sums<float, true, kPotTab>(B, G, s);
            else if (B->split) batch_march_sums<float, true, kPotHill>(B, G, s);
            else if (G.pot) batch_march_sums<double, false, kPotTab>(B, G, s);
            else batch_march_sums<double, false, kPotHill>(B, G, s);
        }
        kb_sums_finish<<<B->nruns, kBlock, 0, s>>>(B->runs);
    }
    return hipGetLastError();
}

template <typename T, bool SPLIT>
static void batch_group_reduce(const BatchDev *B, const BatchGroup &G, hipStream_t s)
{
    switch (G.E) {
    case 1: kb_row_reduce<T, SPLIT, 1><<<G.nrblocks, kBlock, 0, s>>>(B->runs, G.rmap); break;
    case 2: kb_row_reduce<T, SPLIT, 2><<<G.nrblocks, kBlock, 0, s>>>(B->runs, G.rmap); break;
    case 4: kb_row_reduce<T, SPLIT, 4><<<G.nrblocks, kBlock, 0, s>>>(B->runs, G.rmap); break;
    default: kb_site_reduce<T, SPLIT><<<G.nrblocks, kBlock, 0, s>>>(B->runs, G.rmap); break;
    }
}

template <typename T, bool SPLIT>
static void batch_group_kick(const BatchDev *B, const BatchGroup &G, long long n, hipStream_t s)
{
    // records: the previous step's in slot n & 1, this step's in the other
    const unsigned sp = (unsigned)(n & 1), rs = sp ^ 1u;
    if (G.march) {
        if (G.pot) batch_march<T, SPLIT, kPotTab>(B, G, sp, s);
        else batch_march<T, SPLIT, kPotHill>(B, G, sp, s);
    } else switch (G.kE) {
    case 1: kb_row_kick<T, SPLIT, 1><<<G.nblocks, kBlock, 0, s>>>(B->runs, G.map); break;
    case 2: kb_row_kick<T, SPLIT, 2><<<G.nblocks, kBlock, 0, s>>>(B->runs, G.map); break;
    case 4: kb_row_kick<T, SPLIT, 4><<<G.nblocks, kBlock, 0, s>>>(B->runs, G.map); break;
    default: kb_site_kick<T, SPLIT><<<G.nblocks, kBlock, 0, s>>>(B->runs, G.map); break;
    }
    if (!G.ntblocks) return;
    if (G.E == 0) kb_tensor<T><<<G.ntblocks, kBlock, 0, s>>>(B->runs, G.tmap, rs);
    else if (tensor_is_march(G.tE)) {
        const int E = G.tE & 7;
        const unsigned nt = 64 * kTensorWaves;
        if (G.tE & kTensorPf) {
            if (E == 1) kb_tensor_march<T, 1, true><<<G.ntblocks, nt, 0, s>>>(B->runs, G.tmap, rs);
            else if (E == 2) kb_tensor_march<T, 2, true><<<G.ntblocks, nt, 0, s>>>(B->runs, G.tmap, rs);
            else if constexpr (sizeof(T) == 4) kb_tensor_march<T, 4, true><<<G.ntblocks, nt, 0, s>>>(B->runs, G.tmap, rs);
        } else {
            if (E == 1) kb_tensor_march<T, 1, false><<<G.ntblocks, nt, 0, s>>>(B->runs, G.tmap, rs);
            else if (E == 2) kb_tensor_march<T, 2, false><<<G.ntblocks, nt, 0, s>>>(B->runs, G.tmap, rs);
            else if constexpr (sizeof(T) == 4) kb_tensor_march<T, 4, false><<<G.ntblocks, nt, 0, s>>>(B->runs, G.tmap, rs);
        }
    }
    else if (G.tE == 1) kb_row_tensor<T, 1><<<G.ntblocks, kBlock, 0, s>>>(B->runs, G.tmap, rs);
    else if (G.tE == 2) kb_row_tensor<T, 2><<<G.ntblocks, kBlock, 0, s>>>(B->runs, G.tmap, rs);
    else kb_row_tensor<T, 4><<<G.ntblocks, kBlock, 0, s>>>(B->runs, G.tmap, rs);
}

template <typename T, bool SPLIT>
static void batch_group_resident(const BatchDev *B, const BatchGroup &G, int count, hipStream_t s)
{
    if (G.pot) kb_resident<T, SPLIT, kPotTab><<<G.nblocks, kResThreads, 0, s>>>(B->runs, G.map, count);
    else kb_resident<T, SPLIT, kPotHill><<<G.nblocks, kResThreads, 0, s>>>(B->runs, G.map, count);
}

hipError_t batch_step(const BatchDev *B, long long n, hipStream_t s);

hipError_t batch_steps(const BatchDev *B, long long n0, int count, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    for (const BatchGroup &G : B->groups) {
        if (!G.res) continue;
        if (B->split) batch_group_resident<float, true>(B, G, count, s);
        else batch_group_resident<double, false>(B, G, count, s);
    }
    if (B->has_plain || B->has_march)
        for (int q = 0; q < count; ++q) {
            const hipError_t e = batch_step(B, n0 + q, s);
            if (e != hipSuccess) return e;
        }
    return hipGetLastError();
}

hipError_t batch_step(const BatchDev *B, long long n, hipStream_t s)
{
    if (B->has_plain) {
        for (const BatchGroup &G : B->groups) {
            if (G.march || G.res) continue;
            if (B->split) batch_group_reduce<float, true>(B, G, s);
            else batch_group_reduce<double, false>(B, G, s);
        }
        kb_friedmann_finish<<<B->nruns, kBlock, 0, s>>>(B->runs);
    }
    for (const BatchGroup &G : B->groups) {
        if (G.res) continue;
        if (B->split) batch_group_kick<float, true>(B, G, n, s);
        else batch_group_kick<double, false>(B, G, n, s);
    }
    if (B->has_plain) kb_step_finish<<<(B->nruns + 63) / 64, 64, 0, s>>>(B->runs, B->nruns);
    return hipGetLastError();
}

hipError_t batch_loop_end(const BatchDev *B, long long steps, hipStream_t s)
{
    if (B->has_march) kb_loop_end<<<B->nruns, kBlock, 0, s>>>(B->runs, (unsigned)(steps & 1));
    return hipGetLastError();
}

}  // namespace ie
