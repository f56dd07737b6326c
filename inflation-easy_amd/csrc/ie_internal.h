// ie_internal.h -- device state, shared device helpers and launcher declarations
// of the gfx950 lattice engine.  Not part of the public ABI (include/ie_api.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

#include "ie_api.h"
#include "ie_tabindex.h"

namespace ie {

constexpr int kBlock = 256;      // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kMaxRed = 8;       // reduction slots per block

// Scalars of one loop step as a pure function of the background between steps
// and the sums of the current f in DevBg::red (run_evolution_loop's step
// dt astep^{s-1}, the implicit Friedmann half-kick of evolve_derivs, the kick
// coefficients, and for split storage the homogeneous mode's kick and drift;
// src/evolution.cpp:153-181, 256-283, 938).
struct StepScalars {
    double d, ad2, ad_half, ad_next, friction, potnorm, laplnorm, srcamp, vpmean, pibar_next, phibar_next;
};

// Background of one run, resident on the device.  Kernels of a step read it at
// their start; one-workgroup finish kernels (or the last workgroup of an output
// reduction, by arrival counter) update it, so a whole step sequence runs without
// the host.
struct DevBg {
    double t, a, ad, ad2, astep, Ne, phiref, hubble_init;
    double d;                                   // step of the current kick/drift
    double laplnorm, friction, potnorm, srcamp;
    double phibar, pibar;                       // homogeneous mode (split storage)
    double vpmean;                              // <V'(phi)> over the lattice (split)
    double red[kMaxRed];                        // results of the last reduction
    double dt, af, rescale_s, dx, inv_gs;       // constants of the run
    StepScalars nx;                             // scalars of the next fused loop step: set by the
                                                // kernel that closes the previous step (or sums the
                                                // f a loop starts from), read by the step's kernels
    long long steps;
    int done, error, parity, pad0;
    unsigned int counter;
    unsigned int pad;
    int marched;                                // step record of a fused loop: its step has run
    int hill_sums;                              // fused loops of the hilltop potential sum f and f^2 in
                                                // place of V and V' (1: FP64 storage, 2: split storage,
                                                // f the fluctuation about phibar); 0: V and V'
    double hV, hV4, hV2;                        // the hilltop's V0/B^2, V0 (1-ns)/4 /B^2, V0 (1-ns)/2 /B^2
};

// Potential parameters (src/potential.cpp:77-83 analytic hilltop; tabulated
// interpolation :11-68).  Values are in units of B^2 = rescale_B^2.
struct Pot {
    int numerical;
    int n;
    double V0, ns, B2;
    const double *x, *v, *dv;   // device table, descending x
    // constants of the lattice kernels: V0/B^2, V0 (1-ns)/4 /B^2, V0 (1-ns)/2 /B^2, 1/B^2
    double hV, hV4, hV2, invB2;
    // the table's O(1) interval index with the lattice kernels' form of V/B^2,
    // V'/B^2 and their slopes per interval
    TabIndex T;
};

// Device pointers of one lattice's fields.  Storage of f, fd is double (FP64) or
// float (split: fluctuation about the homogeneous mode in DevBg).
struct Fields {
    void *f[2];
    void *fd[2];            // f and fd: current buffer [parity], a loop step writes [parity ^ 1]
    float *h[2][6];
    float *hd[6];
    double *dn;
    int N, logN;
    int sigw;
    int kick_e, tensor_e;   // kick sites per lane / tensor shape code (0: default)
    int march_pf;           // fused step kernel variant: 1 loads the next plane early
    DevBg *rec;             // the two step records of a fused loop (device)
};

// ---------------------------------------------------------------------- helpers

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Deterministic sum of NV values over the first kBlock threads of the workgroup
// (all of it but in the fused step kernel, whose larger workgroup passes zeros
// from its other waves); result valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *smem)
{
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0 && wave < kWaves) {
#pragma unroll
        for (int q = 0; q < NV; ++q) smem[q * kWaves + wave] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int w = 0; w < kWaves; ++w) s += smem[q * kWaves + w];
            v[q] = s;
        }
    }
    __syncthreads();
}

// Arrival counter: returns true in every thread of the last workgroup to arrive.
// Preceded by a release fence so the partials of all workgroups are visible.
__device__ __forceinline__ bool last_block(unsigned int *counter, unsigned int nblocks)
{
    __shared__ int is_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();  // release this workgroup's partials (device scope, all XCDs)
        const unsigned int prev = atomicAdd(counter, 1u);
        is_last = (prev == nblocks - 1);
        if (is_last) *counter = 0;
    }
    __syncthreads();
    if (is_last) __threadfence();  // acquire in every thread before reading partials
    return is_last != 0;
}

// Fixed-order sum of `nblocks` partials of NV values by the last workgroup.
template <int NV>
__device__ __forceinline__ void sum_partials(const double *partials, int nblocks, double (&out)[NV],
                                             double *smem)
{
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] = 0.0;
    for (int b = threadIdx.x; b < nblocks && threadIdx.x < kBlock; b += kBlock) {
#pragma unroll
        for (int q = 0; q < NV; ++q)
            acc[q] += partials[(size_t)b * NV + q];
    }
    block_sum<NV>(acc, smem);
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = acc[q];
}

// ----------------------------------------------------------------- potential

__device__ __forceinline__ int pot_bracket(const Pot &P, double v, int *err)
{
    return tab_bisect_ref(P.x, P.n, v, err);
}

__device__ __forceinline__ double pot_interp(const Pot &P, double v, int l, const double *y)
{
    const double x0 = P.x[l - 1], x1 = P.x[l], y0 = y[l - 1], y1 = y[l];
    return y1 + (v - x1) * (y0 - y1) / (x0 - x1);
}

// V/B^2 and V'/B^2 at v with the reference's expressions; *err set when v
// leaves the table.
__device__ __forceinline__ void pot_eval(const Pot &P, double v, double *V, double *dV, int *err)
{
    if (P.numerical) {
        const int l = pot_bracket(P, v, err);
        *V = pot_interp(P, v, l, P.v) / P.B2;
        *dV = pot_interp(P, v, l, P.dv) / P.B2;
    } else {
        *V = P.V0 * (1.0 - (1.0 - P.ns) / 4.0 * (v * v)) / P.B2;
        *dV = -P.V0 * (1.0 - P.ns) / 2.0 * v / P.B2;
    }
}

// The potential as the lattice kernels evaluate it: the division by B^2 folded
// into constants, and for a table the interpolation y_l + (v - x_l) s_l through
// the interval's precomputed slope (a rounding or two from the reference's
// y_l + (v - x_l)(y_{l-1} - y_l)/(x_{l-1} - x_l), below the lattice sums' own
// order dependence).
// A lane's K values are looked up together.  The neighbouring sites of a wave
// lie in a few neighbouring intervals: the first active lane's hint (the
// interval of a nearby earlier value) names the wave's interval, whose record is
// read once for the wave through the scalar cache (the tables are never written
// by a kernel, so they are read as constant memory); a wave across grid points
// reads its neighbours' too, and then the three intervals beyond them on one
// side.  The values inside these interpolate from scalar registers: a per-lane
// copy of a shared record costs the vector memory path as much as a field row,
// and would bound the step.  A value outside them takes its own hint's record,
// then the O(1) index of ie_tabindex.h, then (a rounding at an index edge, a
// table end) the bisection.
typedef const __attribute__((address_space(4))) double *ConstDoubles;

__device__ __forceinline__ TabRec tab_rec_uniform(const TabRec *rec, int l)
{
    const ConstDoubles p = (ConstDoubles)(rec + l);
    return TabRec{p[0], p[1], p[2], p[3], p[4], p[5]};
}

struct PotView {
    Pot P;

    // Values in the three intervals around lu (records rm above, ru, rp below,
    // read through the scalar cache).  V and V' are continuous and linear on each
    // interval, so a value's interpolation is ru's line plus a hinge at each of
    // ru's grid points:
    //   y(v) = y_u + d s_u + max(v - xhi_u, 0) (s_m - s_u) + min(d, 0) (s_p - s_u),
    //   d = v - xlo_u,
    // which is the line of the value's own interval in another rounding order,
    // with no per-value compare or select.  At a table end the missing
    // neighbour is ru itself (no hinge: the window then ends at ru's edge).
    template <int K, bool WANT_V>
    __device__ __forceinline__ void hinges(const TabRec &rm, const TabRec &ru, const TabRec &rp, const double (&v)[K],
                                           double (&V)[K], double (&dv)[K]) const
    {
        const double gm = rm.sdv - ru.sdv, gp = rp.sdv - ru.sdv;
        const double hm = rm.sv - ru.sv, hp = rp.sv - ru.sv;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const double d = v[q] - ru.xlo, up = fmax(v[q] - ru.xhi, 0.0), dn = fmin(d, 0.0);
            dv[q] = ru.ydv + d * ru.sdv + up * gm + dn * gp;
            if (WANT_V) V[q] = ru.yv + d * ru.sv + up * hm + dn * hp;
        }
    }

    // The wave's interval for its next lookup: that of lane 0's first value, one
    // of the window's three.
    __device__ __forceinline__ int follow(double v0, int lu, const TabRec &ru) const
    {
        return v0 > ru.xhi ? lu - 1 : (v0 <= ru.xlo ? lu + 1 : lu);
    }

    // Whether every value of the wave lies in the window rm .. rp.
    template <int K>
    __device__ __forceinline__ bool in_window(const TabRec &rm, const TabRec &rp, const double (&v)[K]) const
    {
        bool win = true;
#pragma unroll
        for (int q = 0; q < K; ++q) win = win & (v[q] <= rm.xhi) & (v[q] > rp.xlo);   // false for NaN
        return __all(win);
    }

    // V/B^2 (WANT_V) and V'/B^2 at the K values v, hints l updated.
    template <int K, bool WANT_V>
    __device__ __forceinline__ void lookup(const double (&v)[K], int (&l)[K], double (&V)[K], double (&dv)[K],
                                           int *err) const
    {
        // the wave's interval lu and its two neighbours, in scalar registers: the
        // common case, every value of the wave in one of them (read in one round:
        // a second round of scalar loads would add its latency to the lookup)
        const int lu = __builtin_amdgcn_readfirstlane(l[0]);
        const int cm = lu > 1 ? lu - 1 : lu, cp = lu < P.n - 1 ? lu + 1 : lu;
        const TabRec ru = tab_rec_uniform(P.T.rec, lu), rm = tab_rec_uniform(P.T.rec, cm),
                     rp = tab_rec_uniform(P.T.rec, cp);
        if (in_window<K>(rm, rp, v)) {
            hinges<K, WANT_V>(rm, ru, rp, v, V, dv);
            const int nl = follow(v[0], lu, ru);
#pragma unroll
            for (int q = 0; q < K; ++q) l[q] = nl;
            return;
        }
        // otherwise (a dense stretch of the grid, where a wave's values spread
        // over many intervals) every value by its own lookup
#pragma unroll
        for (int q = 0; q < K; ++q) lookup_value<WANT_V>(v[q], l[q], V[q], dv[q], err);
    }
    // Sums of V/B^2 and V'/B^2 over the K values v, added to sV and sdV (hints l,
    // updated), for lattice sums.  Values in the window of three intervals
    // around lu take the hinge form summed: K y_u + s_u sum d + (s_m - s_u) sum
    // max(v - xhi_u, 0) + (s_p - s_u) sum min(d, 0) (the sum of their
    // interpolations in another rounding order).  A wave with values outside
    // the window takes lookup() one value at a time.
    template <int K>
    __device__ __forceinline__ void sums(const double (&v)[K], int (&l)[K], double &sV, double &sdV, int *err) const
    {
        const int lu = __builtin_amdgcn_readfirstlane(l[0]);
        const int cm = lu > 1 ? lu - 1 : lu, cp = lu < P.n - 1 ? lu + 1 : lu;
        const TabRec ru = tab_rec_uniform(P.T.rec, lu), rm = tab_rec_uniform(P.T.rec, cm),
                     rp = tab_rec_uniform(P.T.rec, cp);
        if (in_window<K>(rm, rp, v)) {
            double sd = 0.0, su = 0.0, sn = 0.0;
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const double d = v[q] - ru.xlo;
                sd += d;
                su += fmax(v[q] - ru.xhi, 0.0);
                sn += fmin(d, 0.0);
            }
            const double f = (double)K;
            sV += f * ru.yv + ru.sv * sd + (rm.sv - ru.sv) * su + (rp.sv - ru.sv) * sn;
            sdV += f * ru.ydv + ru.sdv * sd + (rm.sdv - ru.sdv) * su + (rp.sdv - ru.sdv) * sn;
            const int nl = follow(v[0], lu, ru);
#pragma unroll
            for (int q = 0; q < K; ++q) l[q] = nl;
            return;
        }
#pragma unroll
        for (int q = 0; q < K; ++q) {
            double a, b;
            lookup_value<true>(v[q], l[q], a, b, err);
            sV += a;
            sdV += b;
        }
    }
    // One value outside its wave's window: the O(1) index (its two levels in
    // LDS when the kernel staged them), then, at an index edge or a table end, the
    // bisection.  The value's own hint is not tried: where values leave their
    // wave's window the grid is dense and a neighbouring site's interval rarely
    // holds them, and each try costs a dependent record load.
    template <bool WANT_V>
    __device__ __forceinline__ void lookup_value(double v, int &l, double &V, double &dv, int *err) const
    {
        int lq = tab_candidate(P.T, v);
        TabRec r = P.T.vrec[lq];
        if (!tab_in(r, v)) {
            lq = tab_bisect_ref(P.x, P.n, v, err);
            r = P.T.vrec[lq];
        }
        l = lq;
        dv = r.ydv + (v - r.xlo) * r.sdv;
        if (WANT_V) V = r.yv + (v - r.xlo) * r.sv;
    }
    // V'/B^2 at the K values v (hints l, updated)
    template <int K>
    __device__ __forceinline__ void dV(const double (&v)[K], int (&l)[K], double (&dv)[K], int *err) const
    {
        if (!P.numerical) {
#pragma unroll
            for (int q = 0; q < K; ++q) dv[q] = -P.hV2 * v[q];
            return;
        }
        double unused[K];
        lookup<K, false>(v, l, unused, dv, err);
    }
    // V/B^2 and V'/B^2 at the K values v (hints l, updated)
    template <int K>
    __device__ __forceinline__ void eval(const double (&v)[K], int (&l)[K], double (&V)[K], double (&dv)[K],
                                         int *err) const
    {
        if (!P.numerical) {
#pragma unroll
            for (int q = 0; q < K; ++q) {
                V[q] = P.hV - P.hV4 * (v[q] * v[q]);
                dv[q] = -P.hV2 * v[q];
            }
            return;
        }
        lookup<K, true>(v, l, V, dv, err);
    }
    // one value from its hint (updated)
    __device__ __forceinline__ double dV1(double v, int &hint, int *err) const
    {
        double x[1] = {v}, d[1];
        int l[1] = {hint};
        dV<1>(x, l, d, err);
        hint = l[0];
        return d[0];
    }
    __device__ __forceinline__ void eval1(double v, double *V, double *dv, int &hint, int *err) const
    {
        double x[1] = {v}, a[1], d[1];
        int l[1] = {hint};
        eval<1>(x, l, a, d, err);
        hint = l[0];
        *V = a[0];
        *dv = d[0];
    }
    // a first hint for the interval of v
    __device__ __forceinline__ int start(double v) const { return P.numerical ? tab_candidate(P.T, v) : 0; }
};

__device__ __forceinline__ PotView pot_view(const Pot &P) { return PotView{P}; }

// Dynamic LDS of the tabulated march launches: the two index levels of the run's
// table (for the inputs/ table: 35 KB; an index above kTabIndexMax stays in
// global memory), and, with four sites per lane (N = 256: one workgroup per CU
// either way), the interval records after it (55 KB more), which the per-value
// lookups of a dense stretch then read from LDS instead of L2 (256^3 split step
// there 0.133 -> 0.108 ms).  With two sites per lane (N = 128) the records would
// leave one workgroup per CU instead of two: measured slower at the start of a
// run (0.0179 -> 0.0194 ms) and faster in the dense stretch (0.0272 -> 0.0257),
// no gain over a whole run, so they stay in L2 there.
constexpr unsigned kTabIndexMax = 48 * 1024, kTabIndexLds = kTabIndexMax, kTabRecordsLds = 112 * 1024;

__host__ __device__ __forceinline__ unsigned tab_index_bytes(const TabIndex &T)
{
    return (unsigned)(T.k1 * sizeof(TabBucket) + T.nsub * sizeof(TabSub));
}

// Copies the index levels of Pv's table into `lds` (every thread of the
// workgroup takes part) when they fit in `budget` bytes, then the interval
// records when they fit too, and points Pv at the copies; the caller's next
// barrier publishes them.
__device__ __forceinline__ void stage_tab_index(PotView &Pv, char *lds, unsigned nthreads, unsigned budget)
{
    static_assert(sizeof(TabBucket) == 8 && sizeof(TabSub) == 16, "index record sizes");
    const unsigned ib = tab_index_bytes(Pv.P.T);
    if (!Pv.P.numerical || ib > kTabIndexMax || ib > budget) return;
    // l1 (k1 a power of two >= 16, so l2 starts 16-byte aligned)
    const unsigned k1 = (unsigned)Pv.P.T.k1, nsub = (unsigned)Pv.P.T.nsub;
    const uint2 *s1 = (const uint2 *)Pv.P.T.l1;
    const uint4 *s2 = (const uint4 *)Pv.P.T.l2;
    uint2 *d1 = (uint2 *)lds;
    uint4 *d2 = (uint4 *)(lds + k1 * sizeof(TabBucket));
    for (unsigned q = threadIdx.x; q < k1; q += nthreads) d1[q] = s1[q];
    for (unsigned q = threadIdx.x; q < nsub; q += nthreads) d2[q] = s2[q];
    Pv.P.T.l1 = (const TabBucket *)d1;
    Pv.P.T.l2 = (const TabSub *)d2;
    // the records too where the budget holds them (16-byte units after the index)
    const unsigned rb = (unsigned)Pv.P.T.n * (unsigned)sizeof(TabRec);
    if (ib + rb <= budget) {
        const uint4 *s3 = (const uint4 *)Pv.P.T.rec;
        uint4 *d3 = (uint4 *)(lds + ib);
        for (unsigned q = threadIdx.x; q < rb / 16; q += nthreads) d3[q] = s3[q];
        Pv.P.T.vrec = (const TabRec *)d3;
    }
}

// Exact sums on the grid of one binade (the float mean of saveN, the lattice-
// order FP64 sums of the post-inflation set-up): an accumulator that is k
// units of the binade's grid takes a step to k + n, or at a rounding tie
// (k + f + 1/2) to the even one of k + f and k + f + 1; a run of such steps is
// k -> k + A + (t ? (k + c) mod 2 : 0) (t: the run holds a tie; after its first
// tie every later tie's parity is known), and runs compose into runs.
struct GridRun {
    long long A;
    int f;   // bit 0: t, bit 1: c
};
__device__ __forceinline__ GridRun grid_step(long long n) { return GridRun{n, 0}; }
__device__ __forceinline__ GridRun grid_tie(long long f) { return GridRun{f, 1 | (int)((f & 1) << 1)}; }
__device__ __forceinline__ GridRun grid_compose(GridRun a, GridRun b)
{
    const int ta = a.f & 1, tb = b.f & 1, ca = (a.f >> 1) & 1, cb = (b.f >> 1) & 1;
    GridRun r;
    r.A = a.A + b.A + ((ta & tb) ? (int)((a.A + ca + cb) & 1) : 0);
    const int c = ta ? ca : (tb ? (int)((a.A + cb) & 1) : 0);
    r.f = (ta | tb) | (c << 1);
    return r;
}
__device__ __forceinline__ long long grid_apply(GridRun r, long long k)
{
    return k + r.A + ((r.f & 1) ? ((k + ((r.f >> 1) & 1)) & 1) : 0);
}

// ---------------------------------------------------------------- launchers
// (ie_lattice.hip)

enum KickMode { KICK = 1, DRIFT = 2, KICKDRIFT = 3 };

// Workgroups of the step kernels of an N^3 lattice (the same alone and batched,
// so a lattice's fixed-order sums do not depend on its company).
unsigned step_blocks(int N);
// Sites per lane of the row kernels (0 below N = 64) and of a kick / tensor launch
// asked for `want` sites per lane.
int row_elems(int N);
int shape_elems(int N, int want);
// Tensor step of loop steps (N >= 64): Fields::tensor_e is a shape code, 0 or
// 1/2/4 for the row kernel, kTensorMarch | E (E = 1, 2, 4) for the march kernel,
// plus kTensorPf for its variant that loads the next plane early.
constexpr int kTensorMarch = 16;
constexpr int kTensorPf = 8;
bool tensor_is_march(int code);
int tensor_march_elems(int N, int code, int split);
unsigned tensor_march_blocks(int N, int E);
// Partial Friedmann sums of the current f, one triple per workgroup.
hipError_t launch_reduce(const Fields &F, int split, const Pot &P, DevBg *bg, double *partials, int loop,
                         hipStream_t s);
// One workgroup: fixed-order total of the partials and the pre-kick scalar update
// of evolve_derivs (src/evolution.cpp:153-167).  loop = 0: explicit step d_arg;
// loop = 1: loop step d = dt astep^{s-1}, skipped once done.
hipError_t launch_friedmann_finish(const Fields &F, DevBg *bg, const double *partials, double d_arg, int loop,
                                   hipStream_t s);
// Kick / drift / fused kick+drift of f, fd; of h, hd (tensor step).  The tensor
// step's scalars come from bg (fresh = 0) or, in a fused loop, from the step
// record in slot fresh - 1 (step_record).
hipError_t launch_scalar_step(const Fields &F, int split, const Pot &P, DevBg *bg, int mode, double d_arg, int loop,
                              hipStream_t s);
hipError_t launch_tensor_step(const Fields &F, int split, DevBg *bg, int mode, double d_arg, int loop, hipStream_t s,
                              int fresh = 0);
// Fused loop steps for 128 <= N <= 256: the march kernel closes the previous step
// (the scalar updates from its partial sums) into the loop's step records
// (Fields::rec), then kicks, drifts and sums the new f for the next step's
// Friedmann update.  A loop starts with launch_march_begin (a reduce of the
// current f into bg->red and the first record); then launch_fused_step per step
// (march, and the tensor step with the same scalars when sigw); then
// launch_loop_end, which closes the last step into bg.
bool march_ok(int N);
unsigned march_blocks(int N);
hipError_t launch_march_begin(const Fields &F, int split, const Pot &P, DevBg *bg, double *partials, hipStream_t s);
hipError_t launch_march(const Fields &F, int split, const Pot &P, DevBg *bg, double *partials, long long n,
                        hipStream_t s);
// launch_tensor_step's `fresh` for step n of a fused loop: that step's record
int step_record(long long n);
hipError_t launch_loop_end(const Fields &F, int split, DevBg *bg, const double *partials, long long steps,
                           hipStream_t s);
hipError_t launch_fused_step(const Fields &F, int split, const Pot &P, DevBg *bg, double *partials, long long n,
                             hipStream_t s);
// One thread: the scalar updates that close a kick, a drift or a loop step.
hipError_t launch_step_finish(DevBg *bg, int split, int mode, double d_arg, int loop, hipStream_t s);
// launch_reduce + launch_friedmann_finish.
hipError_t launch_friedmann_reduce(const Fields &F, int split, const Pot &P, DevBg *bg, double *partials,
                                   double d_arg, int loop, hipStream_t s);
// Scalar step, tensor step (when sigw) and step finish.  loop = 1: a step of the
// device-resident loop (skipped once done, step from DevBg); loop = 0: explicit
// call with step d_arg (drift) or the step of the last reduce.
// part: -1 all launches, 0 scalar step only (with the finish when !sigw),
// 1 tensor step and finish.
hipError_t launch_kick_drift(const Fields &F, int split, const Pot &P, DevBg *bg, int mode, double d_arg,
                             int loop, hipStream_t s, int part = -1);
// Sums {sum f, sum fd, sum f^2} (split: about the homogeneous mode), kinetic,
// gradient and potential sums into bg->red[0..5].
hipError_t launch_field_sums(const Fields &F, int split, const Pot &P, DevBg *bg, double *partials,
                             hipStream_t s);
hipError_t launch_split_upload(const double *f, const double *fd, Fields &F, DevBg *bg, double *partials,
                               hipStream_t s);
hipError_t launch_reconstitute(const Fields &F, int split, int which, const DevBg *bg, double *out,
                               hipStream_t s);
int reduce_blocks(size_t gs);

// Batched loop steps over many lattices of one storage kind (ie_lattice.hip).
struct BatchRun {
    Fields F;
    Pot P;
    DevBg *bg;
    double *partials;
};
struct BatchDev;
hipError_t batch_build(const BatchRun *runs, int n, int split, BatchDev **out);
// `count` loop steps of every run from loop step n0: resident groups in one
// launch, the other groups one step at a time
hipError_t batch_steps(const BatchDev *B, long long n0, int count, hipStream_t s);
bool resident_ok(int N, int sigw, int split, int numerical);
// Minimum lattices of a potential kind for the resident path (0 restores these
// defaults).  Measured per step, 64^3 split storage, graph replay: hilltop
// resident 0.116 ms from 1 to 16 lattices, 0.134 at 64, 0.147 at 128, 0.221 at
// 256, against 0.012, 0.035, 0.120, 0.241, 0.470 ms for the per-step kernels
// (1, 16, 64, 128, 256 lattices); tabulated resident 0.31-0.32 ms from 1 to 256
// lattices against 0.014, 0.039, 0.123, 0.260, 0.513.
constexpr int kResMinHill = 72, kResMinTab = 152;
void set_resident_min(int hill, int tab);
void batch_free(BatchDev *B);
hipError_t batch_loop_begin(const BatchDev *B, hipStream_t s);
hipError_t batch_step(const BatchDev *B, long long n, hipStream_t s);        // step n of the loop
hipError_t batch_loop_end(const BatchDev *B, long long steps, hipStream_t s);  // after its steps

// deltaN phase and field utilities (ie_deltaN.hip).  `work` is a device buffer of
// at least 4096 doubles.
hipError_t dn_init(const Fields &F, const double *f, double *fd, double *dn, const Pot &P, double apow, int *error,
                   hipStream_t s);
hipError_t dn_fieldsN(size_t gs, double *f, const double *fd, double *dn, const Pot &P, int mono, double phiref,
                      double Vref, double d, hipStream_t s);
hipError_t dn_steps(size_t gs, double *f, double *fd, double *dn, const Pot &P, int mono, double phiref, double Vref,
                    double dN, long long nsteps, int *error, hipStream_t s);
hipError_t dn_rk4_steps(size_t gs, double *f, double *fd, double *dn, const Pot &P, int mono, double phiref,
                        double Vref, double h, long long nsteps, int *error, hipStream_t s);
// One RK45 trial step of every site into (f5, fd5, dn5); *out receives the
// fixed-order sum of the weighted squared errors (work: >= 2048 partials).
hipError_t dn_rk45_trial(size_t gs, const double *f, const double *fd, const double *dn, double *f5, double *fd5,
                         double *dn5, const Pot &P, int mono, double phiref, double Vref, double h, double abs_tol,
                         double rel_tol, double *work, double *out, int *error, hipStream_t s);
hipError_t dn_phiref(size_t gs, const double *f, const Pot &P, int mono, double *work, double *out, hipStream_t s);
hipError_t dn_float_mean(const double *dn, size_t gs, double cutoff, float *out, hipStream_t s);
hipError_t dn_zeta(double *dn, size_t gs, const float *nmean, double cutoff, int zero, hipStream_t s);
hipError_t field_sum(const double *g, size_t gs, double *work, double *out, hipStream_t s);
hipError_t field_center(double *g, size_t gs, double *work, double *sum_out, hipStream_t s);
hipError_t field_minmax(const double *g, size_t gs, double *work, hipStream_t s, int *nb_out);
hipError_t field_histogram(const double *g, size_t gs, double bmin, double bmax, double df, int nbins,
                           unsigned long long *counts, hipStream_t s);

// ------------------------------------------------------------ row helpers
// Shared by the row kernels of the inflation and post-inflation steps.

constexpr int kXcd = 8;          // XCDs of the MI355X; workgroup b runs on XCD b mod 8

// E consecutive values of one lane, loaded and stored as one vector.
template <typename T, int E>
struct alignas(sizeof(T) * E) RowVec {
    T v[E];
};

template <typename T, int E>
__device__ __forceinline__ void ld_row(const T *__restrict__ row, unsigned lane, T (&o)[E])
{
    const RowVec<T, E> x = reinterpret_cast<const RowVec<T, E> *>(row)[lane];
#pragma unroll
    for (int e = 0; e < E; ++e) o[e] = x.v[e];
}

template <typename T, int E>
__device__ __forceinline__ void st_row(T *__restrict__ row, unsigned lane, const T (&v)[E])
{
    RowVec<T, E> x;
#pragma unroll
    for (int e = 0; e < E; ++e) x.v[e] = v[e];
    reinterpret_cast<RowVec<T, E> *>(row)[lane] = x;
}

// Buffer resource over `bytes` bytes of a field (raw dword access; an access past
// the end reads 0 and writes nothing).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}

typedef unsigned ie_v2u __attribute__((ext_vector_type(2)));
typedef unsigned ie_v4u __attribute__((ext_vector_type(4)));

// A lane's E values of the row at byte offset `soff` (wave-uniform, a scalar
// register) through a buffer resource: no per-lane address arithmetic.
template <typename T, int E>
__device__ __forceinline__ void bld_row(__amdgpu_buffer_rsrc_t r, unsigned soff, unsigned lane, T (&o)[E])
{
    constexpr unsigned B = (unsigned)sizeof(T) * E;
    const unsigned vo = lane * B;
    RowVec<T, E> x;
    if constexpr (B == 4) {
        x = __builtin_bit_cast(RowVec<T, E>, __builtin_amdgcn_raw_buffer_load_b32(r, vo, soff, 0));
    } else if constexpr (B == 8) {
        x = __builtin_bit_cast(RowVec<T, E>, __builtin_amdgcn_raw_buffer_load_b64(r, vo, soff, 0));
    } else if constexpr (B == 16) {
        x = __builtin_bit_cast(RowVec<T, E>, __builtin_amdgcn_raw_buffer_load_b128(r, vo, soff, 0));
    } else {
        static_assert(B == 32, "rows of 4, 8, 16 or 32 bytes per lane");
        struct alignas(32) Two {
            ie_v4u a, b;
        } t;
        t.a = __builtin_amdgcn_raw_buffer_load_b128(r, vo, soff, 0);
        t.b = __builtin_amdgcn_raw_buffer_load_b128(r, vo + 16, soff, 0);
        x = __builtin_bit_cast(RowVec<T, E>, t);
    }
#pragma unroll
    for (int e = 0; e < E; ++e) o[e] = x.v[e];
}

template <typename T, int E>
__device__ __forceinline__ void bst_row(__amdgpu_buffer_rsrc_t r, unsigned soff, unsigned lane, const T (&v)[E])
{
    constexpr unsigned B = (unsigned)sizeof(T) * E;
    const unsigned vo = lane * B;
    RowVec<T, E> x;
#pragma unroll
    for (int e = 0; e < E; ++e) x.v[e] = v[e];
    if constexpr (B == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, vo, soff, 0);
    } else if constexpr (B == 8) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(ie_v2u, x), r, vo, soff, 0);
    } else if constexpr (B == 16) {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(ie_v4u, x), r, vo, soff, 0);
    } else {
        static_assert(B == 32, "rows of 4, 8, 16 or 32 bytes per lane");
        struct alignas(32) Two {
            ie_v4u a, b;
        };
        const Two t = __builtin_bit_cast(Two, x);
        __builtin_amdgcn_raw_buffer_store_b128(t.a, r, vo, soff, 0);
        __builtin_amdgcn_raw_buffer_store_b128(t.b, r, vo + 16, soff, 0);
    }
}

// Work of one wave: rows j0 .. j0+R-1 of plane i, k-segment [k0, k0 + 64 E).
struct RowWalk {
    unsigned i, j0, k0, R, segs;
};

// G workgroups of 4 waves per lattice; units (i, segment, j-block) in plane-major
// order, XCD x owning the planes [x N/8, (x+1) N/8).
__device__ __forceinline__ RowWalk row_walk(unsigned N, unsigned G, unsigned blk, int E)
{
    const unsigned segs = N / (64u * (unsigned)E);
    const unsigned waves = 4u * G;
    const unsigned R = (N * N * segs) / waves;
    const unsigned jblocks = N / R;
    const unsigned xcd = blk % kXcd, lb = blk / kXcd;
    const unsigned u = xcd * (waves / kXcd) + lb * 4u + (threadIdx.x >> 6);
    RowWalk w;
    const unsigned t = u / jblocks;
    w.j0 = (u % jblocks) * R;
    w.k0 = (t % segs) * (64u * (unsigned)E);
    w.i = t / segs;
    w.R = R;
    w.segs = segs;
    return w;
}

// The value of the first active lane, in scalar registers (the same in every lane).
__device__ __forceinline__ float first_lane(float x)
{
    return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(x)));
}
__device__ __forceinline__ double first_lane(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// k +- 1 neighbours of the lane's E sites of row `c` (one segment of the row at
// `row`): in-lane, then across lanes, then across segments for N > 64 E.
template <typename T, int E>
__device__ __forceinline__ void k_neighbours(const T (&c)[E], const T *__restrict__ row, const RowWalk &w,
                                             unsigned N, unsigned lane, T &kp_last, T &km_first)
{
    kp_last = __shfl(c[0], (int)((lane + 1) & 63));
    km_first = __shfl(c[E - 1], (int)((lane + 63) & 63));
    if (w.segs > 1) {
        // the two edge sites, loaded by every lane and kept wave-uniform, then
        // selected per lane without a branch
        const unsigned m = N - 1;
        const T lo = first_lane(row[(w.k0 - 1) & m]), hi = first_lane(row[(w.k0 + 64u * E) & m]);
        km_first = lane == 0 ? lo : km_first;
        kp_last = lane == 63 ? hi : kp_last;
    }
}

// Post-inflation steps (ie_postinf.hip).  Scalars of one leapfrog step of
// evolve_derivs_post_inflation + evolve_fields, computed by the host in the
// reference's order (src/evolution.cpp:1437-1564).
struct PostStep {
    double d, dd, A, c1ad, a;                      // kick step, drift step (0: none), w laplnorm, (3(1+w)+s) a', a
    double laplnorm, c2ad, srcamp, invH, coeffU;   // tensor kick: laplnorm, (2+s) a', 2 a^-2s, a/a', 4/(3(1+w))
    double hx, inv_dx2, cross_den;                 // 0.5/dx, 1/(dx dx), 4 dx dx
};
// FP64 Phi, Phi' and FP32 h_ij, h_ij' of a step: read f, fd, h; write f_out,
// fd_out, h_out; hd in place.  post_drift uses f, fd, h, hd only (in place).
struct PostState {
    double *f, *f_out, *fd, *fd_out;
    float *h[6], *h_out[6], *hd[6];
};
#pragma clang fp contract(off)
// Stencil of one site for the post-inflation tensor source, and the source
// S_ij (xx yy zz xy xz yz) in the reference's operation order
// (src/evolution.cpp:1474-1503, 549-570); contraction off, as the reference build.
struct SiteStencil {
    double c;                   // Phi(x)
    double fxp, fxm, fyp, fym, fzp, fzm;      // Phi at the six faces
    double fxpyp, fxpym, fxmyp, fxmym;        // Phi(i +- 1, j +- 1, k)
    double fxpzp, fxpzm, fxmzp, fxmzm;        // Phi(i +- 1, j, k +- 1)
    double fypzp, fymzm;                      // Phi(i, j+1, k+1), Phi(i, j-1, k-1)
    double dc, dxp, dxm, dyp, dym, dzp, dzm;  // kicked Phi' at x and its faces
};

__device__ __forceinline__ void post_source(const SiteStencil &q, const PostStep &P, double S[6])
{
    const double gx = (q.fxp - q.fxm) * P.hx;
    const double gy = (q.fyp - q.fym) * P.hx;
    const double gz = (q.fzp - q.fzm) * P.hx;
    const double gdx = (q.dxp - q.dxm) * P.hx;
    const double gdy = (q.dyp - q.dym) * P.hx;
    const double gdz = (q.dzp - q.dzm) * P.hx;
    const double Hxx = (q.fxp - 2.0 * q.c + q.fxm) * P.inv_dx2;
    const double Hyy = (q.fyp - 2.0 * q.c + q.fym) * P.inv_dx2;
    const double Hzz = (q.fzp - 2.0 * q.c + q.fzm) * P.inv_dx2;
    const double Hxy = (q.fxpyp - q.fxpym - q.fxmyp + q.fxmym) / P.cross_den;
    const double Hxz = (q.fxpzp - q.fxpzm - q.fxmzp + q.fxmzm) / P.cross_den;
    const double Hyz = (q.fypzp - q.fymzm - q.fypzp + q.fymzm) / P.cross_den;
    const double dux = gdx * P.invH + gx;
    const double duy = gdy * P.invH + gy;
    const double duz = gdz * P.invH + gz;
    const double Phi = q.c;
    S[0] = 4.0 * Phi * Hxx + 2.0 * gx * gx - P.coeffU * (dux * dux);
    S[1] = 4.0 * Phi * Hyy + 2.0 * gy * gy - P.coeffU * (duy * duy);
    S[2] = 4.0 * Phi * Hzz + 2.0 * gz * gz - P.coeffU * (duz * duz);
    S[3] = 4.0 * Phi * Hxy + 2.0 * gx * gy - P.coeffU * (dux * duy);
    S[4] = 4.0 * Phi * Hxz + 2.0 * gx * gz - P.coeffU * (dux * duz);
    S[5] = 4.0 * Phi * Hyz + 2.0 * gy * gz - P.coeffU * (duy * duz);
}

#pragma clang fp contract(fast)

int post_row_elems(int N);
unsigned post_row_blocks(int N, int E, int PW, unsigned Rmax);
// rows: the row tensor kernel even where the tensor march takes the lattice
hipError_t post_step(const PostState &S, int N, int logN, const PostStep &P, hipStream_t s, bool rows = false);
hipError_t post_drift(const PostState &S, size_t gs, double d, hipStream_t s);
// serial sums of g1 (and g2 when non-null) in lattice order into out[0], out[1]
hipError_t serial_sums(const double *g1, const double *g2, size_t gs, double *out, hipStream_t s);
hipError_t post_phi(double *f, const double *src, size_t gs, double P, double mean, double den, int divide,
                    hipStream_t s);
// {sum f, sum fd, sum f^2} in fixed order (work: >= 3072 doubles)
hipError_t post_sums3(const double *f, const double *fd, size_t gs, double *work, double *out, hipStream_t s);

// Diagnostics (ie_diag.hip).  Bispectrum shell sums {re, im, count} x numbins
// of the half spectrum in a plan slot (the caller zeroes its k = 0 mode).
struct FftPlan;
hipError_t bispectrum_sums(const FftPlan *plan, int slot, int N, double *out_host, hipStream_t s);
// LOG map of f into dn with its masked histogram counts (counts_dev: nbins
// uint64, device); f, fd are float fluctuations about fbar, pbar (split) or
// doubles.  work: >= 4096 doubles.
struct LogMapResult {
    double factt, fmean, bmin, df;
    int valid;
};
hipError_t log_map(const void *f, const void *fd, int split, size_t gs, double fbar, double pbar, double a,
                   double ad, double s_rescale, double eta, int nbins, double *dn, double *work,
                   unsigned long long *counts_dev, LogMapResult *res, hipStream_t s);
// float store of a double field (split storage after a k-space filter)
hipError_t store_float(const double *src, float *dst, size_t gs, hipStream_t s);

// RK4 / RK45 inflation steps on FP64 storage (ie_rk.hip).  Stage buffers are
// allocated once per run (stages x (2 N^3 doubles + 12 N^3 floats with tensors)).
struct RkWork;
hipError_t rk_create(int N, int logN, int sigw, int stages, RkWork **out);
void rk_free(RkWork *w);
// post = 1: later steps use the post-inflation right-hand side with this w
// (src/evolution.cpp:520-592), as the reference's ScopedPostInflationRhsMode
void rk_set_post(RkWork *w, int post, double omega);
hipError_t rk4_step(RkWork *w, const Pot &P, DevBg *bg, DevBg *bg_host, double *f, double *fd, float *const *h,
                    float *const *hd, double step, hipStream_t s);
hipError_t rk45_step(RkWork *w, const Pot &P, DevBg *bg, DevBg *bg_host, double *f, double *fd, float *const *h,
                     float *const *hd, double step, double abs_tol, double rel_tol, double safety, int *accepted,
                     double *factor, double *err_norm, hipStream_t s);

// The message ie_last_error(NULL) returns on the calling thread (entry points
// without a run handle).
void set_unbound_error(const std::string &msg);

}  // namespace ie
