// ie_deltaN.hip -- separate-universe deltaN map on gfx950.
//
// The deltaN phase of the reference (run_deltaN_loop, src/evolution.cpp:1291-1375)
// integrates an independent ODE per site in e-fold time:
//   pi' = -(3 - pi^2/2)(pi + V'/V),  phi' = pi,  deltaN' = 1   while the site is
// short of the phiref surface (src/evolution.cpp:1004-1037).  There is no
// neighbour coupling, so k_deltaN_steps keeps (phi, pi, deltaN) of a site in
// registers for all steps: one read and one write of the state per launch instead
// of two lattice sweeps per step.  A site that reaches the surface stops: its phi
// and deltaN never change again in the reference either, while its pi is still
// kicked every step with the ratio V'/V frozen at its final phi, as the reference
// does (evolve_derivsN runs over every site); the post-inflation phase starts
// from those velocities.
//
// The site arithmetic keeps the reference's operation order with FP contraction
// off, so the map equals the reference's bit for bit given the same input state.
// The tabulated potential is staged in LDS once per workgroup and searched with
// a per-site hint (the register analogue of lstart, src/potential.cpp:40-42); the
// bracket found is the unique l with x[l-1] >= phi > x[l], as in the reference.
#include <math.h>

#include "ie_internal.h"
#include "ie_numeric.h"

namespace ie {

namespace {

constexpr int kMaxTable = 2048;   // table rows staged in dynamic LDS (4 x 2048 doubles = 64 KB)

struct DnParams {
    int mono;          // 1 monotonic, -1 antimonotonic, 0 generic
    double phiref, Vref;
    Pot P;
};

#pragma clang fp contract(off)

// The reference's expressions (src/potential.cpp:77-83), their division by B^2
// through the run's reciprocal of B^2 (div_by: the same correctly rounded result).
__device__ __forceinline__ double hill_V(const Pot &P, double v)
{
    return div_by(P.V0 * (1.0 - (1.0 - P.ns) / 4.0 * (v * v)), P.B2, P.invB2);
}
__device__ __forceinline__ double hill_dV(const Pot &P, double v)
{
    return div_by(-P.V0 * (1.0 - P.ns) / 2.0 * v, P.B2, P.invB2);
}

// The table as a kernel reads it: x, V, V' and the reciprocals of the interval
// widths x[l-1] - x[l] (rdx[l], or null: divide).
struct DnTable {
    const double *x, *v, *dv, *rdx;
    int n;
};

// A table of up to kMaxTable rows staged in the workgroup's dynamic LDS (4 n
// doubles, dn_table_lds); a longer one read from global memory.  Every thread of
// the workgroup calls it (a uniform barrier when staging).
__device__ __forceinline__ DnTable stage_table(const DnParams &D)
{
    extern __shared__ double dn_tab[];
    const int n = D.P.numerical ? D.P.n : 0;
    DnTable T{D.P.x, D.P.v, D.P.dv, nullptr, n};
    if (n > 0 && n <= kMaxTable) {
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            dn_tab[e] = D.P.x[e];
            dn_tab[n + e] = D.P.v[e];
            dn_tab[2 * n + e] = D.P.dv[e];
            dn_tab[3 * n + e] = e > 0 ? 1.0 / (D.P.x[e - 1] - D.P.x[e]) : 0.0;
        }
        __syncthreads();
        T = DnTable{dn_tab, dn_tab + n, dn_tab + 2 * n, dn_tab + 3 * n, n};
    }
    return T;
}

__host__ inline size_t dn_table_lds(const Pot &P)
{
    return P.numerical && P.n > 0 && P.n <= kMaxTable ? (size_t)4 * P.n * sizeof(double) : 0;
}

__device__ __forceinline__ DnTable global_table(const DnParams &D)
{
    return DnTable{D.P.x, D.P.v, D.P.dv, nullptr, D.P.numerical ? D.P.n : 0};
}

// Hinted bracket walk: returns l with x[l-1] >= v > x[l].
__device__ __forceinline__ int walk(const double *x, int n, double v, int l)
{
    if (l < 1) l = 1;
    if (l > n - 1) l = n - 1;
    while (l < n - 1 && x[l] >= v) ++l;
    while (l > 1 && x[l - 1] < v) --l;
    return l;
}

// interpolate_from_table (src/potential.cpp:52-68): y1 + ((v - x1)(y0 - y1)) / (x0 - x1),
// the division through the interval's reciprocal when staged.
__device__ __forceinline__ double interp(const DnTable &T, const double *y, double v, int l)
{
    const double x0 = T.x[l - 1], x1 = T.x[l], y0 = y[l - 1], y1 = y[l];
    const double num = (v - x1) * (y0 - y1), dx = x0 - x1;
    return y1 + (T.rdx ? div_by(num, dx, T.rdx[l]) : num / dx);
}

// V / B^2 and V' / B^2 of the table at v in interval l
__device__ __forceinline__ double tab_V(const DnParams &D, const DnTable &T, double v, int l)
{
    return div_by(interp(T, T.v, v, l), D.P.B2, D.P.invB2);
}
__device__ __forceinline__ double tab_dV(const DnParams &D, const DnTable &T, double v, int l)
{
    return div_by(interp(T, T.dv, v, l), D.P.B2, D.P.invB2);
}

__device__ __forceinline__ bool active(const DnParams &D, const DnTable &T, double v, int *hint)
{
    if (D.mono > 0) return fabs(v) > fabs(D.phiref);
    if (D.mono < 0) return fabs(v) < fabs(D.phiref);
    double V;
    if (D.P.numerical) {
        *hint = walk(T.x, T.n, v, *hint);
        V = tab_V(D, T, v, *hint);
    } else {
        V = hill_V(D.P, v);
    }
    return V > D.Vref;
}

__device__ __forceinline__ bool in_table(const DnTable &T, double v) { return !(v > T.x[0] || v < T.x[T.n - 1]); }

__global__ __launch_bounds__(kBlock) void k_deltaN_steps(double *__restrict__ f, double *__restrict__ fd,
                                                          double *__restrict__ dn, size_t gs, DnParams D,
                                                          double dN, long long nsteps, int *error)
{
    const DnTable T = stage_table(D);
    const int n = T.n;
    int err = 0;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        double phi = f[id], pi = fd[id], N = dn[id];
        int hint = n / 2;
        if (D.P.numerical && in_table(T, phi)) {
            int lo = 0, hi = n - 1;   // initial bracket by bisection
            while (hi - lo > 1) {
                const int mid = lo + (hi - lo) / 2;
                if (T.x[mid] >= phi) lo = mid;
                else hi = mid;
            }
            hint = hi;
        }
        int ghint = hint;
        bool act = active(D, T, phi, &ghint);
        double ratio = 0.0;
        bool frozen = false;   // phi of a finished site no longer moves: its ratio is fixed
        for (long long s = 0; s < nsteps; ++s) {
            // evolve_derivsN (src/evolution.cpp:1004-1012) on every site, finished
            // or not, with pot_ratio (src/potential.cpp:177-192)
            if (!frozen) {
                if (D.P.numerical) {
                    if (!in_table(T, phi)) {
                        err = 1;
                        break;
                    }
                    hint = walk(T.x, n, phi, hint);
                    ratio = interp(T, T.dv, phi, hint) / interp(T, T.v, phi, hint);
                } else {
                    ratio = hill_dV(D.P, phi) / hill_V(D.P, phi);
                }
                frozen = !act;
            }
            pi += dN * (-(3.0 - 0.5 * (pi * pi)) * (pi + ratio));
            // evolve_fieldsN (src/evolution.cpp:1015-1037): active sites only
            if (act) {
                N += dN;
                phi += dN * pi;
                act = active(D, T, phi, &ghint);
            }
        }
        f[id] = phi;
        fd[id] = pi;
        dn[id] = N;
    }
    if (__any(err) && (threadIdx.x & 63) == 0) atomicOr(error, 1);
}

// RHS of the deltaN system at one state (compute_deltaN_rhs,
// src/evolution.cpp:1065-1113): (dphi, dpi, dN)/dN-time, with the activity test
// at that state.
__device__ __forceinline__ void dn_rhs(const DnParams &D, const DnTable &T, double phi, double pi, int *hint,
                                       int *ghint, double *kf, double *kfd, double *kdn, int *err)
{
    double pot, dpot;
    if (D.P.numerical) {
        if (!in_table(T, phi)) {
            *err = 1;
            pot = 1.0;
            dpot = 0.0;
        } else {
            *hint = walk(T.x, T.n, phi, *hint);
            pot = tab_V(D, T, phi, *hint);
            dpot = tab_dV(D, T, phi, *hint);
        }
    } else {
        pot = hill_V(D.P, phi);
        dpot = hill_dV(D.P, phi);
    }
    const double ratio = dpot / pot;
    bool act;
    if (D.mono > 0) act = fabs(phi) > fabs(D.phiref);
    else if (D.mono < 0) act = fabs(phi) < fabs(D.phiref);
    else act = pot > D.Vref;
    (void)ghint;
    *kfd = -(3.0 - 0.5 * (pi * pi)) * (pi + ratio);
    if (act) {
        *kf = pi;
        *kdn = 1.0;
    } else {
        *kf = 0.0;
        *kdn = 0.0;
    }
}

// Classical RK4 of the deltaN system, nsteps in registers per site
// (rk4_step_deltaN, src/evolution.cpp:1176-1189 with the stage states of
// :1115-1132 and the update of :1146-1174; zero Butcher entries skipped as there).
__global__ __launch_bounds__(kBlock) void k_deltaN_rk4(double *__restrict__ f, double *__restrict__ fd,
                                                        double *__restrict__ dn, size_t gs, DnParams D, double h,
                                                        long long nsteps, int *error)
{
    const DnTable T = stage_table(D);
    const int n = T.n;
    const double A[4][4] = {{0, 0, 0, 0}, {0.5, 0, 0, 0}, {0, 0.5, 0, 0}, {0, 0, 1.0, 0}};
    const double B[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    int err = 0;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        double phi = f[id], pi = fd[id], N = dn[id];
        int hint = n > 0 ? n / 2 : 0, ghint = hint;
        for (long long s = 0; s < nsteps; ++s) {
            double kf[4], kfd[4], kdn[4];
            dn_rhs(D, T, phi, pi, &hint, &ghint, &kf[0], &kfd[0], &kdn[0], &err);
#pragma unroll
            for (int st = 1; st < 4; ++st) {
                double sf = 0.0, sfd = 0.0, sdn = 0.0;
#pragma unroll
                for (int q = 0; q < st; ++q) {
                    const double c = A[st][q];
                    if (c == 0.0) continue;
                    sf += c * kf[q];
                    sfd += c * kfd[q];
                    sdn += c * kdn[q];
                }
                (void)sdn;
                dn_rhs(D, T, phi + h * sf, pi + h * sfd, &hint, &ghint, &kf[st], &kfd[st], &kdn[st],
                       &err);
            }
            double sf = 0.0, sfd = 0.0, sdn = 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                sf += B[q] * kf[q];
                sfd += B[q] * kfd[q];
                sdn += B[q] * kdn[q];
            }
            phi += h * sf;
            pi += h * sfd;
            N += h * sdn;
        }
        f[id] = phi;
        fd[id] = pi;
        dn[id] = N;
    }
    if (__any(err) && (threadIdx.x & 63) == 0) atomicOr(error, 1);
}

// One Dormand-Prince 5(4) trial step of the deltaN system (rk45_step_deltaN,
// src/evolution.cpp:1191-1267).  The system is per site with one global step, so
// the seven stages stay in registers: the fifth-order state goes to (f5, fd5, dn5)
// and each workgroup writes its partial of the weighted squared error
// sum (e / (abs_tol + rel_tol max(|y|, |y5|)))^2 over (phi, pi, N).
__global__ __launch_bounds__(kBlock) void k_deltaN_rk45_trial(
    const double *__restrict__ f, const double *__restrict__ fd, const double *__restrict__ dn,
    double *__restrict__ f5, double *__restrict__ fd5, double *__restrict__ dn5, size_t gs, DnParams D, double h,
    double abs_tol, double rel_tol, double *__restrict__ partials, int *error)
{
    __shared__ double smem[kWaves];
    const DnTable T = stage_table(D);
    const int n = T.n;
    const double A[7][7] = {{0, 0, 0, 0, 0, 0, 0},
                            {1.0 / 5.0, 0, 0, 0, 0, 0, 0},
                            {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0, 0},
                            {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0, 0},
                            {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0, 0},
                            {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0, 0},
                            {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0}};
    const double B[7] = {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0};
    const double E[7] = {35.0 / 384.0 - 5179.0 / 57600.0,     0,
                         500.0 / 1113.0 - 7571.0 / 16695.0,   125.0 / 192.0 - 393.0 / 640.0,
                         -2187.0 / 6784.0 + 92097.0 / 339200.0, 11.0 / 84.0 - 187.0 / 2100.0,
                         -1.0 / 40.0};
    double acc = 0.0;
    int err = 0;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        const double phi = f[id], pi = fd[id], N = dn[id];
        int hint = n > 0 ? n / 2 : 0, ghint = hint;
        double kf[7], kfd[7], kdn[7];
        dn_rhs(D, T, phi, pi, &hint, &ghint, &kf[0], &kfd[0], &kdn[0], &err);
#pragma unroll
        for (int st = 1; st < 7; ++st) {
            double sf = 0.0, sfd = 0.0;
#pragma unroll
            for (int q = 0; q < st; ++q) {
                const double c = A[st][q];
                if (c == 0.0) continue;
                sf += c * kf[q];
                sfd += c * kfd[q];
            }
            dn_rhs(D, T, phi + h * sf, pi + h * sfd, &hint, &ghint, &kf[st], &kfd[st], &kdn[st], &err);
        }
        double y5f = phi, y5fd = pi, y5dn = N, ef = 0.0, efd = 0.0, edn = 0.0;
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            if (B[q] != 0.0) {
                y5f += h * B[q] * kf[q];
                y5fd += h * B[q] * kfd[q];
                y5dn += h * B[q] * kdn[q];
            }
            if (E[q] != 0.0) {
                ef += h * E[q] * kf[q];
                efd += h * E[q] * kfd[q];
                edn += h * E[q] * kdn[q];
            }
        }
        f5[id] = y5f;
        fd5[id] = y5fd;
        dn5[id] = y5dn;
        const double s1 = abs_tol + rel_tol * fmax(fabs(phi), fabs(y5f));
        const double s2 = abs_tol + rel_tol * fmax(fabs(pi), fabs(y5fd));
        const double s3 = abs_tol + rel_tol * fmax(fabs(N), fabs(y5dn));
        acc += (ef / s1) * (ef / s1) + (efd / s2) * (efd / s2) + (edn / s3) * (edn / s3);
    }
    if (__any(err) && (threadIdx.x & 63) == 0) atomicOr(error, 1);
    double v[1] = {acc};
    block_sum<1>(v, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

// evolve_fieldsN(d) alone (src/evolution.cpp:1015-1037).
__global__ __launch_bounds__(kBlock) void k_fieldsN(double *__restrict__ f, const double *__restrict__ fd,
                                                     double *__restrict__ dn, size_t gs, DnParams D, double d)
{
    const DnTable T = global_table(D);
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        int hint = T.n / 2;
        const double v = f[id];
        if (active(D, T, v, &hint)) {
            dn[id] += d;
            f[id] = v + d * fd[id];
        }
    }
}

// initializeN (src/initialize.cpp:219-232): fd -> fd a^{s-1} / sqrt(V/3), dn = 0.
__global__ __launch_bounds__(kBlock) void k_initN(const double *__restrict__ f, double *__restrict__ fd,
                                                   double *__restrict__ dn, size_t gs, Pot P, double apow,
                                                   int *error)
{
    int err = 0;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        const double v = f[id];
        double V;
        if (P.numerical) {
            if (v > P.x[0] || v < P.x[P.n - 1]) {
                err = 1;
                V = 1.0;
            } else {
                int lo = 0, hi = P.n - 1;
                while (hi - lo > 1) {
                    const int mid = lo + (hi - lo) / 2;
                    if (P.x[mid] >= v) lo = mid;
                    else hi = mid;
                }
                const DnTable T{P.x, P.v, P.dv, nullptr, P.n};
                V = div_by(interp(T, P.v, v, hi), P.B2, P.invB2);
            }
        } else {
            V = hill_V(P, v);
        }
        const double H = sqrt(V / 3.0);
        fd[id] = fd[id] * apow / H;
        dn[id] = 0.0;
    }
    if (__any(err) && (threadIdx.x & 63) == 0) atomicOr(error, 1);
}

#pragma clang fp contract(fast)

// Ordered search of get_phiref (src/evolution.cpp:1271-1288): key per site
// (|f| for monotonic/antimonotonic, V(f) otherwise), best key with the lowest
// index wins, which is the value the serial strict-improvement scan returns.
struct Best {
    double key;
    unsigned long long idx;
};

__device__ __forceinline__ bool better(int mono, double a, unsigned long long ia, double b, unsigned long long ib)
{
    if (mono < 0) return a > b || (a == b && ia < ib);
    return a < b || (a == b && ia < ib);
}

__global__ __launch_bounds__(kBlock) void k_phiref(const double *__restrict__ f, size_t gs, DnParams D,
                                                    double *__restrict__ pkey, unsigned long long *__restrict__ pidx)
{
    __shared__ double sk[kBlock];
    __shared__ unsigned long long si[kBlock];
    double bk = 0.0;
    unsigned long long bi = ~0ull;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        const double v = f[id];
        double key;
        if (D.mono != 0) key = fabs(v);
        else if (D.P.numerical) {
            int lo = 0, hi = D.P.n - 1;
            while (hi - lo > 1) {
                const int mid = lo + (hi - lo) / 2;
                if (D.P.x[mid] >= v) lo = mid;
                else hi = mid;
            }
            key = tab_V(D, global_table(D), v, hi);
        } else key = hill_V(D.P, v);
        if (bi == ~0ull || better(D.mono, key, id, bk, bi)) {
            bk = key;
            bi = id;
        }
    }
    sk[threadIdx.x] = bk;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            const int o = threadIdx.x + off;
            if (si[o] != ~0ull && (si[threadIdx.x] == ~0ull || better(D.mono, sk[o], si[o], sk[threadIdx.x], si[threadIdx.x]))) {
                sk[threadIdx.x] = sk[o];
                si[threadIdx.x] = si[o];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pkey[blockIdx.x] = sk[0];
        pidx[blockIdx.x] = si[0];
    }
}

__global__ void k_phiref_final(const double *__restrict__ f, const double *__restrict__ pkey,
                               const unsigned long long *__restrict__ pidx, int nb, int mono, double *out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double bk = 0.0;
    unsigned long long bi = ~0ull;
    for (int b = 0; b < nb; ++b) {
        if (pidx[b] == ~0ull) continue;
        if (bi == ~0ull || better(mono, pkey[b], pidx[b], bk, bi)) {
            bk = pkey[b];
            bi = pidx[b];
        }
    }
    *out = f[bi];
}

// saveN's mean (src/output.cpp:1527-1537): a float accumulator walked in lattice
// order over sites with deltaN <= cutoff, acc = (float)((double)acc + v), then
// divided by N^3 in float.  The order of the roundings is part of the result,
// and it is kept exactly without a serial walk:
//   - while acc lies in one binade [B, 2B) it is k u with u = B 2^-23 and k an
//     integer; a step with 0 <= v < B adds the integer n = rint(r / u), where
//     r = fl64(B + v) - B is v on the double grid of that binade (the double
//     rounding of acc + v, whose parity acc's multiple of u cannot change), as
//     long as k + n < 2^24 and r / u is not a tie (its rounding would depend on
//     k's parity);
//   - a tie (r / u = f + 1/2) rounds k + f + 1/2 to the even one of k + f and
//     k + f + 1: k -> k + f + ((k + f) & 1), which depends on k only through its
//     parity;
//   - so the steps of a binade compose: a run of steps maps k to
//     k + A + (t ? (k + c) & 1 : 0), with A an integer, t whether the run holds
//     a tie and c a parity bit, and two runs compose into a run of the same
//     form (after a run's first tie every later tie's parity is known); one
//     workgroup scans the sites a chunk at a time with this composition and
//     stops at the first step that is not of this kind (it leaves the binade,
//     or has v < 0 or v >= B), which it takes as a float step, then goes on in
//     the new binade (deltaN maps hold values on the dN grid many times over,
//     and a tie used to stop the scan: 1706 stops on a 128^3 tabulated map, 20
//     now); every step of a binade adds n >= 0, so k only grows, and only a
//     chunk with a bad site or whose last k reaches 2^24 is searched for the
//     first stop, partial sum by partial sum;
//   - while acc is 0, negative or not finite, excluded sites (v > cutoff or NaN)
//     leave it alone, so do zero sites while acc is a zero of their sign or +0,
//     and every other included one is a float step.
// The result is the serial walk's bit for bit.
constexpr int kMeanThreads = 1024, kMeanPer = 8, kMeanChunk = kMeanThreads * kMeanPer;

__device__ __forceinline__ float mean_step(float acc, double v, double cutoff)
{
    return v <= cutoff ? (float)((double)acc + v) : acc;
}

__global__ __launch_bounds__(kMeanThreads) void k_float_mean(const double *__restrict__ dn, size_t gs, double cutoff,
                                                              float *out)
{
    constexpr int NW = kMeanThreads / 64;
    constexpr long long kTop = 1ll << 24;
    constexpr int kBadWave = 4;   // flag bit of a wave total (grid_compose reads bits 0-1 only)
    // Chunks alternate between two copies of the wave totals, so a copy is
    // free again one barrier after its last read.
    __shared__ GridRun wtot[2][NW];
    __shared__ unsigned long long s_first;   // ~0 outside a stop's handling
    __shared__ long long s_kfirst;
    __shared__ float s_acc;
    __shared__ unsigned long long s_pos;
    const unsigned tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) {
        s_acc = 0.f;
        s_pos = 0;
        s_first = ~0ull;
    }
    __syncthreads();
    auto load = [&](double (&d)[kMeanPer], size_t b) {
#pragma unroll
        for (int q = 0; q < kMeanPer; ++q) d[q] = b + q < gs ? dn[b + q] : 0.0;
    };
    unsigned p = 0;
    for (;;) {
        const float acc = s_acc;
        const size_t pos = s_pos;
        if (pos >= gs) break;
        // the phase: acc's binade (regular), or a hold (acc 0, negative, not finite)
        const bool regular = acc >= 1.17549435e-38f && acc <= 3.40282347e38f;
        int E = 0;
        if (regular) (void)frexpf(acc, &E);
        // u = 2^(E-24) and 1/u are powers of two in the normal range, so r / u = r * iu
        const double B = regular ? ldexp(1.0, E - 1) : 0.0, u = regular ? ldexp(1.0, E - 24) : 0.0;
        const double iu = regular ? ldexp(1.0, 24 - E) : 0.0;
        long long k = regular ? (long long)((double)acc / u) : 0;
        size_t c0 = pos;
        double cur[kMeanPer];
        load(cur, c0 + (size_t)tid * kMeanPer);
        for (;;) {
            const size_t base = c0 + (size_t)tid * kMeanPer;
            double nxt[kMeanPer];   // the next chunk's values, in flight during this chunk's scan
            load(nxt, base + kMeanChunk);
            // site q's run (a step, a tie or nothing) and whether it stops the
            // scan; computed again where a chunk holds a stop rather than kept
            auto elem = [&](int q, bool &bad) -> GridRun {
                const size_t id = base + q;
                bad = id >= gs;
                if (id >= gs) return GridRun{0, 0};
                const double v = cur[q];
                if (!(v <= cutoff)) return GridRun{0, 0};
                if (!regular || !(v >= 0.0) || !(v < B)) {
                    // a zero added to a zero accumulator leaves it as it is
                    // (+0 + 0 = +0, -0 + -0 = -0): no step, so that a long run
                    // of zero sites at the head of the map is not walked one
                    // float step at a time
                    bad = !(acc == 0.0f && v == 0.0 && !(signbit(acc) && !signbit(v)));
                    return GridRun{0, 0};
                }
                const double qv = ((B + v) - B) * iu;   // in [0, 2^23]: int conversions
                const double fl = floor(qv);
                return qv - fl == 0.5 ? grid_tie((int)fl) : grid_step((int)rint(qv));
            };
            GridRun seg{0, 0};
            bool anybad = false;
#pragma unroll
            for (int q = 0; q < kMeanPer; ++q) {
                bool bad;
                seg = grid_compose(seg, elem(q, bad));
                anybad |= bad;
            }
            // inclusive scan of the threads' runs: in the wave (plain integer
            // sums when none of its runs holds a tie), then over waves
            GridRun incl = seg;
            if (__ballot(seg.f & 1) == 0) {
                long long a = seg.A;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const long long o = __shfl_up(a, d);
                    if ((int)lane >= d) a += o;
                }
                incl.A = a;
            } else {
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    GridRun o;
                    o.A = __shfl_up(incl.A, d);
                    o.f = __shfl_up(incl.f, d);
                    if ((int)lane >= d) incl = grid_compose(o, incl);
                }
            }
            const bool wave_bad = __ballot(anybad) != 0;
            if (lane == 63) wtot[p][wave] = GridRun{incl.A, incl.f | (wave_bad ? kBadWave : 0)};
            __syncthreads();
            // the waves' totals scanned once per wave, in its first NW lanes
            GridRun wt = lane < (unsigned)NW ? wtot[p][lane] : GridRun{0, 0};
            const bool chunk_bad = __ballot(lane < (unsigned)NW && (wt.f & kBadWave)) != 0;
#pragma unroll
            for (int d = 1; d < NW; d <<= 1) {
                GridRun o;
                o.A = __shfl_up(wt.A, d);
                o.f = __shfl_up(wt.f, d);
                if ((int)lane >= d) wt = grid_compose(o, wt);
            }
            GridRun total;
            total.A = __shfl(wt.A, NW - 1);
            total.f = __shfl(wt.f, NW - 1);
            const long long kend = grid_apply(total, k);
            p ^= 1;
            // every step of a binade adds n >= 0, so k only grows: a chunk with
            // no bad site that ends below 2^24 holds no stop
            if (!chunk_bad && !(regular && kend >= kTop)) {
                k = kend;
                c0 += kMeanChunk;
#pragma unroll
                for (int q = 0; q < kMeanPer; ++q) cur[q] = nxt[q];
                continue;
            }
            // the chunk holds a stop: find the first one and its k
            GridRun before, excl;
            before.A = __shfl(wt.A, wave == 0 ? 0 : wave - 1);
            before.f = __shfl(wt.f, wave == 0 ? 0 : wave - 1);
            if (wave == 0) before = GridRun{0, 0};
            excl.A = __shfl_up(incl.A, 1);
            excl.f = __shfl_up(incl.f, 1);
            if (lane == 0) excl = GridRun{0, 0};
            long long K = grid_apply(grid_compose(before, excl), k);
            size_t mine = ~(size_t)0;
            long long kmine = 0;
#pragma unroll
            for (int q = 0; q < kMeanPer; ++q) {
                bool bad;
                const long long Kq = grid_apply(elem(q, bad), K);
                const bool stop = bad || (regular && Kq >= kTop);
                if (stop && mine == ~(size_t)0) {
                    mine = base + q;
                    kmine = K;
                }
                K = Kq;
            }
            if (mine != ~(size_t)0) atomicMin(&s_first, (unsigned long long)mine);
            __syncthreads();
            const unsigned long long first = s_first;
            if (mine == first) s_kfirst = kmine;
            __syncthreads();
            if (tid == 0) {
                const float at = regular ? (float)((double)s_kfirst * u) : acc;
                if (first < gs) {
                    s_acc = mean_step(at, dn[first], cutoff);
                    s_pos = first + 1;
                } else {
                    s_acc = at;
                    s_pos = gs;
                }
                s_first = ~0ull;
            }
            __syncthreads();
            break;
        }
    }
    if (tid == 0) *out = s_acc / (float)gs;
}

// zero = 0: deltaN -= Nmean (src/output.cpp:1539-1542); zero = 1: sites above
// cutoff - Nmean set to 0 (src/output.cpp:1549-1553).
__global__ __launch_bounds__(kBlock) void k_zeta(double *__restrict__ dn, size_t gs, const float *nmean,
                                                  double cutoff, int zero)
{
    const double m = (double)*nmean;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        if (zero) {
            if (dn[id] > (cutoff - m)) dn[id] = 0;
        } else {
            dn[id] -= m;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_sub(double *__restrict__ g, size_t gs, const double *mean_src, double inv)
{
    const double m = mean_src[0] * inv;
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x)
        g[id] -= m;
}

// Plain FP64 sum of a field (fixed order: per-block tree, then block order).
__global__ __launch_bounds__(kBlock) void k_sum(const double *__restrict__ g, size_t gs, double *partials)
{
    __shared__ double smem[kWaves];
    double acc[1] = {0.0};
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x)
        acc[0] += g[id];
    block_sum<1>(acc, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

__global__ void k_sum_final(const double *partials, int nb, double *out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += partials[b];
    *out = s;
}

// min / max (src/output.cpp:1095-1101) and histogram (src/output.cpp:1104-1123).
__global__ __launch_bounds__(kBlock) void k_minmax(const double *__restrict__ g, size_t gs, double *pmin, double *pmax)
{
    __shared__ double smn[kBlock], smx[kBlock];
    double mn = g[0], mx = g[0];
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        mn = g[id] < mn ? g[id] : mn;
        mx = g[id] > mx ? g[id] : mx;
    }
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            smn[threadIdx.x] = fmin(smn[threadIdx.x], smn[threadIdx.x + off]);
            smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pmin[blockIdx.x] = smn[0];
        pmax[blockIdx.x] = smx[0];
    }
}

__global__ __launch_bounds__(kBlock) void k_histogram(const double *__restrict__ g, size_t gs, double bmin,
                                                       double bmax, double df, int nbins,
                                                       unsigned long long *counts)
{
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        const double v = g[id];
        int b = (int)((v - bmin) / df);
        if (v == bmax) b = nbins - 1;
        if (b >= 0 && b < nbins) atomicAdd(&counts[b], 1ull);
    }
}

int grid_for(size_t gs)
{
    size_t b = (gs + kBlock - 1) / kBlock;
    if (b > 4096) b = 4096;
    return (int)(b < 1 ? 1 : b);
}

}  // namespace

hipError_t dn_init(const Fields &F, const double *f, double *fd, double *dn, const Pot &P, double apow, int *error,
                   hipStream_t s)
{
    const size_t gs = (size_t)F.N * F.N * F.N;
    k_initN<<<grid_for(gs), kBlock, 0, s>>>(f, fd, dn, gs, P, apow, error);
    return hipGetLastError();
}

static DnParams params(const Pot &P, int mono, double phiref, double Vref)
{
    DnParams D;
    D.mono = mono;
    D.phiref = phiref;
    D.Vref = Vref;
    D.P = P;
    return D;
}

hipError_t dn_fieldsN(size_t gs, double *f, const double *fd, double *dn, const Pot &P, int mono, double phiref,
                      double Vref, double d, hipStream_t s)
{
    k_fieldsN<<<grid_for(gs), kBlock, 0, s>>>(f, fd, dn, gs, params(P, mono, phiref, Vref), d);
    return hipGetLastError();
}

hipError_t dn_steps(size_t gs, double *f, double *fd, double *dn, const Pot &P, int mono, double phiref, double Vref,
                    double dN, long long nsteps, int *error, hipStream_t s)
{
    size_t b = (gs + kBlock - 1) / kBlock;
    k_deltaN_steps<<<(int)b, kBlock, dn_table_lds(P), s>>>(f, fd, dn, gs, params(P, mono, phiref, Vref), dN, nsteps,
                                                           error);
    return hipGetLastError();
}

hipError_t dn_rk4_steps(size_t gs, double *f, double *fd, double *dn, const Pot &P, int mono, double phiref,
                        double Vref, double h, long long nsteps, int *error, hipStream_t s)
{
    size_t b = (gs + kBlock - 1) / kBlock;
    k_deltaN_rk4<<<(int)b, kBlock, dn_table_lds(P), s>>>(f, fd, dn, gs, params(P, mono, phiref, Vref), h, nsteps, error);
    return hipGetLastError();
}

hipError_t dn_rk45_trial(size_t gs, const double *f, const double *fd, const double *dn, double *f5, double *fd5,
                         double *dn5, const Pot &P, int mono, double phiref, double Vref, double h, double abs_tol,
                         double rel_tol, double *work, double *out, int *error, hipStream_t s)
{
    const int nb = grid_for(gs) > 2048 ? 2048 : grid_for(gs);
    k_deltaN_rk45_trial<<<nb, kBlock, dn_table_lds(P), s>>>(f, fd, dn, f5, fd5, dn5, gs, params(P, mono, phiref, Vref), h, abs_tol,
                                               rel_tol, work, error);
    k_sum_final<<<1, 64, 0, s>>>(work, nb, out);
    return hipGetLastError();
}

hipError_t dn_phiref(size_t gs, const double *f, const Pot &P, int mono, double *work, double *out, hipStream_t s)
{
    const int nb = grid_for(gs) > 1024 ? 1024 : grid_for(gs);
    double *pkey = work;
    unsigned long long *pidx = (unsigned long long *)(work + nb);
    k_phiref<<<nb, kBlock, 0, s>>>(f, gs, params(P, mono, 0.0, 0.0), pkey, pidx);
    k_phiref_final<<<1, 64, 0, s>>>(f, pkey, pidx, nb, mono, out);
    return hipGetLastError();
}

hipError_t dn_float_mean(const double *dn, size_t gs, double cutoff, float *out, hipStream_t s)
{
    k_float_mean<<<1, kMeanThreads, 0, s>>>(dn, gs, cutoff, out);
    return hipGetLastError();
}

hipError_t dn_zeta(double *dn, size_t gs, const float *nmean, double cutoff, int zero, hipStream_t s)
{
    k_zeta<<<grid_for(gs), kBlock, 0, s>>>(dn, gs, nmean, cutoff, zero);
    return hipGetLastError();
}

hipError_t field_sum(const double *g, size_t gs, double *work, double *out, hipStream_t s)
{
    const int nb = grid_for(gs) > 1024 ? 1024 : grid_for(gs);
    k_sum<<<nb, kBlock, 0, s>>>(g, gs, work);
    k_sum_final<<<1, 64, 0, s>>>(work, nb, out);
    return hipGetLastError();
}

hipError_t field_center(double *g, size_t gs, double *work, double *sum_out, hipStream_t s)
{
    hipError_t e = field_sum(g, gs, work, sum_out, s);
    if (e != hipSuccess) return e;
    k_sub<<<grid_for(gs), kBlock, 0, s>>>(g, gs, sum_out, 1.0 / (double)gs);
    return hipGetLastError();
}

hipError_t field_minmax(const double *g, size_t gs, double *work, hipStream_t s, int *nb_out)
{
    const int nb = grid_for(gs) > 1024 ? 1024 : grid_for(gs);
    k_minmax<<<nb, kBlock, 0, s>>>(g, gs, work, work + nb);
    *nb_out = nb;
    return hipGetLastError();
}

hipError_t field_histogram(const double *g, size_t gs, double bmin, double bmax, double df, int nbins,
                           unsigned long long *counts, hipStream_t s)
{
    k_histogram<<<grid_for(gs), kBlock, 0, s>>>(g, gs, bmin, bmax, df, nbins, counts);
    return hipGetLastError();
}

}  // namespace ie
