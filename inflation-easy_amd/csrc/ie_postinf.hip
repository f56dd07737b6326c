// ie_postinf.hip -- post-inflation (fluid era) steps of the gfx950 engine.
//
// After inflation and the deltaN map the reference evolves the Bardeen potential
// Phi in a fluid of equation of state w, and the tensors it sources
// (src/evolution.cpp:1378-1651, src/initialize.cpp:246-291):
//   Phi'' = w laplnorm lapl Phi - (3 (1 + w) + s) (a'/a) Phi'
//   h_ij'' = laplnorm lapl h_ij - (2 + s) (a'/a) h_ij' + 2 a^{-2s} S_ij
//   S_ij = 4 Phi d_i d_j Phi + 2 d_i Phi d_j Phi - 4/(3(1+w)) d_i U d_j U,  U = Phi'/(a'/a) + Phi
// with a'' = -(s - (1 - 3w)/2) a'^2 / a, a closed equation: no lattice sums enter
// the step, so the host advances (a, a') and hands each launch its scalars.
//
// One leapfrog step (evolve_derivs_post_inflation then evolve_fields) is two
// launches over the FP64 Phi, Phi' and the six FP32 h_ij, h_ij':
//   k_post_kick    Phi' += d (...) from the 7-point Laplacian of Phi   24 B/site
//   k_post_tensor  the tensor kick with S_ij from the old Phi (gradient and
//                  Hessian: 15 points) and the kicked Phi' (gradient), then the
//                  drifts of h_ij and Phi                              120 B/site
// Phi, Phi' and h_ij are double-buffered (neighbours are read while the new value
// is written); h_ij' is updated in place.  Rows of N >= 64 lattices are walked
// by waves as in the inflation kernels (vector loads, k +- 1 by lane shuffles,
// XCD-aware plane slabs); N < 64 lattices use one site per thread.
//
// The site arithmetic keeps the reference's operation order, with FP contraction
// off, so every value equals the reference build bit for bit; that includes the
// (y, z) entry of the Hessian, which the reference forms from f(i, j+1, k+1) and
// f(i, j-1, k-1) only (src/evolution.cpp:93-106): ((A - B) - A) + B.
#include <math.h>

#include "ie_internal.h"

#pragma clang fp contract(off)

namespace ie {

namespace {

struct PostTensorArgs {
    const double *f;    // Phi before the step (read)
    const double *fd;   // Phi' after the kick (read)
    double *f_out;      // Phi after the drift
    const float *h[6];
    float *h_out[6];
    float *hd[6];       // in place
};

// hd' = hd + d (laplnorm lapl h - (2+s) a' hd / a + srcAmp S), h' = h + dd hd'
__device__ __forceinline__ void post_tensor_update(float hl, float hc, float hv, double S, const PostStep &P,
                                                   float *hd_new, float *h_new)
{
    const double rhs = P.d * (P.laplnorm * (double)hl - P.c2ad * (double)hv / P.a + P.srcamp * S);
    const float hdn = hv + (float)rhs;
    *hd_new = hdn;
    *h_new = hc + (float)(P.dd * (double)hdn);
}

__device__ __forceinline__ double post_kick_value(double fdv, double lap, const PostStep &P)
{
    return fdv + P.d * (P.A * lap - P.c1ad * fdv / P.a);
}

__device__ __forceinline__ void site_ijk(unsigned id, unsigned logN, unsigned m, unsigned &i, unsigned &j, unsigned &k)
{
    k = id & m;
    j = (id >> logN) & m;
    i = id >> (2 * logN);
}

__device__ __forceinline__ unsigned sidx(unsigned i, unsigned j, unsigned k, unsigned logN, unsigned m)
{
    return ((i & m) << (2 * logN)) | ((j & m) << logN) | (k & m);
}

// Row walks of the row kernels: a workgroup of PW waves takes as many
// consecutive planes of its XCD's slab (wave w plane i0 + w), all walking the
// same R rows j0 .. j0 + R - 1 of one 64E-site segment in step, so the i +- 1
// rows a wave reads are, inside the group, the rows its neighbouring waves load
// at about the same time (L2 hits): per plane the group reads its planes plus
// two, per row walk R + 2 rows.  Measured at 256^3: the kick with 8 planes and
// walks of 32 rows moves 1.04x its algorithmic bytes (111 us; 1.17x and 123 us
// with 4 planes of 8 rows).  The row tensor kernel (N >= 512; the tensor march
// below takes N = 64 .. 256) moved 1.59x its bytes at 256^3 with 4 planes of 8
// rows (680 us, against 470 us and 1.07x for the march).
constexpr int kPostKickPlanes = 8, kPostTensorPlanes = 4;
constexpr unsigned kPostKickRows = 32, kPostTensorRows = 8;   // longest walks

template <int PW>
__device__ __forceinline__ RowWalk post_walk(unsigned N, unsigned blk, int E, unsigned R)
{
    const unsigned segs = N / (64u * (unsigned)E), jblocks = N / R;
    const unsigned xcd = blk % kXcd, lb = blk / kXcd;
    const unsigned t = lb / jblocks;
    RowWalk w;
    w.j0 = (lb % jblocks) * R;
    w.k0 = (t % segs) * (64u * (unsigned)E);
    w.i = xcd * (N / kXcd) + (t / segs) * (unsigned)PW + (threadIdx.x >> 6);
    w.R = R;
    w.segs = segs;
    return w;
}

// ---------------------------------------------------------------- site kernels

__global__ __launch_bounds__(kBlock) void k_post_kick_site(const double *__restrict__ f,
                                                            const double *__restrict__ fd,
                                                            double *__restrict__ fd_out, unsigned N, unsigned logN,
                                                            PostStep P)
{
    const unsigned m = N - 1, gs = N * N * N;
    for (unsigned id = blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += gridDim.x * blockDim.x) {
        unsigned i, j, k;
        site_ijk(id, logN, m, i, j, k);
        const double c = f[id];
        const double lap = f[sidx(i, j, k + 1, logN, m)] + f[sidx(i, j, k - 1, logN, m)] +
                           f[sidx(i, j + 1, k, logN, m)] + f[sidx(i, j - 1, k, logN, m)] +
                           f[sidx(i + 1, j, k, logN, m)] + f[sidx(i - 1, j, k, logN, m)] - 6.0 * c;
        fd_out[id] = post_kick_value(fd[id], lap, P);
    }
}

__global__ __launch_bounds__(kBlock) void k_post_tensor_site(PostTensorArgs T, unsigned N, unsigned logN, PostStep P)
{
    const unsigned m = N - 1, gs = N * N * N;
    const double *f = T.f, *fd = T.fd;
    for (unsigned id = blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += gridDim.x * blockDim.x) {
        unsigned i, j, k;
        site_ijk(id, logN, m, i, j, k);
        SiteStencil q;
        q.c = f[id];
        q.fxp = f[sidx(i + 1, j, k, logN, m)];
        q.fxm = f[sidx(i - 1, j, k, logN, m)];
        q.fyp = f[sidx(i, j + 1, k, logN, m)];
        q.fym = f[sidx(i, j - 1, k, logN, m)];
        q.fzp = f[sidx(i, j, k + 1, logN, m)];
        q.fzm = f[sidx(i, j, k - 1, logN, m)];
        q.fxpyp = f[sidx(i + 1, j + 1, k, logN, m)];
        q.fxpym = f[sidx(i + 1, j - 1, k, logN, m)];
        q.fxmyp = f[sidx(i - 1, j + 1, k, logN, m)];
        q.fxmym = f[sidx(i - 1, j - 1, k, logN, m)];
        q.fxpzp = f[sidx(i + 1, j, k + 1, logN, m)];
        q.fxpzm = f[sidx(i + 1, j, k - 1, logN, m)];
        q.fxmzp = f[sidx(i - 1, j, k + 1, logN, m)];
        q.fxmzm = f[sidx(i - 1, j, k - 1, logN, m)];
        q.fypzp = f[sidx(i, j + 1, k + 1, logN, m)];
        q.fymzm = f[sidx(i, j - 1, k - 1, logN, m)];
        q.dc = fd[id];
        q.dxp = fd[sidx(i + 1, j, k, logN, m)];
        q.dxm = fd[sidx(i - 1, j, k, logN, m)];
        q.dyp = fd[sidx(i, j + 1, k, logN, m)];
        q.dym = fd[sidx(i, j - 1, k, logN, m)];
        q.dzp = fd[sidx(i, j, k + 1, logN, m)];
        q.dzm = fd[sidx(i, j, k - 1, logN, m)];
        double S[6];
        post_source(q, P, S);
#pragma unroll 1
        for (int c = 0; c < 6; ++c) {
            const float *h = T.h[c];
            const float hc = h[id];
            const float hl = h[sidx(i, j, k + 1, logN, m)] + h[sidx(i, j, k - 1, logN, m)] +
                             h[sidx(i, j + 1, k, logN, m)] + h[sidx(i, j - 1, k, logN, m)] +
                             h[sidx(i + 1, j, k, logN, m)] + h[sidx(i - 1, j, k, logN, m)] - 6.0f * hc;
            float hdn, hn;
            post_tensor_update(hl, hc, T.hd[c][id], S[c], P, &hdn, &hn);
            T.hd[c][id] = hdn;
            T.h_out[c][id] = hn;
        }
        T.f_out[id] = q.c + P.dd * q.dc;
    }
}

// ----------------------------------------------------------------- row kernels

// kp / km of element e of a row held as E values per lane
template <typename T, int E>
__device__ __forceinline__ T kplus(const T (&c)[E], int e, T kp_last)
{
    return e < E - 1 ? c[e + 1 < E ? e + 1 : e] : kp_last;
}
template <typename T, int E>
__device__ __forceinline__ T kminus(const T (&c)[E], int e, T km_first)
{
    return e > 0 ? c[e > 0 ? e - 1 : 0] : km_first;
}

template <int E>
__global__ __launch_bounds__(64 * kPostKickPlanes) void k_post_kick_rows(const double *__restrict__ f,
                                                                      const double *__restrict__ fd,
                                                                      double *__restrict__ fd_out, unsigned N,
                                                                      unsigned logN, unsigned R, PostStep P)
{
    const unsigned m = N - 1, lane = threadIdx.x & 63;
    const RowWalk w = post_walk<kPostKickPlanes>(N, blockIdx.x, E, R);
    const unsigned P0 = w.i << (2 * logN), Pp = ((w.i + 1) & m) << (2 * logN), Pm = ((w.i - 1) & m) << (2 * logN);
    double prev[E], cur[E], nxt[E];
    ld_row<double, E>(f + P0 + (((w.j0 - 1) & m) << logN) + w.k0, lane, prev);
    ld_row<double, E>(f + P0 + (w.j0 << logN) + w.k0, lane, cur);
    for (unsigned r = 0; r < w.R; ++r) {
        const unsigned j = w.j0 + r, rj = j << logN;
        double up[E], dn[E], v[E], out[E];
        ld_row<double, E>(f + P0 + (((j + 1) & m) << logN) + w.k0, lane, nxt);
        ld_row<double, E>(f + Pp + rj + w.k0, lane, up);
        ld_row<double, E>(f + Pm + rj + w.k0, lane, dn);
        ld_row<double, E>(fd + P0 + rj + w.k0, lane, v);
        double kp_last, km_first;
        k_neighbours<double, E>(cur, f + P0 + rj, w, N, lane, kp_last, km_first);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const double lap = kplus<double, E>(cur, e, kp_last) + kminus<double, E>(cur, e, km_first) + nxt[e] +
                               prev[e] + up[e] + dn[e] - 6.0 * cur[e];
            out[e] = post_kick_value(v[e], lap, P);
        }
        st_row<double, E>(fd_out + P0 + rj + w.k0, lane, out);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            prev[e] = cur[e];
            cur[e] = nxt[e];
        }
    }
}

// Rows of Phi in planes i-1, i, i+1 (index 0, 1, 2) at j-1, j, j+1 (0, 1, 2)
// slide along j in registers; the kicked Phi' rows of plane i and the rows of
// the six h_ij of plane i likewise.
template <int E>
__global__ __launch_bounds__(64 * kPostTensorPlanes) void k_post_tensor_rows(PostTensorArgs T, unsigned N, unsigned logN,
                                                                        unsigned R, PostStep P)
{
    const unsigned m = N - 1, lane = threadIdx.x & 63;
    const RowWalk w = post_walk<kPostTensorPlanes>(N, blockIdx.x, E, R);
    const unsigned pl[3] = {((w.i - 1) & m) << (2 * logN), w.i << (2 * logN), ((w.i + 1) & m) << (2 * logN)};
    const double *f = T.f, *fd = T.fd;
    double F[3][3][E], D[3][E];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        ld_row<double, E>(f + pl[a] + (((w.j0 - 1) & m) << logN) + w.k0, lane, F[a][0]);
        ld_row<double, E>(f + pl[a] + (w.j0 << logN) + w.k0, lane, F[a][1]);
    }
    ld_row<double, E>(fd + pl[1] + (((w.j0 - 1) & m) << logN) + w.k0, lane, D[0]);
    ld_row<double, E>(fd + pl[1] + (w.j0 << logN) + w.k0, lane, D[1]);
    float HJ[6][2][E];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        ld_row<float, E>(T.h[c] + pl[1] + (((w.j0 - 1) & m) << logN) + w.k0, lane, HJ[c][0]);
        ld_row<float, E>(T.h[c] + pl[1] + (w.j0 << logN) + w.k0, lane, HJ[c][1]);
    }
    for (unsigned r = 0; r < w.R; ++r) {
        const unsigned j = w.j0 + r, rj = j << logN, rjp = ((j + 1) & m) << logN, rjm = ((j - 1) & m) << logN;
#pragma unroll
        for (int a = 0; a < 3; ++a) ld_row<double, E>(f + pl[a] + rjp + w.k0, lane, F[a][2]);
        double Dup[E], Ddn[E];
        ld_row<double, E>(fd + pl[1] + rjp + w.k0, lane, D[2]);
        ld_row<double, E>(fd + pl[2] + rj + w.k0, lane, Dup);
        ld_row<double, E>(fd + pl[0] + rj + w.k0, lane, Ddn);
        // k +- 1 of the rows that need them
        double c_kp, c_km, xp_kp, xp_km, xm_kp, xm_km, yp_kp, yp_km, ym_kp, ym_km, d_kp, d_km;
        k_neighbours<double, E>(F[1][1], f + pl[1] + rj, w, N, lane, c_kp, c_km);
        k_neighbours<double, E>(F[2][1], f + pl[2] + rj, w, N, lane, xp_kp, xp_km);
        k_neighbours<double, E>(F[0][1], f + pl[0] + rj, w, N, lane, xm_kp, xm_km);
        k_neighbours<double, E>(F[1][2], f + pl[1] + rjp, w, N, lane, yp_kp, yp_km);
        k_neighbours<double, E>(F[1][0], f + pl[1] + rjm, w, N, lane, ym_kp, ym_km);
        k_neighbours<double, E>(D[1], fd + pl[1] + rj, w, N, lane, d_kp, d_km);
        double S[E][6], fnew[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            SiteStencil q;
            q.c = F[1][1][e];
            q.fxp = F[2][1][e];
            q.fxm = F[0][1][e];
            q.fyp = F[1][2][e];
            q.fym = F[1][0][e];
            q.fzp = kplus<double, E>(F[1][1], e, c_kp);
            q.fzm = kminus<double, E>(F[1][1], e, c_km);
            q.fxpyp = F[2][2][e];
            q.fxpym = F[2][0][e];
            q.fxmyp = F[0][2][e];
            q.fxmym = F[0][0][e];
            q.fxpzp = kplus<double, E>(F[2][1], e, xp_kp);
            q.fxpzm = kminus<double, E>(F[2][1], e, xp_km);
            q.fxmzp = kplus<double, E>(F[0][1], e, xm_kp);
            q.fxmzm = kminus<double, E>(F[0][1], e, xm_km);
            q.fypzp = kplus<double, E>(F[1][2], e, yp_kp);
            q.fymzm = kminus<double, E>(F[1][0], e, ym_km);
            q.dc = D[1][e];
            q.dxp = Dup[e];
            q.dxm = Ddn[e];
            q.dyp = D[2][e];
            q.dym = D[0][e];
            q.dzp = kplus<double, E>(D[1], e, d_kp);
            q.dzm = kminus<double, E>(D[1], e, d_km);
            post_source(q, P, S[e]);
            fnew[e] = q.c + P.dd * q.dc;
        }
        st_row<double, E>(T.f_out + pl[1] + rj + w.k0, lane, fnew);
        (void)yp_km;
        (void)ym_kp;
        // the six components one after the other, each with its j-1 and j rows
        // held from the previous row (HJ[0] is the current component's: the
        // windows rotate by one per component, so the loop indexes no register
        // array); the next component's rows are loaded while one is computed
        float nj[E], ni[E], nm[E], nv[E];
        ld_row<float, E>(T.h[0] + pl[1] + rjp + w.k0, lane, nj);
        ld_row<float, E>(T.h[0] + pl[2] + rj + w.k0, lane, ni);
        ld_row<float, E>(T.h[0] + pl[0] + rj + w.k0, lane, nm);
        ld_row<float, E>(T.hd[0] + pl[1] + rj + w.k0, lane, nv);
#pragma unroll 1
        for (int c = 0; c < 6; ++c) {
            float hjp[E], hip[E], him[E], hv[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                hjp[e] = nj[e];
                hip[e] = ni[e];
                him[e] = nm[e];
                hv[e] = nv[e];
            }
            const int cn = c < 5 ? c + 1 : 5;
            ld_row<float, E>(T.h[cn] + pl[1] + rjp + w.k0, lane, nj);
            ld_row<float, E>(T.h[cn] + pl[2] + rj + w.k0, lane, ni);
            ld_row<float, E>(T.h[cn] + pl[0] + rj + w.k0, lane, nm);
            ld_row<float, E>(T.hd[cn] + pl[1] + rj + w.k0, lane, nv);
            float hkp, hkm;
            k_neighbours<float, E>(HJ[0][1], T.h[c] + pl[1] + rj, w, N, lane, hkp, hkm);
            float hdn[E], hn[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const float hl = kplus<float, E>(HJ[0][1], e, hkp) + kminus<float, E>(HJ[0][1], e, hkm) + hjp[e] +
                                 HJ[0][0][e] + hip[e] + him[e] - 6.0f * HJ[0][1][e];
                double Sc = S[e][0];
#pragma unroll
                for (int q = 1; q < 6; ++q) Sc = c == q ? S[e][q] : Sc;
                post_tensor_update(hl, HJ[0][1][e], hv[e], Sc, P, &hdn[e], &hn[e]);
            }
            st_row<float, E>(T.hd[c] + pl[1] + rj + w.k0, lane, hdn);
            st_row<float, E>(T.h_out[c] + pl[1] + rj + w.k0, lane, hn);
            float keep[2][E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                keep[0][e] = HJ[0][1][e];
                keep[1][e] = hjp[e];
            }
#pragma unroll
            for (int q = 0; q < 5; ++q)
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    HJ[q][0][e] = HJ[q + 1][0][e];
                    HJ[q][1][e] = HJ[q + 1][1][e];
                }
#pragma unroll
            for (int e = 0; e < E; ++e) {
                HJ[5][0][e] = keep[0][e];
                HJ[5][1][e] = keep[1][e];
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                F[a][0][e] = F[a][1][e];
                F[a][1][e] = F[a][2][e];
            }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            D[0][e] = D[1][e];
            D[1][e] = D[2][e];
        }
    }
}

// ------------------------------------------------------ tensor march kernel
//
// The tensor step for N = 64, 128, 256 (one 64E-site segment per row, E = N / 64):
// a workgroup of 8 waves owns rows j0 .. j0 + 7 and marches along i over TI
// planes of its XCD's slab.  Each wave keeps the i-1, i, i+1 rows of Phi, Phi'
// and the six h_ij in registers (the x faces, and the k faces by lane shuffles)
// and puts its rows in LDS, where the neighbouring waves read their j +- 1 rows:
// Phi for planes i-1, i, i+1 (the xy Hessian entry reads Phi(i +- 1, j +- 1)),
// Phi' and h_ij for plane i.  Waves 0 and 7 also bring the tile's halo rows
// j0 - 1 and j0 + 8.  Every row comes from HBM about once (the halo rows are the
// neighbouring tiles' own rows, read at about the same time on the same XCD), so
// the step moves about its 120 B/site, against the row walk's re-reads of the
// i +- 1 planes.  The LDS holds one plane per field (Phi: a ring of three), so a
// plane step has two barriers: after the writes, and after the reads.  The site
// arithmetic is the row kernels' (post_source, post_tensor_update), so the values
// are the same bits.
constexpr int kPostMarchWaves = 8;

template <int E>
struct PostTile {
    double f[3][kPostMarchWaves + 2][64 * E];   // Phi rows of planes i-1, i, i+1 (slot: plane mod 3)
    double fd[kPostMarchWaves + 2][64 * E];     // Phi' rows of plane i
    float h[6][kPostMarchWaves + 2][64 * E];    // h_ij rows of plane i
};

template <typename T, int E>
__device__ __forceinline__ void lds_row_put(T *row, unsigned lane, const T (&v)[E])
{
#pragma unroll
    for (int e = 0; e < E; ++e) row[lane * E + e] = v[e];
}

// The values past the lane's ends of a whole row held E per lane: k = 64E wraps
// to k = 0 (lane 63 to lane 0).
template <typename T, int E>
__device__ __forceinline__ void k_ends(const T (&c)[E], unsigned lane, T &nxt, T &prv)
{
    nxt = __shfl(c[0], (int)((lane + 1) & 63));
    prv = __shfl(c[E - 1], (int)((lane + 63) & 63));
}

template <int E>
__global__ __launch_bounds__(64 * kPostMarchWaves) void k_post_tensor_march(PostTensorArgs T, unsigned N,
                                                                             unsigned logN, unsigned TI, PostStep P)
{
    constexpr unsigned W = kPostMarchWaves, S = 64u * E;
    __shared__ PostTile<E> L;
    const unsigned m = N - 1, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned jblocks = N / W, xcd = blockIdx.x % kXcd, lb = blockIdx.x / kXcd;
    const unsigned j0 = (lb % jblocks) * W, i0 = xcd * (N / kXcd) + (lb / jblocks) * TI;
    const unsigned rj = (j0 + w) << logN, P2 = 2 * logN;
    // waves 0 and 7 also carry the halo rows j0 - 1 and j0 + 8 (LDS rows 0 and 9)
    const bool halo = w == 0 || w == W - 1;
    const unsigned rh = ((w == 0 ? j0 - 1 : j0 + W) & m) << logN;
    const unsigned sh = w == 0 ? 0u : W + 1;
    const double *__restrict__ f = T.f;
    const double *__restrict__ fd = T.fd;
    auto pl = [&](unsigned i) { return (i & m) << P2; };

    double fdn[E], fcur[E];   // Phi rows i-1, i
    double ddn[E], dcur[E];   // Phi' rows i-1, i
    float hdn[6][E], hcur[6][E];
    {
        const unsigned pm = pl(i0 - 1), p0 = pl(i0);
        ld_row<double, E>(f + pm + rj, lane, fdn);
        ld_row<double, E>(f + p0 + rj, lane, fcur);
        ld_row<double, E>(fd + pm + rj, lane, ddn);
        ld_row<double, E>(fd + p0 + rj, lane, dcur);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            ld_row<float, E>(T.h[c] + pm + rj, lane, hdn[c]);
            ld_row<float, E>(T.h[c] + p0 + rj, lane, hcur[c]);
        }
        // Phi planes i0 - 1 and i0 into their ring slots (0 and 1)
        lds_row_put<double, E>(L.f[0][w + 1], lane, fdn);
        lds_row_put<double, E>(L.f[1][w + 1], lane, fcur);
        if (halo) {
            double t0[E], t1[E];
            ld_row<double, E>(f + pm + rh, lane, t0);
            ld_row<double, E>(f + p0 + rh, lane, t1);
            lds_row_put<double, E>(L.f[0][sh], lane, t0);
            lds_row_put<double, E>(L.f[1][sh], lane, t1);
        }
    }
    for (unsigned t = 0; t < TI; ++t) {
        const unsigned p = i0 + t, pi = pl(p), pu = pl(p + 1);
        const unsigned sm = t % 3, s0 = (t + 1) % 3, sp = (t + 2) % 3;
        // this plane's loads: the wave's Phi and Phi' rows of plane i+1; the
        // halo waves' rows of Phi (plane i+1), Phi' and h (plane i) for the LDS
        double fup[E], dup[E];
        ld_row<double, E>(f + pu + rj, lane, fup);
        ld_row<double, E>(fd + pu + rj, lane, dup);
        lds_row_put<double, E>(L.fd[w + 1], lane, dcur);
#pragma unroll
        for (int c = 0; c < 6; ++c) lds_row_put<float, E>(L.h[c][w + 1], lane, hcur[c]);
        if (halo) {
            double fh[E], dh[E];
            float hh[6][E];
            ld_row<double, E>(f + pu + rh, lane, fh);
            ld_row<double, E>(fd + pi + rh, lane, dh);
#pragma unroll
            for (int c = 0; c < 6; ++c) ld_row<float, E>(T.h[c] + pi + rh, lane, hh[c]);
            lds_row_put<double, E>(L.f[sp][sh], lane, fh);
            lds_row_put<double, E>(L.fd[sh], lane, dh);
#pragma unroll
            for (int c = 0; c < 6; ++c) lds_row_put<float, E>(L.h[c][sh], lane, hh[c]);
        }
        lds_row_put<double, E>(L.f[sp][w + 1], lane, fup);
        __syncthreads();
        double Sv[E][6], fnew[E];
        double c_n, c_p, u_n, u_p, d_n, d_p, v_n, v_p;
        k_ends<double, E>(fcur, lane, c_n, c_p);
        k_ends<double, E>(fup, lane, u_n, u_p);
        k_ends<double, E>(fdn, lane, d_n, d_p);
        k_ends<double, E>(dcur, lane, v_n, v_p);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const unsigned x = lane * E + e, xp = (x + 1) & (S - 1), xm = (x - 1) & (S - 1);
            SiteStencil q;
            q.c = fcur[e];
            q.fxp = fup[e];
            q.fxm = fdn[e];
            q.fyp = L.f[s0][w + 2][x];
            q.fym = L.f[s0][w][x];
            q.fzp = kplus<double, E>(fcur, e, c_n);
            q.fzm = kminus<double, E>(fcur, e, c_p);
            q.fxpyp = L.f[sp][w + 2][x];
            q.fxpym = L.f[sp][w][x];
            q.fxmyp = L.f[sm][w + 2][x];
            q.fxmym = L.f[sm][w][x];
            q.fxpzp = kplus<double, E>(fup, e, u_n);
            q.fxpzm = kminus<double, E>(fup, e, u_p);
            q.fxmzp = kplus<double, E>(fdn, e, d_n);
            q.fxmzm = kminus<double, E>(fdn, e, d_p);
            q.fypzp = L.f[s0][w + 2][xp];
            q.fymzm = L.f[s0][w][xm];
            q.dc = dcur[e];
            q.dxp = dup[e];
            q.dxm = ddn[e];
            q.dyp = L.fd[w + 2][x];
            q.dym = L.fd[w][x];
            q.dzp = kplus<double, E>(dcur, e, v_n);
            q.dzm = kminus<double, E>(dcur, e, v_p);
            post_source(q, P, Sv[e]);
            fnew[e] = q.c + P.dd * q.dc;
        }
        st_row<double, E>(T.f_out + pi + rj, lane, fnew);
        float hup[6][E];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            float hv[E], hdnew[E], hn[E];
            ld_row<float, E>(T.h[c] + pu + rj, lane, hup[c]);
            ld_row<float, E>(T.hd[c] + pi + rj, lane, hv);
            float h_n, h_p;
            k_ends<float, E>(hcur[c], lane, h_n, h_p);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const unsigned x = lane * E + e;
                const float hl = kplus<float, E>(hcur[c], e, h_n) + kminus<float, E>(hcur[c], e, h_p) +
                                 L.h[c][w + 2][x] + L.h[c][w][x] + hup[c][e] + hdn[c][e] - 6.0f * hcur[c][e];
                post_tensor_update(hl, hcur[c][e], hv[e], Sv[e][c], P, &hdnew[e], &hn[e]);
            }
            st_row<float, E>(T.hd[c] + pi + rj, lane, hdnew);
            st_row<float, E>(T.h_out[c] + pi + rj, lane, hn);
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            fdn[e] = fcur[e];
            fcur[e] = fup[e];
            ddn[e] = dcur[e];
            dcur[e] = dup[e];
        }
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                hdn[c][e] = hcur[c][e];
                hcur[c][e] = hup[c][e];
            }
    }
}

// --------------------------------------------------------- drift and set-up

// evolve_fields on the post-inflation state (src/evolution.cpp:260-284)
__global__ __launch_bounds__(kBlock) void k_post_drift(double *__restrict__ f, const double *__restrict__ fd,
                                                        PostTensorArgs T, size_t gs, double d)
{
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        f[id] += d * fd[id];
#pragma unroll
        for (int c = 0; c < 6; ++c) T.h_out[c][id] += (float)(d * (double)T.hd[c][id]);
    }
}

// The reference's set-up sums walk the lattice in order with one FP64
// accumulator each (src/initialize.cpp:261-263, 270-276); the order of the
// roundings is part of the result, and one workgroup keeps it exactly:
//   - while acc lies in one binade [2^e, 2^(e+1)) (either sign) it is A u with
//     u = 2^(e-52) and A an integer, 2^52 <= |A| < 2^53; a step adds v, and with
//     x = v / u (exact: u is a power of two) the rounded sum is (A + rint(x)) u
//     as long as x is not a half-integer (a tie, whose rounding depends on the
//     parity of A) and the exact sum stays in the binade, which holds when
//     2^52 < |A + rint(x)| < 2^53 - 1;
//   - a tie (x a half-integer) rounds to the even one of A + floor(x) and
//     A + floor(x) + 1, a map of A through its parity; runs of steps compose
//     (GridRun), so a round of 4096 values (8 waves x 512, 8 consecutive per
//     lane, staged in LDS) is a prefix scan under that composition: every wave
//     scans its block, the waves' totals give each block its start, and every
//     partial sum is checked (a centred deltaN map, whose values lie on the dN
//     grid, has a tie in every 30 steps or so);
//   - at the round's first step that is not of this kind (out of the binade,
//     |x| >= 2^51, NaN; or acc 0, subnormal or not finite) wave 0 takes
//     FP64 steps from the LDS copy (every lane reads the same value: a
//     broadcast), in runs of 64 to the end of that step's block, and the round
//     is scanned again from there in the new binade;
//   - each wave loads its block of the next round while one is summed.
// The result is the serial walk's bit for bit.  Where the sum seldom leaves its
// binade a round costs one scan instead of 4096 dependent adds; a map centred
// on zero changes binade often near the sum's zero crossings and takes about
// one step in eight serially (256^3).
constexpr int kOrdWaves = 8, kOrdPer = 8, kOrdBlock = 64 * kOrdPer, kOrdRound = kOrdWaves * kOrdBlock;
constexpr int kOrdRun = 64;

struct OrdLds {
    double data[kOrdRound];
    GridRun tot[kOrdWaves];
    long long kb[kOrdWaves];
    unsigned first[kOrdWaves];
    double acc;
    unsigned pos;
};

__device__ __forceinline__ double pow2(int k)
{
    return __longlong_as_double((long long)(1023 + k) << 52);
}

// FP64 steps over data[p ..] in runs of 64 until p >= target (target <= count);
// a run that would pass count (the round's last values) goes one step at a time.
__device__ __forceinline__ double ordered_steps(double acc, const double *data, unsigned &p, unsigned target,
                                                unsigned count)
{
    while (p < target) {
        if (p + kOrdRun <= count) {
#pragma unroll 16
            for (int k = 0; k < kOrdRun; ++k) acc = acc + data[p + k];
            p += kOrdRun;
        } else {
            for (; p < count; ++p) acc = acc + data[p];
        }
    }
    return acc;
}

__device__ __forceinline__ double ordered_sum(const double *__restrict__ g, size_t gs, OrdLds &S)
{
    constexpr double kMagic = 0x1.8p52;   // t = x + kMagic rounds x to an integer (|x| < 2^51)
    constexpr long long kLo = 1ll << 52, kHi = (1ll << 53) - 1;
    constexpr unsigned kNone = ~0u;
    const unsigned tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const unsigned mine = w * kOrdBlock + lane * kOrdPer;   // this lane's first value in a round
    double nxt[kOrdPer];
    auto fetch = [&](size_t r0) {
#pragma unroll
        for (int q = 0; q < kOrdPer; ++q) {
            const size_t id = r0 + mine + q;
            nxt[q] = id < gs ? g[id] : 0.0;
        }
    };
    fetch(0);
    double acc = 0.0;   // the same in every thread
    for (size_t r0 = 0; r0 < gs; r0 += kOrdRound) {
        const unsigned count = (unsigned)(gs - r0 < (size_t)kOrdRound ? gs - r0 : (size_t)kOrdRound);
        double v[kOrdPer];
#pragma unroll
        for (int q = 0; q < kOrdPer; ++q) {
            v[q] = nxt[q];
            S.data[mine + q] = v[q];
        }
        fetch(r0 + kOrdRound);
        __syncthreads();
        unsigned pos = 0;
        while (pos < count) {
            const int ef = (int)((__double_as_longlong(acc) >> 52) & 0x7ff);
            unsigned from = pos;
            if (ef >= 63 && ef < 2047) {   // normal, |acc| >= 2^-960, finite
                const int e = ef - 1023;
                const double iu = pow2(52 - e), u = pow2(e - 52);
                const long long A = (long long)(acc * iu);
                GridRun st[kOrdPer];
                bool bad[kOrdPer], lt = false;
#pragma unroll
                for (int q = 0; q < kOrdPer; ++q) {
                    const unsigned id = mine + q;
                    const bool in = id >= pos && id < count;
                    const double x = v[q] * iu;
                    const bool ok = fabs(x) < 0x1p51;
                    const double t = x + kMagic;
                    const double r = t - kMagic;
                    const bool tie = in && ok && fabs(r - x) == 0.5;
                    // n = rint(x); at a tie the lower neighbour is n - (n > x)
                    const long long n = __double_as_longlong(t) - __double_as_longlong(kMagic);
                    st[q] = !(in && ok) ? GridRun{0, 0} : tie ? grid_tie(n - (r > x ? 1 : 0)) : grid_step(n);
                    bad[q] = in && !ok;
                    lt = lt | tie;
                }
                // a wave without ties takes plain integer sums; with ties, the
                // runs' composition
                const bool wt = __any(lt);
                GridRun seg{0, 0}, incl, excl{0, 0};
                if (wt) {
#pragma unroll
                    for (int q = 0; q < kOrdPer; ++q) seg = grid_compose(seg, st[q]);
                    incl = seg;
#pragma unroll
                    for (int d = 1; d < 64; d <<= 1) {
                        GridRun o;
                        o.A = __shfl_up(incl.A, d);
                        o.f = __shfl_up(incl.f, d);
                        if ((int)lane >= d) incl = grid_compose(o, incl);
                    }
                    excl.A = __shfl_up(incl.A, 1);
                    excl.f = __shfl_up(incl.f, 1);
                } else {
#pragma unroll
                    for (int q = 0; q < kOrdPer; ++q) seg.A += st[q].A;
                    incl = seg;
#pragma unroll
                    for (int d = 1; d < 64; d <<= 1) {
                        const long long o = __shfl_up(incl.A, d);
                        if ((int)lane >= d) incl.A += o;
                    }
                    excl.A = incl.A - seg.A;
                }
                if (lane == 0) excl = GridRun{0, 0};
                if (lane == 63) S.tot[w] = incl;
                __syncthreads();
                // the waves' totals: plain sums of the integers when no wave holds a
                // tie, else composed across lanes 0 .. 7 (one wave's total each)
                int tied = 0;
#pragma unroll
                for (int q = 0; q < kOrdWaves; ++q) tied |= S.tot[q].f;
                GridRun before{0, 0}, total{0, 0};
                if (tied & 1) {
                    GridRun tq = lane < (unsigned)kOrdWaves ? S.tot[lane] : GridRun{0, 0};
#pragma unroll
                    for (int d = 1; d < kOrdWaves; d <<= 1) {
                        GridRun o;
                        o.A = __shfl_up(tq.A, d);
                        o.f = __shfl_up(tq.f, d);
                        if ((int)lane >= d) tq = grid_compose(o, tq);
                    }
                    total = GridRun{__shfl(tq.A, kOrdWaves - 1), __shfl(tq.f, kOrdWaves - 1)};
                    before = GridRun{__shfl(tq.A, (int)w - 1), __shfl(tq.f, (int)w - 1)};
                    if (w == 0) before = GridRun{0, 0};
                } else {
#pragma unroll
                    for (int q = 0; q < kOrdWaves; ++q) {
                        const long long t = S.tot[q].A;
                        before.A += q < (int)w ? t : 0;
                        total.A += t;
                    }
                }
                long long K = grid_apply(grid_compose(before, excl), A), Kb = 0;
                unsigned first = kNone;
#pragma unroll
                for (int q = 0; q < kOrdPer; ++q) {
                    const unsigned id = mine + q;
                    if (id >= pos && id < count) {
                        const long long Kq = wt ? grid_apply(st[q], K) : K + st[q].A, aK = Kq < 0 ? -Kq : Kq;
                        if ((bad[q] || aK <= kLo || aK >= kHi) && first == kNone) {
                            first = id;
                            Kb = K;
                        }
                        K = Kq;
                    }
                }
                const unsigned long long m = __ballot(first != kNone);
                const int l = m ? __builtin_ctzll(m) : 0;
                const unsigned fw = (unsigned)__shfl((int)first, l);
                const long long kw = __shfl(Kb, l);
                if (lane == 0) {
                    S.first[w] = m ? fw : kNone;
                    S.kb[w] = kw;
                }
                __syncthreads();
                int ws = -1;
#pragma unroll
                for (int q = kOrdWaves - 1; q >= 0; --q)
                    if (S.first[q] != kNone) ws = q;
                if (ws < 0) {
                    acc = (double)grid_apply(total, A) * u;
                    break;
                }
                from = S.first[ws];
                acc = (double)S.kb[ws] * u;
            }
            // FP64 steps by wave 0 to the end of the block holding `from`
            if (w == 0) {
                unsigned p = from;
                const unsigned target = (from / kOrdBlock + 1) * kOrdBlock;
                acc = ordered_steps(acc, S.data, p, target < count ? target : count, count);
                if (lane == 0) {
                    S.acc = acc;
                    S.pos = p;
                }
            }
            __syncthreads();
            acc = S.acc;
            pos = S.pos;
            __syncthreads();
        }
    }
    return acc;
}

template <bool TWO>
__global__ __launch_bounds__(64 * kOrdWaves) void k_serial_sums(const double *__restrict__ g1,
                                                                 const double *__restrict__ g2, size_t gs,
                                                                 double *out)
{
    __shared__ OrdLds S;
    const double s1 = ordered_sum(g1, gs, S);
    __syncthreads();
    const double s2 = TWO ? ordered_sum(g2, gs, S) : 0.0;
    if (threadIdx.x == 0) {
        out[0] = s1;
        out[1] = s2;
    }
}

// Phi = P (src - mean) [/ den] (src/initialize.cpp:264-265, 278-281)
__global__ __launch_bounds__(kBlock) void k_post_phi(double *__restrict__ f, const double *__restrict__ src, size_t gs,
                                                      double P, double mean, double den, int divide)
{
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        const double v = P * (src[id] - mean);
        f[id] = divide ? v / den : v;
    }
}

// sums of Phi, Phi' and Phi^2 for meansvars (fixed order: block tree, then
// the blocks in order by one workgroup)
__global__ __launch_bounds__(kBlock) void k_post_sums3(const double *__restrict__ f, const double *__restrict__ fd,
                                                        size_t gs, double *__restrict__ partials)
{
    __shared__ double smem[3 * kWaves];
    double acc[3] = {0.0, 0.0, 0.0};
    for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += (size_t)gridDim.x * blockDim.x) {
        const double v = f[id];
        acc[0] += v;
        acc[1] += fd[id];
        acc[2] += v * v;
    }
    block_sum<3>(acc, smem);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) partials[(size_t)blockIdx.x * 3 + q] = acc[q];
}

__global__ __launch_bounds__(kBlock) void k_post_sums3_final(const double *__restrict__ partials, int nb,
                                                              double *__restrict__ out)
{
    __shared__ double smem[3 * kWaves];
    double tot[3];
    sum_partials<3>(partials, nb, tot, smem);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) out[q] = tot[q];
}

unsigned site_grid(size_t gs)
{
    size_t b = (gs + kBlock - 1) / kBlock;
    if (b > 4096) b = 4096;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

// sites per lane of the row kernels: 2 from N = 128 (measured at 256^3: the
// tensor kernel 0.82 ms per step with 2, 1.06 ms with 1)
int post_row_elems(int N) { return N < 64 ? 0 : (N < 128 ? 1 : 2); }

// Rows per walk: the longest walk (up to 32 rows) that still gives at least 256
// workgroups, so that short walks' R + 2 row reads matter only at small N.
unsigned post_row_walk(int N, int E, int PW, unsigned Rmax)
{
    const size_t segs = (size_t)N / (64 * (size_t)E);
    unsigned R = Rmax;
    while (R > 2 && (size_t)N * N * segs / ((size_t)PW * R) < 256) R >>= 1;
    return R;
}

unsigned post_row_blocks(int N, int E, int PW, unsigned Rmax)
{
    const size_t segs = (size_t)N / (64 * (size_t)E);
    return (unsigned)((size_t)N * N * segs / ((size_t)PW * post_row_walk(N, E, PW, Rmax)));
}

// The tensor march takes lattices whose rows are one segment of E = N / 64 sites
// per lane (N = 64, 128, 256).  Planes per tile: the XCD's whole slab, halved
// until the grid has 256 workgroups (one per CU at N = 256).
int post_march_elems(int N) { return (N == 64 || N == 128 || N == 256) ? N / 64 : 0; }

unsigned post_march_blocks(int N, unsigned TI)
{
    return (unsigned)(kXcd * (N / kPostMarchWaves) * ((N / kXcd) / TI));
}

unsigned post_march_planes(int N)
{
    unsigned TI = (unsigned)(N / kXcd);
    while (TI > 2 && post_march_blocks(N, TI) < 256) TI >>= 1;
    return TI;
}

hipError_t post_step(const PostState &S, int N, int logN, const PostStep &P, hipStream_t s, bool rows)
{
    PostTensorArgs T;
    T.f = S.f;
    T.fd = S.fd_out;
    T.f_out = S.f_out;
    for (int c = 0; c < 6; ++c) {
        T.h[c] = S.h[c];
        T.h_out[c] = S.h_out[c];
        T.hd[c] = S.hd[c];
    }
    const int E = post_row_elems(N);
    if (E == 0) {
        const unsigned G = site_grid((size_t)N * N * N);
        k_post_kick_site<<<G, kBlock, 0, s>>>(S.f, S.fd, S.fd_out, (unsigned)N, (unsigned)logN, P);
        k_post_tensor_site<<<G, kBlock, 0, s>>>(T, (unsigned)N, (unsigned)logN, P);
    } else {
        constexpr int KP = kPostKickPlanes, TP = kPostTensorPlanes;
        constexpr unsigned KR = kPostKickRows, TR = kPostTensorRows;
        k_post_kick_rows<1><<<post_row_blocks(N, 1, KP, KR), 64 * KP, 0, s>>>(
            S.f, S.fd, S.fd_out, (unsigned)N, (unsigned)logN, post_row_walk(N, 1, KP, KR), P);
        const int EM = rows ? 0 : post_march_elems(N);
        if (EM) {
            const unsigned TI = post_march_planes(N), G = post_march_blocks(N, TI), nt = 64 * kPostMarchWaves;
            if (EM == 1) k_post_tensor_march<1><<<G, nt, 0, s>>>(T, (unsigned)N, (unsigned)logN, TI, P);
            else if (EM == 2) k_post_tensor_march<2><<<G, nt, 0, s>>>(T, (unsigned)N, (unsigned)logN, TI, P);
            else k_post_tensor_march<4><<<G, nt, 0, s>>>(T, (unsigned)N, (unsigned)logN, TI, P);
        } else if (E == 2)
            k_post_tensor_rows<2><<<post_row_blocks(N, 2, TP, TR), 64 * TP, 0, s>>>(
                T, (unsigned)N, (unsigned)logN, post_row_walk(N, 2, TP, TR), P);
        else
            k_post_tensor_rows<1><<<post_row_blocks(N, 1, TP, TR), 64 * TP, 0, s>>>(
                T, (unsigned)N, (unsigned)logN, post_row_walk(N, 1, TP, TR), P);
    }
    return hipGetLastError();
}

hipError_t post_drift(const PostState &S, size_t gs, double d, hipStream_t s)
{
    PostTensorArgs T{};
    for (int c = 0; c < 6; ++c) {
        T.h_out[c] = S.h[c];
        T.hd[c] = S.hd[c];
    }
    k_post_drift<<<site_grid(gs), kBlock, 0, s>>>(S.f, S.fd, T, gs, d);
    return hipGetLastError();
}

hipError_t serial_sums(const double *g1, const double *g2, size_t gs, double *out, hipStream_t s)
{
    if (g2) k_serial_sums<true><<<1, 64 * kOrdWaves, 0, s>>>(g1, g2, gs, out);
    else k_serial_sums<false><<<1, 64 * kOrdWaves, 0, s>>>(g1, nullptr, gs, out);
    return hipGetLastError();
}

hipError_t post_phi(double *f, const double *src, size_t gs, double P, double mean, double den, int divide,
                    hipStream_t s)
{
    k_post_phi<<<site_grid(gs), kBlock, 0, s>>>(f, src, gs, P, mean, den, divide);
    return hipGetLastError();
}

hipError_t post_sums3(const double *f, const double *fd, size_t gs, double *work, double *out, hipStream_t s)
{
    const unsigned nb = site_grid(gs) > 1024 ? 1024 : site_grid(gs);
    k_post_sums3<<<nb, kBlock, 0, s>>>(f, fd, gs, work);
    k_post_sums3_final<<<1, kBlock, 0, s>>>(work, (int)nb, out);
    return hipGetLastError();
}

}  // namespace ie
