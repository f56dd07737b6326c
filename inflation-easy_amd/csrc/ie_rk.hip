// ie_rk.hip -- RK4 and RK45 (Dormand-Prince) inflation steps on gfx950.
//
// The reference's secondary integrators (src/evolution.cpp:286-904): stage states
// y + h sum_q A[s][q] k_q, a right-hand side with fused energy sums, the weighted
// RK4 update, and for RK45 the fifth-order state with the embedded error in a
// weighted RMS norm over every variable and an accept/reject decision.  FP64
// storage only.  Per stage:
//   k_rk_combine  y_s = y + h sum A k        (pointwise; fields and tensors)
//   k_rk_rhs      k_s = f(y_s) and the sums of -f lapl f and V(f); its last
//                 workgroup writes (a', a'') of the stage
// then k_rk_update (RK4) or k_rk45_final (RK45: y5, error sums; the last
// workgroup forms the norm, the step factor and the decision).
#include <math.h>

#include <vector>

#include "ie_internal.h"

// no contraction: the stage sums, updates and right-hand sides keep the
// reference's rounding (the post-inflation right-hand side is bitwise equal)
#pragma clang fp contract(off)

namespace ie {

namespace {

constexpr int kMaxStages = 7;

struct RkState {
    // fields of y (f, fd) and of the stage state (ftmp, fdtmp), FP64
    const double *f, *fd;
    double *ftmp, *fdtmp;
    const float *h[6], *hd[6];
    float *htmp[6], *hdtmp[6];
    double *kf[kMaxStages], *kfd[kMaxStages];
    float *kh[kMaxStages][6], *khd[kMaxStages][6];
    int N, logN, sigw;
};

struct RkCoef {
    double A[kMaxStages][kMaxStages];
    double B[kMaxStages];
    double E[kMaxStages];
};

// Scalars of an RK step on the device: the stage derivatives of the background.
struct RkScalars {
    double ka[kMaxStages], kad[kMaxStages];
    double err_norm, a5, ad5, factor;
    unsigned counter;
    int accepted;
};

__device__ __forceinline__ unsigned gsize(const RkState &S) { return (unsigned)S.N * S.N * S.N; }

// stage state y_s = y + h sum_{q<s} A[s][q] k_q (src/evolution.cpp:639-669)
__global__ __launch_bounds__(kBlock) void k_rk_combine(RkState S, const RkCoef *C, double h, int stage)
{
    const unsigned gs = gsize(S);
    for (unsigned id = blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += gridDim.x * blockDim.x) {
        double sf = 0.0, sfd = 0.0;
        for (int q = 0; q < stage; ++q) {
            const double c = C->A[stage][q];
            if (c == 0.0) continue;
            sf += c * S.kf[q][id];
            sfd += c * S.kfd[q][id];
        }
        S.ftmp[id] = S.f[id] + h * sf;
        S.fdtmp[id] = S.fd[id] + h * sfd;
        if (S.sigw)
            for (int c6 = 0; c6 < 6; ++c6) {
                double sh = 0.0, shd = 0.0;
                for (int q = 0; q < stage; ++q) {
                    const double c = C->A[stage][q];
                    if (c == 0.0) continue;
                    sh += c * (double)S.kh[q][c6][id];
                    shd += c * (double)S.khd[q][c6][id];
                }
                S.htmp[c6][id] = (float)((double)S.h[c6][id] + h * sh);
                S.hdtmp[c6][id] = (float)((double)S.hd[c6][id] + h * shd);
            }
    }
}

__device__ __forceinline__ void nb(unsigned id, unsigned N, unsigned logN, unsigned n[6])
{
    const unsigned m = N - 1;
    const unsigned k = id & m, j = (id >> logN) & m, i = id >> (2 * logN);
    const unsigned bk = id - k, bj = id - (j << logN), bi = id - (i << (2 * logN));
    n[0] = bk + ((k + 1) & m);
    n[1] = bk + ((k - 1) & m);
    n[2] = bj + (((j + 1) & m) << logN);
    n[3] = bj + (((j - 1) & m) << logN);
    n[4] = bi + (((i + 1) & m) << (2 * logN));
    n[5] = bi + (((i - 1) & m) << (2 * logN));
}

// right-hand side at the stage state (compute_inflation_rhs, src/evolution.cpp:419-518)
__global__ __launch_bounds__(kBlock) void k_rk_rhs(RkState S, Pot P, const RkCoef *C, RkScalars *R, DevBg *bg,
                                                   double h, int stage, int use_y, double *partials)
{
    __shared__ double smem[2 * kWaves];
    // background of the stage (src/evolution.cpp:672-688)
    double a_s = bg->a, ad_s = bg->ad;
    for (int q = 0; q < stage; ++q) {
        const double c = C->A[stage][q];
        if (c == 0.0) continue;
        a_s += h * c * R->ka[q];
        ad_s += h * c * R->kad[q];
    }
    const double s = bg->rescale_s, dx = bg->dx;
    const double laplnorm = pow(a_s, -2.0 * s) / (dx * dx);
    const double friction = (2.0 + s) * ad_s / a_s;
    const double potnorm = pow(a_s, 2.0 - 2.0 * s);
    const double srcAmp = 2.0 * pow(a_s, -2.0 * s);
    const double *fs = use_y ? S.f : S.ftmp;
    const double *fds = use_y ? S.fd : S.fdtmp;
    const unsigned gs = gsize(S), N = (unsigned)S.N, logN = (unsigned)S.logN;
    double acc[2] = {0.0, 0.0};
    int err = 0;
    for (unsigned id = blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += gridDim.x * blockDim.x) {
        unsigned n[6];
        nb(id, N, logN, n);
        const double fh = fs[id];
        const double lf = fs[n[0]] + fs[n[1]] + fs[n[2]] + fs[n[3]] + fs[n[4]] + fs[n[5]] - 6.0 * fh;
        double V, dV;
        pot_eval(P, fh, &V, &dV, &err);
        S.kf[stage][id] = fds[id];
        S.kfd[stage][id] = laplnorm * lf - friction * fds[id] - potnorm * dV;
        acc[0] -= fh * lf;
        acc[1] += V;
        if (S.sigw) {
            const double hx = 0.5 / dx;
            const double gx = (fs[n[4]] - fs[n[5]]) * hx, gy = (fs[n[2]] - fs[n[3]]) * hx,
                         gz = (fs[n[0]] - fs[n[1]]) * hx;
            const double T[6] = {gx * gx, gy * gy, gz * gz, gx * gy, gx * gz, gy * gz};
            for (int c6 = 0; c6 < 6; ++c6) {
                const float *hh = use_y ? S.h[c6] : S.htmp[c6];
                const float *hdh = use_y ? S.hd[c6] : S.hdtmp[c6];
                const float lh = hh[n[0]] + hh[n[1]] + hh[n[2]] + hh[n[3]] + hh[n[4]] + hh[n[5]] - 6.0f * hh[id];
                S.kh[stage][c6][id] = hdh[id];
                S.khd[stage][c6][id] =
                    (float)(laplnorm * (double)lh - friction * (double)hdh[id] + srcAmp * T[c6]);
            }
        }
    }
    if (__any(err) && (threadIdx.x & 63) == 0) atomicOr(&bg->error, 1);
    block_sum<2>(acc, smem);
    if (threadIdx.x == 0) {
        partials[blockIdx.x * 2] = acc[0];
        partials[blockIdx.x * 2 + 1] = acc[1];
    }
    if (!last_block(&R->counter, gridDim.x)) return;
    double tot[2];
    sum_partials<2>(partials, gridDim.x, tot, smem);
    if (threadIdx.x != 0) return;
    const double gsd = (double)gs;
    const double grad = 0.5 * tot[0] * ((1.0 / (a_s * dx)) * (1.0 / (a_s * dx))) / gsd;
    const double pavg = tot[1] / gsd;
    const double source = 2.0 * grad / 3.0 + pavg;
    R->ka[stage] = ad_s;
    R->kad[stage] = pow(a_s, 3.0 - 2.0 * s) * source - (s + 1.0) * (ad_s * ad_s) / a_s;
}

// post-inflation right-hand side at the stage state (compute_inflation_rhs's
// post branch, src/evolution.cpp:520-592): no sums; a'' from (a, a') alone
__global__ __launch_bounds__(kBlock) void k_rk_rhs_post(RkState S, const RkCoef *C, RkScalars *R, const DevBg *bg,
                                                        double h, int stage, int use_y, double omega)
{
    double a_s = bg->a, ad_s = bg->ad;
    for (int q = 0; q < stage; ++q) {
        const double c = C->A[stage][q];
        if (c == 0.0) continue;
        a_s += h * c * R->ka[q];
        ad_s += h * c * R->kad[q];
    }
    const double s = bg->rescale_s, dx = bg->dx, w = omega;
    PostStep P{};
    P.laplnorm = pow(a_s, -2.0 * s) / (dx * dx);
    const double scalar_friction = (3.0 * (1.0 + w) + s) * ad_s / a_s;
    const double tensor_friction = (2.0 + s) * ad_s / a_s;
    const double Htilde = ad_s / a_s;
    P.invH = 1.0 / Htilde;
    P.coeffU = 4.0 / (3.0 * (1.0 + w));
    P.srcamp = 2.0 * pow(a_s, -2.0 * s);
    P.hx = 0.5 / dx;
    P.inv_dx2 = 1.0 / (dx * dx);
    P.cross_den = 4.0 * dx * dx;
    const double *fs = use_y ? S.f : S.ftmp;
    const double *fds = use_y ? S.fd : S.fdtmp;
    const unsigned gs = gsize(S), N = (unsigned)S.N, logN = (unsigned)S.logN, m = N - 1;
    for (unsigned id = blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += gridDim.x * blockDim.x) {
        unsigned n[6];
        nb(id, N, logN, n);
        const unsigned k = id & m, j = (id >> logN) & m, i = id >> (2 * logN);
        auto at = [&](unsigned ii, unsigned jj, unsigned kk) {
            return ((ii & m) << (2 * logN)) | ((jj & m) << logN) | (kk & m);
        };
        SiteStencil q;
        q.c = fs[id];
        q.fzp = fs[n[0]];
        q.fzm = fs[n[1]];
        q.fyp = fs[n[2]];
        q.fym = fs[n[3]];
        q.fxp = fs[n[4]];
        q.fxm = fs[n[5]];
        q.fxpyp = fs[at(i + 1, j + 1, k)];
        q.fxpym = fs[at(i + 1, j - 1, k)];
        q.fxmyp = fs[at(i - 1, j + 1, k)];
        q.fxmym = fs[at(i - 1, j - 1, k)];
        q.fxpzp = fs[at(i + 1, j, k + 1)];
        q.fxpzm = fs[at(i + 1, j, k - 1)];
        q.fxmzp = fs[at(i - 1, j, k + 1)];
        q.fxmzm = fs[at(i - 1, j, k - 1)];
        q.fypzp = fs[at(i, j + 1, k + 1)];
        q.fymzm = fs[at(i, j - 1, k - 1)];
        q.dc = fds[id];
        q.dzp = fds[n[0]];
        q.dzm = fds[n[1]];
        q.dyp = fds[n[2]];
        q.dym = fds[n[3]];
        q.dxp = fds[n[4]];
        q.dxm = fds[n[5]];
        const double lf = q.fzp + q.fzm + q.fyp + q.fym + q.fxp + q.fxm - 6.0 * q.c;
        S.kf[stage][id] = q.dc;
        S.kfd[stage][id] = w * P.laplnorm * lf - scalar_friction * q.dc;
        double Ssrc[6];
        post_source(q, P, Ssrc);
        for (int c6 = 0; c6 < 6; ++c6) {
            const float *hh = use_y ? S.h[c6] : S.htmp[c6];
            const float *hdh = use_y ? S.hd[c6] : S.hdtmp[c6];
            const float lh = hh[n[0]] + hh[n[1]] + hh[n[2]] + hh[n[3]] + hh[n[4]] + hh[n[5]] - 6.0f * hh[id];
            S.kh[stage][c6][id] = hdh[id];
            S.khd[stage][c6][id] =
                (float)(P.laplnorm * (double)lh - tensor_friction * (double)hdh[id] + P.srcamp * Ssrc[c6]);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        R->ka[stage] = ad_s;
        R->kad[stage] = -(s - 0.5 * (1.0 - 3.0 * w)) * (ad_s * ad_s) / a_s;
    }
}

// RK4 weighted update in place (src/evolution.cpp:715-758)
struct TensorOut {
    float *h[6], *hd[6];
};

__global__ __launch_bounds__(kBlock) void k_rk_update(RkState S, double *f, double *fd, TensorOut T, const RkCoef *C,
                                                      double h, int stages)
{
    const unsigned gs = gsize(S);
    for (unsigned id = blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += gridDim.x * blockDim.x) {
        double sf = 0.0, sfd = 0.0;
        for (int q = 0; q < stages; ++q) {
            const double b = C->B[q];
            if (b == 0.0) continue;
            sf += b * S.kf[q][id];
            sfd += b * S.kfd[q][id];
        }
        f[id] += h * sf;
        fd[id] += h * sfd;
        if (S.sigw)
            for (int c6 = 0; c6 < 6; ++c6) {
                double sh = 0.0, shd = 0.0;
                for (int q = 0; q < stages; ++q) {
                    const double b = C->B[q];
                    if (b == 0.0) continue;
                    sh += b * (double)S.kh[q][c6][id];
                    shd += b * (double)S.khd[q][c6][id];
                }
                T.h[c6][id] = (float)((double)T.h[c6][id] + h * sh);
                T.hd[c6][id] = (float)((double)T.hd[c6][id] + h * shd);
            }
    }
}

// RK45: fifth-order state into the stage buffers, embedded error in a weighted
// RMS norm (src/evolution.cpp:800-904); the last workgroup decides.
__global__ __launch_bounds__(kBlock) void k_rk45_final(RkState S, const RkCoef *C, RkScalars *R, const DevBg *bg,
                                                       double h, double abs_tol, double rel_tol, double safety,
                                                       double *partials)
{
    __shared__ double smem[kWaves];
    const unsigned gs = gsize(S);
    double acc[1] = {0.0};
    for (unsigned id = blockIdx.x * blockDim.x + threadIdx.x; id < gs; id += gridDim.x * blockDim.x) {
        double y5f = S.f[id], y5fd = S.fd[id], ef = 0.0, efd = 0.0;
        for (int q = 0; q < 7; ++q) {
            const double b = C->B[q], e = C->E[q];
            if (b != 0.0) {
                y5f += h * b * S.kf[q][id];
                y5fd += h * b * S.kfd[q][id];
            }
            if (e != 0.0) {
                ef += h * e * S.kf[q][id];
                efd += h * e * S.kfd[q][id];
            }
        }
        S.ftmp[id] = y5f;
        S.fdtmp[id] = y5fd;
        const double sf = abs_tol + rel_tol * fmax(fabs(S.f[id]), fabs(y5f));
        const double sfd = abs_tol + rel_tol * fmax(fabs(S.fd[id]), fabs(y5fd));
        acc[0] += (ef / sf) * (ef / sf) + (efd / sfd) * (efd / sfd);
        if (S.sigw)
            for (int c6 = 0; c6 < 6; ++c6) {
                double y5h = (double)S.h[c6][id], y5hd = (double)S.hd[c6][id], eh = 0.0, ehd = 0.0;
                for (int q = 0; q < 7; ++q) {
                    const double b = C->B[q], e = C->E[q];
                    if (b != 0.0) {
                        y5h += h * b * (double)S.kh[q][c6][id];
                        y5hd += h * b * (double)S.khd[q][c6][id];
                    }
                    if (e != 0.0) {
                        eh += h * e * (double)S.kh[q][c6][id];
                        ehd += h * e * (double)S.khd[q][c6][id];
                    }
                }
                S.htmp[c6][id] = (float)y5h;
                S.hdtmp[c6][id] = (float)y5hd;
                const double sh = abs_tol + rel_tol * fmax(fabs((double)S.h[c6][id]), fabs(y5h));
                const double shd = abs_tol + rel_tol * fmax(fabs((double)S.hd[c6][id]), fabs(y5hd));
                acc[0] += (eh / sh) * (eh / sh) + (ehd / shd) * (ehd / shd);
            }
    }
    block_sum<1>(acc, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
    if (!last_block(&R->counter, gridDim.x)) return;
    double tot[1];
    sum_partials<1>(partials, gridDim.x, tot, smem);
    if (threadIdx.x != 0) return;
    double a5 = bg->a, ad5 = bg->ad, erra = 0.0, errad = 0.0;
    for (int q = 0; q < 7; ++q) {
        const double b = C->B[q], e = C->E[q];
        if (b != 0.0) {
            a5 += h * b * R->ka[q];
            ad5 += h * b * R->kad[q];
        }
        if (e != 0.0) {
            erra += h * e * R->ka[q];
            errad += h * e * R->kad[q];
        }
    }
    double err = tot[0];
    const double sa = abs_tol + rel_tol * fmax(fabs(bg->a), fabs(a5));
    const double sad = abs_tol + rel_tol * fmax(fabs(bg->ad), fabs(ad5));
    err += (erra / sa) * (erra / sa) + (errad / sad) * (errad / sad);
    const double nvars = (double)(2 * (size_t)gs + 2 + (S.sigw ? 12 * (size_t)gs : 0));
    const double norm = sqrt(err / nvars);
    const double safe = norm > 1e-16 ? norm : 1e-16;
    double factor = safety * pow(safe, -0.2);
    factor = factor < 0.1 ? 0.1 : (factor > 5.0 ? 5.0 : factor);
    R->err_norm = norm;
    R->a5 = a5;
    R->ad5 = ad5;
    R->factor = factor;
    R->accepted = norm <= 1.0;
}

}  // namespace

// host side ---------------------------------------------------------------------

struct RkWork {
    int stages = 0, sigw = 0;
    int post = 0;               // post-inflation right-hand side
    double omega = 0.0;
    size_t gs = 0;
    std::vector<void *> allocs;
    RkState S{};
    RkCoef *coef = nullptr;     // device
    RkScalars *scal = nullptr;  // device
    double *partials = nullptr;
};

static hipError_t rk_alloc(RkWork *w, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) w->allocs.push_back(*p);
    return e;
}

void rk_set_post(RkWork *w, int post, double omega)
{
    w->post = post;
    w->omega = omega;
}

void rk_free(RkWork *w)
{
    if (!w) return;
    for (void *p : w->allocs) (void)hipFree(p);
    delete w;
}

hipError_t rk_create(int N, int logN, int sigw, int stages, RkWork **out)
{
    RkWork *w = new RkWork();
    w->stages = stages;
    w->sigw = sigw;
    w->gs = (size_t)N * N * N;
    w->S.N = N;
    w->S.logN = logN;
    w->S.sigw = sigw;
    hipError_t e = hipSuccess;
    auto A = [&](void **p, size_t b) {
        if (e == hipSuccess) e = rk_alloc(w, p, b);
    };
    A((void **)&w->S.ftmp, w->gs * 8);
    A((void **)&w->S.fdtmp, w->gs * 8);
    for (int s = 0; s < stages; ++s) {
        A((void **)&w->S.kf[s], w->gs * 8);
        A((void **)&w->S.kfd[s], w->gs * 8);
        if (sigw)
            for (int c = 0; c < 6; ++c) {
                A((void **)&w->S.kh[s][c], w->gs * 4);
                A((void **)&w->S.khd[s][c], w->gs * 4);
            }
    }
    if (sigw)
        for (int c = 0; c < 6; ++c) {
            A((void **)&w->S.htmp[c], w->gs * 4);
            A((void **)&w->S.hdtmp[c], w->gs * 4);
        }
    A((void **)&w->coef, sizeof(RkCoef));
    A((void **)&w->scal, sizeof(RkScalars));
    A((void **)&w->partials, 2048 * 2 * 8);
    if (e == hipSuccess) e = hipMemset(w->scal, 0, sizeof(RkScalars));
    if (e != hipSuccess) {
        rk_free(w);
        return e;
    }
    // Butcher tables (src/evolution.cpp:373-414, 595-637)
    RkCoef c{};
    if (stages == 4) {
        c.A[1][0] = 0.5;
        c.A[2][1] = 0.5;
        c.A[3][2] = 1.0;
        c.B[0] = 1.0 / 6.0;
        c.B[1] = 1.0 / 3.0;
        c.B[2] = 1.0 / 3.0;
        c.B[3] = 1.0 / 6.0;
    } else {
        const double a21 = 1.0 / 5.0, a31 = 3.0 / 40.0, a32 = 9.0 / 40.0, a41 = 44.0 / 45.0, a42 = -56.0 / 15.0,
                     a43 = 32.0 / 9.0, a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0,
                     a54 = -212.0 / 729.0, a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0,
                     a64 = 49.0 / 176.0, a65 = -5103.0 / 18656.0;
        const double b1 = 35.0 / 384.0, b3 = 500.0 / 1113.0, b4 = 125.0 / 192.0, b5 = -2187.0 / 6784.0,
                     b6 = 11.0 / 84.0;
        const double bs1 = 5179.0 / 57600.0, bs3 = 7571.0 / 16695.0, bs4 = 393.0 / 640.0,
                     bs5 = -92097.0 / 339200.0, bs6 = 187.0 / 2100.0, bs7 = 1.0 / 40.0;
        const double A[7][7] = {{0, 0, 0, 0, 0, 0, 0},           {a21, 0, 0, 0, 0, 0, 0},
                                {a31, a32, 0, 0, 0, 0, 0},       {a41, a42, a43, 0, 0, 0, 0},
                                {a51, a52, a53, a54, 0, 0, 0},   {a61, a62, a63, a64, a65, 0, 0},
                                {b1, 0, b3, b4, b5, b6, 0}};
        for (int i = 0; i < 7; ++i)
            for (int j = 0; j < 7; ++j) c.A[i][j] = A[i][j];
        const double B[7] = {b1, 0, b3, b4, b5, b6, 0};
        const double E[7] = {b1 - bs1, 0, b3 - bs3, b4 - bs4, b5 - bs5, b6 - bs6, -bs7};
        for (int i = 0; i < 7; ++i) {
            c.B[i] = B[i];
            c.E[i] = E[i];
        }
    }
    e = hipMemcpy(w->coef, &c, sizeof(c), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        rk_free(w);
        return e;
    }
    *out = w;
    return hipSuccess;
}

static int rk_grid(size_t gs)
{
    size_t b = gs / kBlock;
    if (b > 2048) b = 2048;
    return (int)(b < 1 ? 1 : b);
}

// Stages of one RK step from y = (f, fd, h, hd) into w's stage buffers.
static hipError_t rk_stages(RkWork *w, const Pot &P, DevBg *bg, double h, hipStream_t s)
{
    const int G = rk_grid(w->gs);
    for (int st = 0; st < w->stages; ++st) {
        if (st > 0) k_rk_combine<<<G, kBlock, 0, s>>>(w->S, w->coef, h, st);
        if (w->post) k_rk_rhs_post<<<G, kBlock, 0, s>>>(w->S, w->coef, w->scal, bg, h, st, st == 0, w->omega);
        else k_rk_rhs<<<G, kBlock, 0, s>>>(w->S, P, w->coef, w->scal, bg, h, st, st == 0, w->partials);
    }
    return hipGetLastError();
}

static void bind_y(RkWork *w, double *f, double *fd, float *const *h, float *const *hd)
{
    w->S.f = f;
    w->S.fd = fd;
    for (int c = 0; c < 6; ++c) {
        w->S.h[c] = w->sigw ? h[c] : nullptr;
        w->S.hd[c] = w->sigw ? hd[c] : nullptr;
    }
}

// One RK4 step of the state in place; the background (t, a, ad) of bg_host is
// advanced by the host from the stage derivatives (src/evolution.cpp:760-779).
hipError_t rk4_step(RkWork *w, const Pot &P, DevBg *bg, DevBg *bg_host, double *f, double *fd, float *const *h,
                    float *const *hd, double step, hipStream_t s)
{
    bind_y(w, f, fd, h, hd);
    hipError_t e = rk_stages(w, P, bg, step, s);
    if (e != hipSuccess) return e;
    TensorOut T{};
    for (int c = 0; c < 6; ++c) {
        T.h[c] = w->sigw ? h[c] : nullptr;
        T.hd[c] = w->sigw ? hd[c] : nullptr;
    }
    k_rk_update<<<rk_grid(w->gs), kBlock, 0, s>>>(w->S, f, fd, T, w->coef, step, 4);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    RkScalars R;
    if ((e = hipMemcpyAsync(&R, w->scal, sizeof(R), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    const double B[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    for (int q = 0; q < 4; ++q) {
        bg_host->a += step * B[q] * R.ka[q];
        bg_host->ad += step * B[q] * R.kad[q];
    }
    bg_host->t += step;
    return hipSuccess;
}

// One Dormand-Prince trial step (src/evolution.cpp:781-904).  On acceptance the
// fifth-order state is copied over the state and bg_host takes (t + h, a5, ad5);
// *factor is the step-size factor of the error norm, *err_norm the norm.
hipError_t rk45_step(RkWork *w, const Pot &P, DevBg *bg, DevBg *bg_host, double *f, double *fd, float *const *h,
                     float *const *hd, double step, double abs_tol, double rel_tol, double safety, int *accepted,
                     double *factor, double *err_norm, hipStream_t s)
{
    bind_y(w, f, fd, h, hd);
    hipError_t e = rk_stages(w, P, bg, step, s);
    if (e != hipSuccess) return e;
    k_rk45_final<<<rk_grid(w->gs), kBlock, 0, s>>>(w->S, w->coef, w->scal, bg, step, abs_tol, rel_tol, safety,
                                                   w->partials);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    RkScalars R;
    if ((e = hipMemcpyAsync(&R, w->scal, sizeof(R), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    *accepted = R.accepted;
    *factor = R.factor;
    *err_norm = R.err_norm;
    if (!R.accepted) return hipSuccess;
    const size_t b8 = w->gs * 8, b4 = w->gs * 4;
    if ((e = hipMemcpyAsync(f, w->S.ftmp, b8, hipMemcpyDeviceToDevice, s)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(fd, w->S.fdtmp, b8, hipMemcpyDeviceToDevice, s)) != hipSuccess) return e;
    if (w->sigw)
        for (int c = 0; c < 6; ++c) {
            if ((e = hipMemcpyAsync(h[c], w->S.htmp[c], b4, hipMemcpyDeviceToDevice, s)) != hipSuccess) return e;
            if ((e = hipMemcpyAsync(hd[c], w->S.hdtmp[c], b4, hipMemcpyDeviceToDevice, s)) != hipSuccess) return e;
        }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    bg_host->a = R.a5;
    bg_host->ad = R.ad5;
    bg_host->t += step;
    return hipSuccess;
}

}  // namespace ie
