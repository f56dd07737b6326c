/*
 * FFTEASY-layout FFT, restated for one real type.  TEST INFRASTRUCTURE ONLY (see
 * ie_oracle.h).  Included twice by ie_oracle.c with IEO_R = double / float and
 * IEO_SFX = d / f.  Follows src/ffteasy.hpp:21-190 operation by operation:
 *   - forward kernel e^{+2 pi i k n / N}, unnormalised; inverse e^{-i} with 1/N per
 *     transformed dimension;
 *   - the real transform packs z-pairs into complex numbers, transforms the
 *     N x N x N/2 complex array, then unpacks with the kz = N/2 plane written to a
 *     separate N x N complex "Nyquist" array.
 */
#define IEO_CAT2(a, b) a##b
#define IEO_CAT(a, b) IEO_CAT2(a, b)
#define IEO_FN(name) IEO_CAT(name, IEO_SFX)

typedef struct { IEO_R re, im; } IEO_FN(ieo_cpx_);

/* 1-D complex radix-2 DIT transform of n points with stride `skip` complex
 * elements (src/ffteasy.hpp:21-70). */
static void IEO_FN(ieo_fftc1_)(IEO_R *data, int n, int skip, int forward)
{
    IEO_FN(ieo_cpx_) *c = (IEO_FN(ieo_cpx_) *)data;
    const IEO_R two_pi = (IEO_R)4 * IEO_ASIN((IEO_R)1);
    int i1, i2, bit;
    /* bit-reversal permutation, incremental reversed counter */
    for (i1 = 1, i2 = 0; i1 < n; ++i1) {
        for (bit = n / 2; i2 >= bit; bit >>= 1) i2 -= bit;
        i2 += bit;
        if (i2 > i1) {
            IEO_FN(ieo_cpx_) tmp = c[i2 * skip];
            c[i2 * skip] = c[i1 * skip];
            c[i1 * skip] = tmp;
        }
    }
    /* butterflies: span doubles each pass, twiddle advanced by repeated rotation */
    for (int span = 2; span <= n; span <<= 1) {
        const IEO_R ang = (IEO_R)forward * two_pi / (IEO_R)span;
        const IEO_R cr = IEO_COS(ang), sr = IEO_SIN(ang);
        IEO_R wr = (IEO_R)1, wi = (IEO_R)0;
        for (int b = 0; b < span / 2; ++b) {
            for (int blk = 0; blk < n / span; ++blk) {
                const int p = (blk * span + b) * skip;
                const int q = p + (span / 2) * skip;
                const IEO_FN(ieo_cpx_) u = c[p], v = c[q];
                c[p].re = u.re + wr * v.re - wi * v.im;
                c[p].im = u.im + wr * v.im + wi * v.re;
                c[q].re = u.re - wr * v.re + wi * v.im;
                c[q].im = u.im - wr * v.im - wi * v.re;
            }
            const IEO_R owr = wr, owi = wi;
            wr = cr * owr - sr * owi;
            wi = cr * owi + sr * owr;
        }
    }
    if (forward < 0) {
        for (i1 = 0; i1 < skip * n; i1 += skip) {
            c[i1].re /= (IEO_R)n;
            c[i1].im /= (IEO_R)n;
        }
    }
}

/* 3-D complex transform over sizes (n0, n1, n2) complex points, last dimension
 * first (src/ffteasy.hpp:74-88).  The 1-D transforms of one dimension touch
 * disjoint elements, so they run on OpenMP threads with each transform's
 * operations unchanged (the same bits as the serial loop). */
static void IEO_FN(ieo_fftc3_)(IEO_R *data, const int size[3], int forward)
{
    int plane = 1, skip = 1;
    const int total = size[0] * size[1] * size[2];
    for (int dim = 2; dim >= 0; --dim) {
        plane *= size[dim];
        const int nlines = total / plane, n = size[dim], sk = skip, pl = plane;
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < nlines; ++i)
            for (int j = 0; j < sk; ++j)
                IEO_FN(ieo_fftc1_)(data + 2 * ((size_t)i * pl + j), n, sk, forward);
        skip *= size[dim];
    }
}

/* Real 3-D transform of an N^3 array in place with separate Nyquist plane
 * (src/ffteasy.hpp:127-190). */
void IEO_FN(ieo_fftrn_)(IEO_R *data, IEO_R *nyq, int N, int forward)
{
    IEO_FN(ieo_cpx_) *c = (IEO_FN(ieo_cpx_) *)data;
    IEO_FN(ieo_cpx_) *cn = (IEO_FN(ieo_cpx_) *)nyq;
    const IEO_R ang = ((IEO_R)4 * IEO_ASIN((IEO_R)1)) / (IEO_R)N;
    const IEO_R cr = IEO_COS(ang), sr = IEO_SIN(ang);
    const int size[3] = {N, N, N / 2};
    const int rowlen = N / 2;
    const int total = N * N * rowlen;
    int ind0 = 0, ind1 = 0;   /* multi-index over the two outer dimensions */
    int neg = 0;              /* start of the row at (-i, -j) */

    if (forward == 1) {
        IEO_FN(ieo_fftc3_)(data, size, 1);
        for (int r = 0; r < total / rowlen; ++r) cn[r] = c[r * rowlen];
    }

    for (int row = 0; row < total; row += rowlen) {
        IEO_R wr = (IEO_R)1, wi = (IEO_R)0;
        for (int b = 1; b < N / 4; ++b) {
            const IEO_R owr = wr, owi = wi;
            wr = cr * owr - sr * owi;
            wi = cr * owi + sr * owr;
            const IEO_FN(ieo_cpx_) u = c[row + b];
            const IEO_FN(ieo_cpx_) v = c[neg + N / 2 - b];
            c[row + b].re = (IEO_R)0.5 * (u.re + v.re + (IEO_R)forward * wr * (u.im + v.im) + wi * (u.re - v.re));
            c[row + b].im = (IEO_R)0.5 * (u.im - v.im - (IEO_R)forward * wr * (u.re - v.re) + wi * (u.im + v.im));
            c[neg + N / 2 - b].re = (IEO_R)0.5 * (u.re + v.re - (IEO_R)forward * wr * (u.im + v.im) - wi * (u.re - v.re));
            c[neg + N / 2 - b].im = (IEO_R)0.5 * (-u.im + v.im - (IEO_R)forward * wr * (u.re - v.re) + wi * (u.im + v.im));
        }
        {
            const IEO_FN(ieo_cpx_) u = c[row];
            const IEO_FN(ieo_cpx_) v = cn[neg / rowlen];
            c[row].re = (IEO_R)0.5 * (u.re + v.re + (IEO_R)forward * (u.im + v.im));
            c[row].im = (IEO_R)0.5 * (u.im - v.im - (IEO_R)forward * (u.re - v.re));
            cn[neg / rowlen].re = (IEO_R)0.5 * (u.re + v.re - (IEO_R)forward * (u.im + v.im));
            cn[neg / rowlen].im = (IEO_R)0.5 * (-u.im + v.im - (IEO_R)forward * (u.re - v.re));
        }
        /* advance (ind0, ind1) and keep `neg` at the row of (-ind0, -ind1) */
        int step = rowlen;
        int dim = 1;
        int *ind[2] = {&ind0, &ind1};
        for (; dim >= 0 && *ind[dim] == size[dim] - 1; --dim) {
            *ind[dim] = 0;
            neg -= step;
            step *= size[dim];
        }
        if (dim >= 0) {
            if (*ind[dim] == 0) neg += step * (size[dim] - 1);
            else neg -= step;
            ++*ind[dim];
        }
    }

    if (forward == -1) IEO_FN(ieo_fftc3_)(data, size, -1);
}

#undef IEO_FN
#undef IEO_CAT
#undef IEO_CAT2
