// ie_tabindex.h -- O(1) bracket lookup of a tabulated potential.
//
// The reference finds the interval of a field value in its descending grid x by
// bisection or by a walk from the previous lookup's interval
// (src/potential.cpp:11-49): the unique l with x[l-1] >= v > x[l].  A lattice
// kernel that walks per site spends its time in divergent loops and dependent
// loads.  This index finds l with three loads and no loop:
//   - level 1: K1 uniform buckets over [x[n-1], x[0]];
//   - level 2: bucket b split into m_b uniform sub-buckets, m_b chosen so that a
//     sub-bucket is narrower than every grid interval it touches, so it holds at
//     most one grid point.  A sub-bucket keeps the interval l of its lowest value
//     and g = x[l-1], the one grid point that can lie inside it: the interval of
//     v is l - (v > g);
//   - the interval's record: its two ends (for the check x[l-1] >= v > x[l]),
//     and V / B^2, V' / B^2 at x[l] with the slopes of the interval.
// A value whose candidate fails the check (within a rounding of a sub-bucket
// edge, at v = x[n-1] exactly, or outside the table) takes the bisection, so the
// interval found is always the reference's; the interpolation
// y_l + (v - x_l) s_l is the lattice kernels' form of src/potential.cpp:52-68.
//
// Plain structs and host/device inline code, so that the host test compiles the
// same candidate arithmetic with g++.
#pragma once

#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#ifndef IE_HD
#if defined(__HIPCC__)
#define IE_HD __host__ __device__
#else
#define IE_HD
#endif
#endif

namespace ie {

// Interval l of the table: its ends and the lattice form of V and V' on it.
struct TabRec {
    double xhi, xlo;   // x[l-1], x[l]
    double ydv, sdv;   // V'(x[l]) / B^2 and its slope over the interval
    double yv, sv;     // V(x[l]) / B^2 and its slope
};

struct TabBucket {
    int32_t off, m;    // first sub-bucket, sub-bucket count
};

struct TabSub {
    double g;          // x[l - 1]
    int32_t l;         // interval of the sub-bucket's lowest value
    int32_t pad;
};

// Device (or host) view of one table's index.
struct TabIndex {
    const TabRec *rec;      // n records (record 0 unused: its check always fails)
    const TabRec *vrec;     // where per-value lookups read the records: an LDS copy, or rec
    const TabBucket *l1;    // k1 buckets
    const TabSub *l2;       // sub-buckets
    const double *x;        // the grid (bisection)
    int n, k1;
    int nsub;               // sub-buckets in l2
    double x0, inv_w;       // x[n-1] and k1 / (x[0] - x[n-1])
    double tmax;            // the largest double below k1
};

// The candidate interval of v (no loads beyond the two index levels).  NaN and
// values outside the table land in an end bucket; their check fails.
IE_HD inline int tab_candidate(const TabIndex &T, double v)
{
    const double t = fmin(fmax((v - T.x0) * T.inv_w, 0.0), T.tmax);
    const int b = (int)t;
    const TabBucket B = T.l1[b];
    int j = (int)((t - (double)b) * (double)B.m);
    j = j < B.m ? j : B.m - 1;
    const TabSub S = T.l2[B.off + j];
    return S.l - (v > S.g ? 1 : 0);
}

IE_HD inline bool tab_in(const TabRec &r, double v) { return v <= r.xhi && v > r.xlo; }

// The reference's bisection (src/potential.cpp:28-37): the interval of v, or
// 1 with *err set outside [x[n-1], x[0]].
IE_HD inline int tab_bisect_ref(const double *x, int n, double v, int *err)
{
    if (v > x[0] || v < x[n - 1]) {
        *err = 1;
        return 1;
    }
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        if (x[mid] >= v) lo = mid;
        else hi = mid;
    }
    return hi;
}

// Host-side build of the index from a descending grid x of n >= 2 points and
// the tabulated V, V' (scaled by invB2 as the lattice kernels use them).
struct TabIndexHost {
    std::vector<TabRec> rec;
    std::vector<TabBucket> l1;
    std::vector<TabSub> l2;
    int k1 = 0;
    double x0 = 0.0, inv_w = 0.0, tmax = 0.0;

    size_t bytes() const
    {
        return rec.size() * sizeof(TabRec) + l1.size() * sizeof(TabBucket) + l2.size() * sizeof(TabSub);
    }
};

// Sub-bucket counts of k1 buckets: a sub-bucket narrower than every interval the
// bucket touches (at most kMaxSub per bucket; a denser bucket leaves some values
// to the bisection).
inline long tab_sub_counts(const double *x, int n, int k1, std::vector<int32_t> *m)
{
    constexpr long kMaxSub = 1 << 14;
    const double x0 = x[n - 1], w = (x[0] - x[n - 1]) / k1;
    if (m) m->assign(k1, 1);
    long total = 0;
    int l = n - 1;   // interval [x[l], x[l-1]] containing the bucket's low end
    for (int b = 0; b < k1; ++b) {
        const double lo = x0 + b * w, hi = lo + w;
        while (l > 1 && x[l - 1] <= lo) --l;
        double ms = x[l - 1] - x[l];
        for (int q = l; q >= 1 && x[q] < hi; --q) ms = fmin(ms, x[q - 1] - x[q]);
        long mb = ms > 0.0 ? (long)floor(w / ms) + 1 : kMaxSub;
        mb = mb < kMaxSub ? mb : kMaxSub;
        if (m) (*m)[b] = (int32_t)mb;
        total += mb;
    }
    return total;
}

// Level-1 size: the power of two in [16, 4096] that keeps the index smallest;
// *subs gets its sub-bucket count.
inline int tab_k1(const double *x, int n, long *subs)
{
    long best = -1;
    int k1 = 16;
    for (int k = 16; k <= 4096; k *= 2) {
        const long s = tab_sub_counts(x, n, k, nullptr);
        const long bytes = k * (long)sizeof(TabBucket) + s * (long)sizeof(TabSub);
        if (best < 0 || bytes < best) {
            best = bytes;
            k1 = k;
            if (subs) *subs = s;
        }
    }
    return k1;
}

inline TabIndexHost tab_index_build(const double *x, const double *v, const double *dv, int n, double invB2)
{
    TabIndexHost H;
    H.k1 = tab_k1(x, n, nullptr);
    std::vector<int32_t> m;
    tab_sub_counts(x, n, H.k1, &m);
    H.x0 = x[n - 1];
    H.inv_w = H.k1 / (x[0] - x[n - 1]);
    H.tmax = nextafter((double)H.k1, 0.0);
    H.rec.resize(n);
    H.rec[0] = TabRec{x[0], x[0], 0.0, 0.0, 0.0, 0.0};
    for (int l = 1; l < n; ++l) {
        const double d = x[l - 1] - x[l];
        H.rec[l] = TabRec{x[l - 1], x[l], dv[l] * invB2, (dv[l - 1] - dv[l]) / d * invB2, v[l] * invB2,
                          (v[l - 1] - v[l]) / d * invB2};
    }
    H.l1.resize(H.k1);
    int32_t off = 0;
    for (int b = 0; b < H.k1; ++b) {
        H.l1[b] = TabBucket{off, m[b]};
        off += m[b];
    }
    H.l2.resize(off);
    int unused = 0;
    for (int b = 0; b < H.k1; ++b)
        for (int j = 0; j < m[b]; ++j) {
            // the lowest value of the sub-bucket, as tab_candidate maps values
            const double s = H.x0 + ((double)b + (double)j / m[b]) / H.inv_w;
            int l = tab_bisect_ref(x, n, fmin(fmax(s, x[n - 1]), x[0]), &unused);
            H.l2[H.l1[b].off + j] = TabSub{x[l - 1], l, 0};
        }
    return H;
}

inline TabIndex tab_view(const TabIndexHost &H, const TabRec *rec, const TabBucket *l1, const TabSub *l2,
                         const double *x, int n)
{
    return TabIndex{rec, rec, l1, l2, x, n, H.k1, (int)H.l2.size(), H.x0, H.inv_w, H.tmax};
}

}  // namespace ie
