// Host check of the tabulated potential's interval index (ie_tabindex.h), built
// with g++ from the header the kernels use.  For a grid read from a file (the
// reference's inputs/ table) and synthetic grids (uniform, geometric, random and
// two-point), every probe value v gets the index's candidate; a candidate that
// passes the check x[l-1] >= v > x[l] must be the bisection's interval, and the
// share of values that fail the check (and take the bisection) is printed per
// grid.  Probes: random values, every grid point, the midpoints, and the two
// neighbouring doubles of each grid point and of each sub-bucket edge.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <random>
#include <vector>

#include "ie_tabindex.h"

using ie::TabIndex;
using ie::TabIndexHost;

namespace {

int g_bad = 0;

struct Tally {
    long probes = 0, slow = 0;
};

void probe(const TabIndex &T, const std::vector<double> &x, double v, Tally &t)
{
    if (!(v <= x[0] && v >= x.back())) return;   // outside the table: not a lookup
    ++t.probes;
    int err = 0;
    const int want = ie::tab_bisect_ref(x.data(), (int)x.size(), v, &err);
    const int l = ie::tab_candidate(T, v);
    if (l < 0 || l >= (int)x.size()) {
        fprintf(stderr, "candidate %d out of range for v = %.17g\n", l, v);
        ++g_bad;
        return;
    }
    if (ie::tab_in(T.rec[l], v)) {
        if (l != want) {
            fprintf(stderr, "candidate %d passes the check but the interval is %d (v = %.17g)\n", l, want, v);
            ++g_bad;
        }
    } else {
        ++t.slow;
    }
}

// Returns the share of probes that take the bisection.
double check_grid(const char *name, const std::vector<double> &x, unsigned seed, double max_slow)
{
    const int n = (int)x.size();
    std::vector<double> v(n), dv(n);
    for (int i = 0; i < n; ++i) {
        v[i] = 1.0 + 0.1 * x[i] * x[i];
        dv[i] = 0.2 * x[i];
    }
    const TabIndexHost H = ie::tab_index_build(x.data(), v.data(), dv.data(), n, 1.0);
    const TabIndex T = ie::tab_view(H, H.rec.data(), H.l1.data(), H.l2.data(), x.data(), n);
    Tally rnd, pts;
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> U(x.back(), x[0]);
    for (int q = 0; q < 400000; ++q) probe(T, x, U(g), rnd);
    for (int i = 0; i < n; ++i) {
        probe(T, x, x[i], pts);
        probe(T, x, nextafter(x[i], INFINITY), pts);
        probe(T, x, nextafter(x[i], -INFINITY), pts);
        if (i > 0) probe(T, x, 0.5 * (x[i] + x[i - 1]), pts);
    }
    for (int b = 0; b < H.k1; ++b)
        for (int j = 0; j < H.l1[b].m; ++j) {
            const double s = H.x0 + ((double)b + (double)j / H.l1[b].m) / H.inv_w;
            probe(T, x, s, pts);
            probe(T, x, nextafter(s, INFINITY), pts);
            probe(T, x, nextafter(s, -INFINITY), pts);
        }
    const double share = (double)rnd.slow / (double)rnd.probes;
    printf("%s: n %d k1 %d subs %zu bytes %zu | random %ld probes, %.2e to bisection | edges %ld probes, %.2e\n",
           name, n, H.k1, H.l2.size(), H.bytes(), rnd.probes, share, pts.probes,
           (double)pts.slow / (double)(pts.probes ? pts.probes : 1));
    if (share > max_slow) {
        fprintf(stderr, "%s: %.2e of random values take the bisection (bound %.1e)\n", name, share, max_slow);
        ++g_bad;
    }
    return share;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc > 1) {
        FILE *f = fopen(argv[1], "rb");
        if (!f) {
            perror(argv[1]);
            return 2;
        }
        std::vector<double> x;
        double d;
        while (fread(&d, sizeof d, 1, f) == 1) x.push_back(d);
        fclose(f);
        check_grid("reference table", x, 1, 1e-4);
    }
    {
        std::vector<double> x(1001);
        for (int i = 0; i <= 1000; ++i) x[i] = 2.0 - 1e-3 * i;
        check_grid("uniform", x, 2, 1e-4);
    }
    {
        std::vector<double> x(300);
        for (int i = 0; i < 300; ++i) x[i] = 1.0 + exp(-0.05 * i);   // spacing 1 : 3e6 from end to end
        check_grid("geometric", x, 3, 1e-4);
    }
    {
        std::mt19937_64 g(4);
        std::uniform_real_distribution<double> U(1e-6, 1e-2);
        std::vector<double> x(2000);
        x[0] = 5.0;
        for (int i = 1; i < 2000; ++i) x[i] = x[i - 1] - U(g);
        check_grid("random spacing", x, 5, 1e-4);
    }
    {
        std::vector<double> x = {1.0, -1.0};
        check_grid("two points", x, 6, 1e-4);
    }
    if (g_bad) {
        fprintf(stderr, "tabindex_test: %d failures\n", g_bad);
        return 1;
    }
    printf("tabindex_test ok\n");
    return 0;
}
