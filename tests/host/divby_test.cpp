// Host check of div_by (inflation-easy_amd/csrc/ie_numeric.h): x / b through the
// correctly rounded reciprocal must equal the IEEE quotient bit for bit.  Divisors:
// random doubles over many decades, the squares of rescale factors B the engine
// uses, and the interval widths of a grid read from a file (the reference's
// inputs/ table); numerators: random values over the quotient range of use,
// and the table's potential values.
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <random>
#include <vector>

#include "ie_numeric.h"

static long g_bad = 0, g_checked = 0;

static void check(double x, double b)
{
    const double want = x / b, got = ie::div_by(x, b, 1.0 / b);
    ++g_checked;
    if (memcmp(&want, &got, sizeof want) != 0) {
        if (g_bad < 10) fprintf(stderr, "x %.17g b %.17g: %.17g != %.17g\n", x, b, got, want);
        ++g_bad;
    }
}

int main(int argc, char **argv)
{
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> U(0.5, 2.0), E(-60.0, 60.0);
    std::vector<double> divisors;
    for (int i = 0; i < 2000; ++i) divisors.push_back(U(g) * std::ldexp(1.0, (int)E(g)));
    for (double B : {1.0, 1e-5, 3.3e-7, 5.5e-5, 0.1234567, 1.0 / 3.0, 7.0}) divisors.push_back(B * B);
    std::vector<double> xs;
    if (argc > 1) {
        FILE *f = fopen(argv[1], "rb");
        if (!f) return 2;
        std::vector<double> t;
        double d;
        while (fread(&d, sizeof d, 1, f) == 1) t.push_back(d);
        fclose(f);
        const size_t n = t.size() / 3;   // x, V, V' of the table
        for (size_t l = 1; l < n; ++l) divisors.push_back(t[l - 1] - t[l]);
        for (size_t l = n; l < t.size(); ++l) xs.push_back(t[l]);
    }
    for (double b : divisors) {
        for (int q = 0; q < 2000; ++q) check(U(g) * std::ldexp(1.0, (int)E(g)) * (q & 1 ? -1.0 : 1.0), b);
        for (double x : xs) check(x, b);
        check(0.0, b);
        check(b, b);
        check(-3.0 * b, b);
    }
    printf("divby_test: %ld quotients, %ld differ\n", g_checked, g_bad);
    return g_bad ? 1 : 0;
}
