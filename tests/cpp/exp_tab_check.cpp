// Host build of hd::exp_tab (hadros_amd/csrc/fastmath.h) against the
// long-double exponential: prints the largest error in ulps of the correctly
// rounded result over a seeded sample, then the special values.
#include "fastmath.h"
#include <cmath>
#include <cstdio>
#include <random>

int main(int argc, char** argv)
{
    const long n = argc > 1 ? std::atol(argv[1]) : 2000000;
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    double max_ulp = 0.0, worst = 0.0;
    for (long i = 0; i < n; ++i) {
        const double t = u(g);
        double x;
        switch (i % 4) {
        case 0: x = -745.0 + 1454.7 * t; break;      // the whole finite range
        case 1: x = -8.0 + 16.0 * t; break;          // near 1
        case 2: x = -1.0e-3 + 2.0e-3 * t; break;     // |r| only
        default: x = -745.0 + 37.0 * t; break;       // subnormal results
        }
        const double a = hd::exp_tab(x, hd::EXP_TAB);
        const long double e = expl(static_cast<long double>(x));
        const double ed = static_cast<double>(e);
        if (ed == 0.0 && a == 0.0) continue;
        const double ulp = std::nextafter(std::fabs(ed), INFINITY) - std::fabs(ed);
        const double err = static_cast<double>(fabsl(static_cast<long double>(a) - e)) / ulp;
        if (err > max_ulp) {
            max_ulp = err;
            worst = x;
        }
    }
    std::printf("max_ulp %.6f %.17g\n", max_ulp, worst);
    std::printf("special %a %a %a %a %a %d\n", hd::exp_tab(0.0, hd::EXP_TAB), hd::exp_tab(1.0, hd::EXP_TAB),
                hd::exp_tab(-INFINITY, hd::EXP_TAB), hd::exp_tab(-746.0, hd::EXP_TAB),
                hd::exp_tab(709.0, hd::EXP_TAB), int(std::isnan(hd::exp_tab(NAN, hd::EXP_TAB))));
    return 0;
}
