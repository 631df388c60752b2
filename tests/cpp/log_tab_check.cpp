// Host build of hd::log_tab (hadros_amd/csrc/fastmath.h) against the
// long-double logarithm: prints the largest error in ulps of the correctly
// rounded result over a seeded sample, then the special values.
#include "fastmath.h"
#include <cmath>
#include <cstdio>
#include <random>

int main(int argc, char** argv)
{
    const long n = argc > 1 ? std::atol(argv[1]) : 2000000;
    std::mt19937_64 g(11);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    double max_ulp = 0.0, worst = 0.0;
    for (long i = 0; i < n; ++i) {
        const double t = u(g);
        double x;
        switch (i % 5) {
        case 0: x = std::exp2(-1074.0 + 2097.0 * t); break;   // the whole range, subnormals included
        case 1: x = 0.5 + 1.5 * t; break;                     // around 1
        case 2: x = 1.0 + (t - 0.5) * 1.0e-6; break;          // next to 1: relative accuracy
        case 3: x = 0.74 + 0.02 * t; break;                   // the normalisation boundary
        default: x = 1.0e-4 + 1.0e3 * t; break;               // radii ratios and redshifts
        }
        const double a = hd::log_tab(x, hd::LOG_TAB);
        const long double e = logl(static_cast<long double>(x));
        const double ed = static_cast<double>(e);
        if (ed == 0.0) {
            if (a != 0.0) max_ulp = 1e9;
            continue;
        }
        const double ulp = std::nextafter(std::fabs(ed), INFINITY) - std::fabs(ed);
        const double err = static_cast<double>(fabsl(static_cast<long double>(a) - e)) / ulp;
        if (err > max_ulp) {
            max_ulp = err;
            worst = x;
        }
    }
    std::printf("max_ulp %.6f %.17g\n", max_ulp, worst);
    std::printf("special %a %a %a %a %d %d %a\n", hd::log_tab(1.0, hd::LOG_TAB), hd::log_tab(2.0, hd::LOG_TAB),
                hd::log_tab(0.0, hd::LOG_TAB), hd::log_tab(INFINITY, hd::LOG_TAB),
                int(std::isnan(hd::log_tab(-1.0, hd::LOG_TAB))), int(std::isnan(hd::log_tab(NAN, hd::LOG_TAB))),
                hd::log_tab(0.75, hd::LOG_TAB));
    return 0;
}
