// Host build of hd::sincos_m, hd::sincos_shift and hd::cos_polar
// (hadros_amd/csrc/fastmath.h) against the long-double sine and cosine: prints
// the largest absolute error (in units of 2^-53) of each over seeded samples,
// the range flags of sincos_shift, then cos_polar's error.
#include "fastmath.h"
#include <cmath>
#include <cstdio>
#include <random>

static double err53(double got, long double want) { return double(fabsl((long double)got - want) * 0x1p53L); }

int main()
{
    std::mt19937_64 g(11);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    double e_fused = 0.0, e_plain = 0.0, e_shift = 0.0;
    for (int i = 0; i < 1000000; ++i) {
        const double x = -12.6 + 25.2 * u(g);   // +-4 pi: theta and its stage offsets
        const long double sx = sinl((long double)x), cx = cosl((long double)x);
        double s, c;
        hd::sincos_m<true>(x, &s, &c);
        e_fused = std::fmax(e_fused, std::fmax(err53(s, sx), err53(c, cx)));
        hd::sincos_m<false>(x, &s, &c);
        e_plain = std::fmax(e_plain, std::fmax(err53(s, sx), err53(c, cx)));
        // a stage: theta in [0, pi], offset |d| <= 1/4
        const double th = 3.14159 * u(g), d = -0.25 + 0.5 * u(g);
        double s0, c0;
        hd::sincos_m<true>(th, &s0, &c0);
        if (!hd::sincos_shift(s0, c0, d, &s, &c)) return 1;
        const long double a = (long double)th + (long double)d;
        e_shift = std::fmax(e_shift, std::fmax(err53(s, sinl(a)), err53(c, cosl(a))));
    }
    double s, c;
    const bool out = hd::sincos_shift(0.5, 0.8, 0.2500001, &s, &c) || hd::sincos_shift(0.5, 0.8, -0.26, &s, &c) ||
                     hd::sincos_shift(0.5, 0.8, NAN, &s, &c);
    std::printf("err53 %.4f %.4f %.4f %d\n", e_fused, e_plain, e_shift, int(out));
    // cos_polar over [0, pi], its ends, near pi/2, and angles just outside
    double e_polar = 0.0;
    for (int i = 0; i < 1000000; ++i) {
        const double th = 3.141592653589793 * u(g);
        e_polar = std::fmax(e_polar, err53(hd::cos_polar(th), cosl((long double)th)));
    }
    const double pts[] = {0.0, 3.141592653589793, 1.5707963267948966, 1.5707963267948966 + 1e-12, 1.5707963267948966 - 3e-9,
                          -1e-3, -2.0, 3.1426, 5.0, 6.28};
    for (double th : pts) e_polar = std::fmax(e_polar, err53(hd::cos_polar(th), cosl((long double)th)));
    std::printf("polar %.4f\n", e_polar);
    return 0;
}
