// fastmath.h -- FP64 elementary functions of the kernels: the transfer
// kernels' table-driven exponential and logarithm, the geodesic kernels'
// sine/cosine and their angle-addition stage form, and refined hardware
// reciprocals.
//
// The exponential:
//
// e^x by Tang's table-driven method (P. T. P. Tang, "Table-driven
// implementation of the exponential function in IEEE floating-point
// arithmetic", ACM TOMS 15(2):144-157, 1989) with a 64-entry table:
//
//   x = (64 m + j) ln2/64 + r,   |r| <= ln2/128,   e^x = 2^m 2^(j/64) e^r,
//
// e^r by its degree-5 Taylor polynomial (truncation below 3.6e-17 relative
// for |r| <= ln2/128).  The reduction x - k ln2/64 uses a two-part ln2/64
// whose high part has 33 significant bits, so k L1 is exact for |k| < 2^17
// (all |x| < 745).  Error: the table entry (0.5 ulp), the polynomial's last
// rounding (0.5 ulp) and the product (0.5 ulp): under 2 ulp.  About 18 VALU
// instructions and one LDS read, against about 22-31 for the general exp,
// whose degree-11 polynomial covers |r| <= ln2/2.
//
// Domain: finite x (x > 709.78 gives +inf) and x = -inf; every x below
// -745.2 gives +0.  The kernels stage EXP_TAB into LDS and pass it as T.
// The same code compiles for the host, where the unit test checks it
// against a long-double exponential.
#pragma once

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define HD_FM_QUAL __host__ __device__ __forceinline__
#else
#include <cmath>
#define HD_FM_QUAL inline
#endif

namespace hd {

// 2^(j/64), j = 0..63, each correctly rounded (from 60-digit decimal arithmetic)
constexpr double EXP_TAB[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0,
};

constexpr double EXP_INV_L = 0x1.71547652b82fep+6;   // 64 / ln 2
constexpr double EXP_L1 = 0x1.62e42fef00000p-7;      // ln2/64, high 33 bits
constexpr double EXP_L2 = 0x1.473de6af278edp-40;     // ln2/64 - EXP_L1
constexpr double EXP_ZERO_BELOW = -745.2;            // e^x < 2^-1075 below this

// exp_tab for arguments known to lie above EXP_ZERO_BELOW (no underflow test)
HD_FM_QUAL double exp_tab_nz(double x, const double* T)
{
    const double k = rint(x * EXP_INV_L);
    const double r = fma(-k, EXP_L2, fma(-k, EXP_L1, x));
    const int ki = int(k);
    const double p =
        fma(r, fma(r, fma(r, fma(r, fma(r, 1.0 / 120.0, 1.0 / 24.0), 1.0 / 6.0), 0.5), 1.0), 1.0);
    return ldexp(T[ki & 63] * p, ki >> 6);
}

HD_FM_QUAL double exp_tab(double x, const double* T)
{
    const double k = rint(x * EXP_INV_L);
    const double r = fma(-k, EXP_L2, fma(-k, EXP_L1, x));
    const int ki = int(k);
    const double p =
        fma(r, fma(r, fma(r, fma(r, fma(r, 1.0 / 120.0, 1.0 / 24.0), 1.0 / 6.0), 0.5), 1.0), 1.0);
    const double v = ldexp(T[ki & 63] * p, ki >> 6);
    return x < EXP_ZERO_BELOW ? 0.0 : v;
}

// ------------------------------------------------------------------ log
// ln x by Tang's table-driven method (P. T. P. Tang, "Table-driven
// implementation of the logarithm function in IEEE floating-point
// arithmetic", ACM TOMS 16(4):378-400, 1990), the mantissa normalised to
// [0.75, 1.5) so that k ln2 and ln F never cancel by more than a factor 2.4:
//
//   x = 2^k m,  m in [0.75, 1.5),  F = 1 + j/64 nearest m (j = -16..32),
//   r = (m - F) / F,  |r| <= 1/96,  ln x = k ln2 + ln F + ln(1 + r).
//
// m - F is exact (Sterbenz), r = (m - F) (1/F) with the tabulated reciprocal
// (exact for F = 1, so ln x keeps its relative accuracy next to x = 1),
// ln(1 + r) by its degree-9 Taylor polynomial (truncation below 2e-18
// relative), and k ln2 in two parts whose high part has 42 significant bits
// (k LOG_LN2_HI exact for |k| < 2^11).  Error under 2 ulp (tested against a
// long-double logarithm on the host build).  About 25 VALU instructions and
// one 16-byte LDS read, against about 90 for the general log.
//
// Domain: every double; 0 gives -inf, +inf gives +inf, x < 0 and NaN give
// NaN.  The kernels stage LOG_TAB into LDS and pass it as L.

// (1/F, ln F) for F = 1 + j/64, j = -16..32, each correctly rounded (from
// 60-digit decimal arithmetic)
constexpr int LOG_TAB_N = 49;
constexpr double LOG_TAB[2 * LOG_TAB_N] = {
    0x1.5555555555555p+0, -0x1.269621134db92p-2, 0x1.4e5e0a72f0539p+0, -0x1.1178e8227e47cp-2,
    0x1.47ae147ae147bp+0, -0x1.f991c6cb3b379p-3, 0x1.4141414141414p+0, -0x1.d1037f2655e7bp-3,
    0x1.3b13b13b13b14p+0, -0x1.a93ed3c8ad9e3p-3, 0x1.3521cfb2b78c1p+0, -0x1.823c16551a3c2p-3,
    0x1.2f684bda12f68p+0, -0x1.5bf406b543db2p-3, 0x1.29e4129e4129ep+0, -0x1.365fcb0159016p-3,
    0x1.2492492492492p+0, -0x1.1178e8227e47cp-3, 0x1.1f7047dc11f70p+0, -0x1.da727638446a2p-4,
    0x1.1a7b9611a7b96p+0, -0x1.9335e5d594989p-4, 0x1.15b1e5f75270dp+0, -0x1.4d3115d207eacp-4,
    0x1.1111111111111p+0, -0x1.08598b59e3a07p-4, 0x1.0c9714fbcda3bp+0, -0x1.894aa149fb343p-5,
    0x1.0842108421084p+0, -0x1.0415d89e74444p-5, 0x1.0410410410410p+0, -0x1.0205658935847p-6,
    0x1.0000000000000p+0, 0x0.0p+0, 0x1.f81f81f81f820p-1, 0x1.fc0a8b0fc03e4p-7,
    0x1.f07c1f07c1f08p-1, 0x1.f829b0e783300p-6, 0x1.e9131abf0b767p-1, 0x1.77458f632dcfcp-5,
    0x1.e1e1e1e1e1e1ep-1, 0x1.f0a30c01162a6p-5, 0x1.dae6076b981dbp-1, 0x1.341d7961bd1d1p-4,
    0x1.d41d41d41d41dp-1, 0x1.6f0d28ae56b4cp-4, 0x1.cd85689039b0bp-1, 0x1.a926d3a4ad563p-4,
    0x1.c71c71c71c71cp-1, 0x1.e27076e2af2e6p-4, 0x1.c0e070381c0e0p-1, 0x1.0d77e7cd08e59p-3,
    0x1.bacf914c1bad0p-1, 0x1.29552f81ff523p-3, 0x1.b4e81b4e81b4fp-1, 0x1.44d2b6ccb7d1ep-3,
    0x1.af286bca1af28p-1, 0x1.5ff3070a793d4p-3, 0x1.a98ef606a63bep-1, 0x1.7ab890210d909p-3,
    0x1.a41a41a41a41ap-1, 0x1.9525a9cf456b4p-3, 0x1.9ec8e951033d9p-1, 0x1.af3c94e80bff3p-3,
    0x1.999999999999ap-1, 0x1.c8ff7c79a9a22p-3, 0x1.948b0fcd6e9e0p-1, 0x1.e27076e2af2e6p-3,
    0x1.8f9c18f9c18fap-1, 0x1.fb9186d5e3e2bp-3, 0x1.8acb90f6bf3aap-1, 0x1.0a324e27390e3p-2,
    0x1.8618618618618p-1, 0x1.1675cababa60ep-2, 0x1.8181818181818p-1, 0x1.22941fbcf7966p-2,
    0x1.7d05f417d05f4p-1, 0x1.2e8e2bae11d31p-2, 0x1.78a4c8178a4c8p-1, 0x1.3a64c556945eap-2,
    0x1.745d1745d1746p-1, 0x1.4618bc21c5ec2p-2, 0x1.702e05c0b8170p-1, 0x1.51aad872df82dp-2,
    0x1.6c16c16c16c17p-1, 0x1.5d1bdbf5809cap-2, 0x1.6816816816817p-1, 0x1.686c81e9b14afp-2,
    0x1.642c8590b2164p-1, 0x1.739d7f6bbd007p-2, 0x1.6058160581606p-1, 0x1.7eaf83b82afc3p-2,
    0x1.5c9882b931057p-1, 0x1.89a3386c1425bp-2, 0x1.58ed2308158edp-1, 0x1.947941c2116fbp-2,
    0x1.5555555555555p-1, 0x1.9f323ecbf984cp-2,
};
constexpr double LOG_LN2_HI = 0x1.62e42fefa3800p-1;   // ln 2, high 42 bits
constexpr double LOG_LN2_LO = 0x1.ef35793c76730p-45;  // ln 2 - LOG_LN2_HI

HD_FM_QUAL double log_tab(double x, const double* L)
{
    int e;
    const double m0 = frexp(x, &e);   // x = m0 2^e, m0 in [0.5, 1) for finite x > 0
    const bool lo_half = m0 < 0.75;
    const double m = lo_half ? 2.0 * m0 : m0;
    const double k = double(lo_half ? e - 1 : e);
    const double jf = rint((m - 1.0) * 64.0);
    const int j = int(jf) + 16;
    const double f = m - (1.0 + jf * (1.0 / 64.0));
    const double r = f * L[2 * j];
    double q = 1.0 / 9.0;
    q = fma(r, q, -1.0 / 8.0);
    q = fma(r, q, 1.0 / 7.0);
    q = fma(r, q, -1.0 / 6.0);
    q = fma(r, q, 1.0 / 5.0);
    q = fma(r, q, -1.0 / 4.0);
    q = fma(r, q, 1.0 / 3.0);
    q = fma(r, q, -0.5);
    const double p = fma(r * r, q, r);
    const double hi = fma(k, LOG_LN2_HI, L[2 * j + 1]);
    const double lo = fma(k, LOG_LN2_LO, p);
    const double v = hi + lo;
    // specials: frexp leaves 0, inf and NaN unchanged (and m0 < 0 for x < 0)
    return (x > 0.0 && x < INFINITY) ? v : (x == 0.0 ? -INFINITY : (x == INFINITY ? x : NAN));
}

// ------------------------------------------------------------------ sincos
// sin and cos of a moderate argument (|x| < 2^20 pi/2; the path only feeds
// angles theta, theta +- 1e-5 that stay within a few pi).  Cody-Waite
// reduction by pi/2 split into P1 + P2 + P3 (33 + 33 + 53 significant bits, so
// n * P1 and n * P2 are exact), then the Taylor series of sin and cos on
// [-pi/4, pi/4] to degree 17 / 18 (truncation < 1e-19 relative).  About 1 ulp,
// and far fewer instructions and registers than a general double sincos with
// a large-argument reduction path.  FUSED (the default) evaluates the
// reduction and the Horner steps as fused multiply-adds, 20 % fewer
// instructions; the written multiply/add form is kept for the density's cos
// theta, where the fused form's schedule made the transfer kernel spill.
template <bool FUSED = true>
HD_FM_QUAL void sincos_m(double x, double* sp, double* cp)
{
    constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    constexpr double P1 = 0x1.921fb54400000p+0;
    constexpr double P2 = 0x1.0b4611a600000p-34;
    constexpr double P3 = 0x1.3198a2e037073p-69;
    const double fn = rint(x * TWO_OVER_PI);
    double y, sn, cs;
    if constexpr (FUSED) {
        y = fma(-fn, P3, fma(-fn, P2, fma(-fn, P1, x)));   // n P1 and n P2 exact
        const double z = y * y;
        // Horner steps as fused multiply-adds (one rounding each)
        // sin y = y + y z (-1/3! + z (1/5! - z (1/7! ...)))
        double ps = 1.0 / 355687428096000.0;
        ps = fma(z, ps, -1.0 / 1307674368000.0);
        ps = fma(z, ps, 1.0 / 6227020800.0);
        ps = fma(z, ps, -1.0 / 39916800.0);
        ps = fma(z, ps, 1.0 / 362880.0);
        ps = fma(z, ps, -1.0 / 5040.0);
        ps = fma(z, ps, 1.0 / 120.0);
        ps = fma(z, ps, -1.0 / 6.0);
        sn = fma(y * z, ps, y);
        // cos y = 1 - z/2 + z^2 (1/4! - z (1/6! ...))
        double pc = -1.0 / 6402373705728000.0;
        pc = fma(z, pc, 1.0 / 20922789888000.0);
        pc = fma(z, pc, -1.0 / 87178291200.0);
        pc = fma(z, pc, 1.0 / 479001600.0);
        pc = fma(z, pc, -1.0 / 3628800.0);
        pc = fma(z, pc, 1.0 / 40320.0);
        pc = fma(z, pc, -1.0 / 720.0);
        pc = fma(z, pc, 1.0 / 24.0);
        const double hz = 0.5 * z;
        const double w = 1.0 - hz;
        cs = w + fma(z * z, pc, (1.0 - w) - hz);
    } else {
        y = ((x - fn * P1) - fn * P2) - fn * P3;
        const double z = y * y;
        const double ps = -1.0 / 6.0 + z * (1.0 / 120.0 + z * (-1.0 / 5040.0 + z * (1.0 / 362880.0 + z * (
                          -1.0 / 39916800.0 + z * (1.0 / 6227020800.0 + z * (-1.0 / 1307674368000.0 + z * (
                          1.0 / 355687428096000.0)))))));
        sn = y + y * z * ps;
        const double pc = 1.0 / 24.0 + z * (-1.0 / 720.0 + z * (1.0 / 40320.0 + z * (-1.0 / 3628800.0 + z * (
                          1.0 / 479001600.0 + z * (-1.0 / 87178291200.0 + z * (1.0 / 20922789888000.0 + z * (
                          -1.0 / 6402373705728000.0)))))));
        const double hz = 0.5 * z;
        const double w = 1.0 - hz;
        cs = w + (((1.0 - w) - hz) + z * z * pc);
    }
    const int q = int(static_cast<long long>(fn) & 3);
    const double s0 = (q & 1) ? cs : sn;
    const double c0 = (q & 1) ? sn : cs;
    *sp = (q & 2) ? -s0 : s0;
    *cp = ((q + 1) & 2) ? -c0 : c0;
}

// cos of a polar angle th, as sin y with y = pi/2 - t in [-pi/2, pi/2] for
// t = th folded into [0, pi] (|th|, then 2 pi - |th| above pi: exact for
// |th| <= 2 pi, the range of a ray's theta and far beyond it): y =
// (PIO2_HI - t) + PIO2_LO (the first difference is exact for t >= pi/4 by
// Sterbenz, within half an ulp of y below), then the odd Taylor series of
// sin y to degree 23 (truncation below 2e-18 for |y| <= pi/2).  About 30
// instructions against sincos_m's 60 in its plain form; within 2 x 2^-53 of
// the long-double cosine.
HD_FM_QUAL double cos_polar(double th)
{
    constexpr double PIO2_HI = 0x1.921fb54442d18p+0;   // pi/2 rounded to double
    constexpr double PIO2_LO = 0x1.1a62633145c07p-54;  // pi/2 - PIO2_HI
    const double a = fabs(th);
    const double t = (a > 2.0 * PIO2_HI) ? 4.0 * PIO2_HI - a : a;
    const double y = (PIO2_HI - t) + PIO2_LO;
    const double z = y * y;
    // written multiplies and adds, as sincos_m's plain form: fused, the
    // transfer kernel's schedule spilled
    const double ps = -1.0 / 6.0 + z * (1.0 / 120.0 + z * (-1.0 / 5040.0 + z * (1.0 / 362880.0 + z * (
                      -1.0 / 39916800.0 + z * (1.0 / 6227020800.0 + z * (-1.0 / 1307674368000.0 + z * (
                      1.0 / 355687428096000.0 + z * (-1.0 / 121645100408832000.0 + z * (
                      1.0 / 51090942171709440000.0 + z * (-1.0 / 25852016738884976640000.0))))))))));
    return y + y * z * ps;
}

// sin and cos of x + d from s0 = sin x, c0 = cos x, for |d| <= 1/4: the
// Taylor series of sin d to degree 11 and of cos d to degree 12 (truncation
// below 1e-17 relative), then the angle-addition formulas; about 20
// instructions against sincos_m's 35.  Returns false for |d| > 1/4 or NaN,
// whose results are to be discarded.
HD_FM_QUAL bool sincos_shift(double s0, double c0, double d, double* sp, double* cp)
{
    const double z = d * d;
    double ps = -1.0 / 39916800.0;
    ps = fma(z, ps, 1.0 / 362880.0);
    ps = fma(z, ps, -1.0 / 5040.0);
    ps = fma(z, ps, 1.0 / 120.0);
    ps = fma(z, ps, -1.0 / 6.0);
    const double sd = fma(d * z, ps, d);
    double pc = 1.0 / 479001600.0;
    pc = fma(z, pc, -1.0 / 3628800.0);
    pc = fma(z, pc, 1.0 / 40320.0);
    pc = fma(z, pc, -1.0 / 720.0);
    pc = fma(z, pc, 1.0 / 24.0);
    pc = fma(z, pc, -0.5);
    const double cdm1 = z * pc;   // cos d - 1
    *sp = fma(s0, cdm1, fma(c0, sd, s0));
    *cp = fma(c0, cdm1, fma(-s0, sd, c0));
    return fabs(d) <= 0.25;
}

#if defined(__HIP__)
// 1/x for finite normal x whose reciprocal is normal (no denormal or
// overflow scaling, unlike the IEEE division sequence): the hardware estimate
// refined by two Newton-Raphson steps r <- r + r (1 - x r), each of which
// doubles the correct bits; within an ulp of the correctly rounded value.
// x = 0 gives +-inf, as the division does.
__device__ __forceinline__ double fast_rcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    r = fma(fma(-x, r, 1.0), r, r);
    return r;
}

// 1/sqrt(x) for finite normal x > 0: the hardware estimate refined by two
// Newton-Raphson steps y <- y + y (1 - x y^2) / 2.  x < 0 gives NaN.
__device__ __forceinline__ double fast_rsqrt(double x)
{
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double e = fma(-x * y, y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}

// ln x for finite x > 0 to the accuracy of a float, for logarithms that are
// then rounded to float: x = m 2^e exactly (frexp, m in [0.5, 1)), and
// ln x = (e + log2 m) ln 2 with the hardware FP32 log2 of m (about 1e-7
// absolute, below the float rounding of the result).  About 8 instructions
// against log_tab's 25 and its table read.
__device__ __forceinline__ double ln_f32(double x)
{
    int e;
    const double m = frexp(x, &e);
    return (double(e) + double(__builtin_amdgcn_logf(float(m)))) * 0.69314718055994530942;
}

// e^x to the accuracy of a float, for values that are then rounded to float
// or enter only float-accurate results: x log2 e = k + f with k = floor and
// f in [0, 1), e^x = 2^k 2^f with the hardware FP32 exp2 of f (about 1e-7
// relative); 0 below about -745, +inf above about 709.  About 8 instructions
// against exp_tab's 17 and its table read.
__device__ __forceinline__ double exp_f32(double x)
{
    const double u = x * 1.4426950408889634074;
    const double t = (u < -1100.0) ? -1100.0 : ((u > 1100.0) ? 1100.0 : u);   // NaN stays NaN
    const double k = floor(t);
    return ldexp(double(__builtin_amdgcn_exp2f(float(t - k))), int(k));
}
#endif

}  // namespace hd
