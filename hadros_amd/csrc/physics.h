// physics.h -- device-side physics of the per-pixel path for gfx950.
//
// One ray per lane: every function here is a register-resident FP64 (or
// FP32 for the precision form's transfer) routine with no memory traffic
// except the sigma table, which the kernels stage in LDS.
//
// Reference correspondence (file:line of the reference sources):
//   Kerr metric / inverse / lapse / omega   src/kerr_metric.cpp:10-144
//   rhs (finite differences, REF mode)      src/kerr_geodesic.cpp:47-135
//   RKF45 tableau and REF step control      src/kerr_geodesic.cpp:194-329
//   RK4 fallback                            src/kerr_geodesic.cpp:137-192
//   camera initial state, dl, redshift      src/kerr_camera.cpp:14-186
//   density morphologies                    src/torus_profile.cpp:93-272
//   source / spectral weights, emissivity   src/radiative_transfer.cpp:64-395
//   transfer per path point                 src/radiative_transfer.cpp:501-672
//   sigma log-log interpolation             src/sigma_table.cpp:80-124
#pragma once

#include <hip/hip_runtime.h>

#include "fastmath.h"

#include "../../include/hadros.h"

namespace hd {

constexpr double PI = 3.141592653589793238462643383279502884;
constexpr double M_U_G = 1.66053906660e-24;   // atomic mass unit [g]

// Products feeding a sum may fuse into FMAs (one rounding instead of two) in
// the functions that open with HD_CONTRACT: the transfer's terms, compared with
// the oracle to a tolerance.  The CFG integrator and the per-point geometry
// write their FMAs out instead (the compiler's choice of which products to
// fuse depends on the surrounding code, and every kernel that traces a ray
// must round it alike); the REF-mode integrator keeps the written operation
// order.
#define HD_CONTRACT _Pragma("clang fp contract(fast)")

__host__ __device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ double dclamp(double v, double lo, double hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// sincos_m and sincos_shift are in fastmath.h.

// ------------------------------------------------------------------ params
// Normalised TorusProfile (constructor clamps of src/torus_profile.cpp:12-42
// applied on the host).
struct DevMedium {
    int kind;
    int powerlaw_disk, funnel, envelope;
    double rho0, r0, sr, H, p, fd, ft, env0, ealpha, rmin, rmax, floor_;
    double inv_r0, inv_sr, inv_H, inv_rmax, inv_ft;   // reciprocals: the hot path multiplies
};

struct DevSource {
    int kind;
    int ring_ok;          // emissivity_collapsar_ring guard (src/radiative_transfer.cpp:343-349)
    double rc, sr, tw, norm, ft, q, s, rho_ref, cmin, cmax;
    double gdr, gdth;   // shock-layer steps, already clamped to >= 1e-4 and >= 1e-5
    double rho_norm_c;    // torus.rho(r_center, pi/2), a per-scene constant
    double inv_rc, inv_sr, inv_tw;
};

// Spectral factor of every bin: ln S(E) = -gam ln E - E * invcut, valid if spec_ok.
// Monochromatic: gam = source.powerlaw, invcut = 1/emax (src/radiative_transfer.cpp:184-192);
// powerlaw: gam, 0; powerlaw_cutoff: gam, 1/ecut.
struct DevRt {
    DevMedium med;
    DevSource src;
    int nb;               // bins integrated (0..16)
    int mono;
    int spec_ok;
    int n_sigma;
    double gam, invcut;
    double rg_cm;         // r_g in cm (include/constants.hpp:17-20)
    double r_h_rt;        // 1 + sqrt(1 - clamp(a,0,0.999999)^2) (src/radiative_transfer.cpp:485-489)
    double e_lo, e_hi;    // sigma table range
    double sig_lo;        // ln E of the first table node
    double sig_inv_d;     // (n - 1) / (ln E_last - ln E_first) when the nodes are uniform in ln E, else 0
    int sig_exact;        // the nodes lie within 1e-12 of the spacing of that uniform grid
    int sig_pow;          // every node lies on one power law sigma = sigma_0 (E / E_0)^sig_s
    double sig_s;         // its exponent
    double emid[HADROS_MAX_BINS], lnemid[HADROS_MAX_BINS], de[HADROS_MAX_BINS], w[HADROS_MAX_BINS];
    double sigb[HADROS_MAX_BINS];   // sigma at each bin's E_mid, for sig_pow tables
};

struct DevCam {
    double a, r_obs, th_obs, tanhalf, r_max, h, tol;
    int nx, ny, mode, max_steps;
    int pad_rt;             // scene index of this camera in a batch
    int sub_stride, sub_offset, sub_nx, sub_ny;   // traced pixels i = offset + stride p (batch only)
    // observer constants of KerrCamera::initial_state (src/kerr_camera.cpp:110-186)
    double g_tt, g_tph, g_rr, g_thth, g_phph, alpha, omega;
    double r_h;
};

// ------------------------------------------------------------------ metric
struct Metric5 {
    double tt, tph, rr, thth, phph;
};

// KerrMetric::inverse_metric (src/kerr_metric.cpp:77-116) given sin, cos of theta.
__device__ __forceinline__ Metric5 inv_metric(double a, double r, double s, double c)
{
    const double sig = r * r + a * a * c * c;
    const double del = r * r - 2.0 * r + a * a;
    const double s2 = s * s;
    const double bigA = (r * r + a * a) * (r * r + a * a) - a * a * del * s2;
    Metric5 g;
    g.tt = -bigA / (sig * del);
    g.tph = -2.0 * a * r / (sig * del);
    g.rr = del / sig;
    g.thth = 1.0 / sig;
    g.phph = (del - a * a * s2) / (sig * del * s2);
    return g;
}

// KerrMetric::metric (src/kerr_metric.cpp:34-75).
__device__ __forceinline__ Metric5 metric(double a, double r, double th)
{
    double s, c;
    sincos_m(th, &s, &c);
    const double sig = r * r + a * a * c * c;
    const double del = r * r - 2.0 * r + a * a;
    const double s2 = s * s;
    Metric5 g;
    g.tt = -(1.0 - 2.0 * r / sig);
    g.tph = -2.0 * a * r * s2 / sig;
    g.rr = sig / del;
    g.thth = sig;
    g.phph = (r * r + a * a + 2.0 * a * a * r * s2 / sig) * s2;
    return g;
}

// lapse and frame-drag together (src/kerr_metric.cpp:118-144)
__device__ __forceinline__ void lapse_omega(double a, double r, double th, double& alpha, double& omega)
{
    double s, c;
    sincos_m(th, &s, &c);
    const double sig = r * r + a * a * c * c;
    const double del = r * r - 2.0 * r + a * a;
    const double s2 = s * s;
    const double bigA = (r * r + a * a) * (r * r + a * a) - a * a * del * s2;
    alpha = sqrt(sig * del / bigA);
    omega = 2.0 * a * r / bigA;
}

// ------------------------------------------------------------------ RHS
// Evolving part of the state: r, theta, phi, p_r, p_theta.  p_t and p_phi have
// exactly zero derivative (src/kerr_geodesic.cpp:116-117) and t feeds nothing.
struct Y5 {
    double r, th, ph, pr, pth;
};

// KerrGeodesic::rhs in REF mode: central differences of g^{mu nu} with
// eps_r = 1e-5 max(1,|r|), eps_th = 1e-5; the accumulation runs over the
// non-zero (mu,nu) pairs in the reference's mu-major order.
__device__ __forceinline__ Y5 rhs_ref(double a, const Y5& y, double pt, double pph)
{
    const double r = y.r, th = y.th;
    double s, c;
    sincos_m(th, &s, &c);
    const Metric5 g0 = inv_metric(a, r, s, c);
    Y5 d;
    d.r = g0.rr * y.pr;
    d.th = g0.thth * y.pth;
    d.ph = g0.tph * pt + g0.phph * pph;

    const double er = 1.0e-5 * dmax(1.0, fabs(r));
    const double et = 1.0e-5;
    const Metric5 gp = inv_metric(a, r + er, s, c);
    const Metric5 gm = inv_metric(a, r - er, s, c);
    double sp, cp, sm, cm;
    sincos_m(th + et, &sp, &cp);
    sincos_m(th - et, &sm, &cm);
    const Metric5 tp = inv_metric(a, r, sp, cp);
    const Metric5 tm = inv_metric(a, r, sm, cm);
    const double two_er = 2.0 * er, two_et = 2.0 * et;
    const double dr_tt = (gp.tt - gm.tt) / two_er, dt_tt = (tp.tt - tm.tt) / two_et;
    const double dr_tp = (gp.tph - gm.tph) / two_er, dt_tp = (tp.tph - tm.tph) / two_et;
    const double dr_rr = (gp.rr - gm.rr) / two_er, dt_rr = (tp.rr - tm.rr) / two_et;
    const double dr_hh = (gp.thth - gm.thth) / two_er, dt_hh = (tp.thth - tm.thth) / two_et;
    const double dr_pp = (gp.phph - gm.phph) / two_er, dt_pp = (tp.phph - tm.phph) / two_et;
    double a_r = 0.0, a_t = 0.0;
    a_r -= 0.5 * dr_tt * pt * pt;      a_t -= 0.5 * dt_tt * pt * pt;
    a_r -= 0.5 * dr_tp * pt * pph;     a_t -= 0.5 * dt_tp * pt * pph;
    a_r -= 0.5 * dr_rr * y.pr * y.pr;  a_t -= 0.5 * dt_rr * y.pr * y.pr;
    a_r -= 0.5 * dr_hh * y.pth * y.pth; a_t -= 0.5 * dt_hh * y.pth * y.pth;
    a_r -= 0.5 * dr_tp * pph * pt;     a_t -= 0.5 * dt_tp * pph * pt;
    a_r -= 0.5 * dr_pp * pph * pph;    a_t -= 0.5 * dt_pp * pph * pph;
    d.pr = a_r;
    d.pth = a_t;
    return d;
}

// Analytic Hamiltonian derivative (CFG mode).  With W = (r^2+a^2) p_t + a p_phi
// and V = p_phi/sin + a sin p_t, 2 Sigma H = N = Delta p_r^2 + p_th^2 - W^2/Delta + V^2
// (the contraction g^{mu nu} p_mu p_nu of src/kerr_metric.cpp:77-116 times Sigma).
// Takes sin and cos of y.th (y.th itself is not read); rhs_cfg computes them.
// The fused multiply-adds are written out (no contraction pragma), so every
// kernel that inlines this rounds alike whatever code surrounds it.
__device__ __forceinline__ Y5 rhs_cfg_sc(double a, const Y5& y, double s, double c, double pt, double pph)
{
    const double r = y.r, a2 = a * a, apt = a * pt;
    const double r2 = r * r;
    const double sig = fma(a2 * c, c, r2);
    const double ra = r2 + a2;
    const double del = fma(-2.0, r, ra);
    const double W = fma(ra, pt, a * pph);
    // one reciprocal for 1/s, 1/Sigma, 1/Delta
    const double sd = sig * del;
    const double inv_all = fast_rcp(s * sd);
    const double inv_s = sd * inv_all;
    const double s_ia = s * inv_all;          // 1 / (Sigma Delta)
    const double inv_sig = s_ia * del;
    const double pinv = pph * inv_s;          // p_phi / sin
    const double V = fma(apt, s, pinv);
    const double Wd = W * (sig * s_ia);       // W / Delta
    const double pr2 = y.pr * y.pr;
    const double dpr = del * y.pr;
    const double N = fma(V, V, fma(-W, Wd, fma(dpr, y.pr, y.pth * y.pth)));
    // dN/dr = Delta' (p_r^2 + (W/Delta)^2) - 4 r p_t W/Delta, Delta' = 2 r - 2
    const double ddel = fma(2.0, r, -2.0);
    const double dNr = fma(-4.0 * pt * r, Wd, ddel * fma(Wd, Wd, pr2));
    // dN/dtheta / 2 = V dV/dtheta, dV/dtheta = cos (a p_t - p_phi / sin^2)
    const double VVth = V * (c * fma(-pinv, inv_s, apt));
    const double h_is = 0.5 * inv_sig;
    const double n_over = (N * h_is) * inv_sig;   // N / (2 Sigma^2)
    Y5 d;
    d.r = dpr * inv_sig;
    d.th = y.pth * inv_sig;
    d.ph = fma(V, inv_s, -a * Wd) * inv_sig;
    // -dH/dr = -(dN/dr) / (2 Sigma) + N Sigma_r / (2 Sigma^2), Sigma_r = 2 r
    d.pr = fma(n_over, 2.0 * r, -dNr * h_is);
    // -dH/dtheta, Sigma_theta = -2 a^2 cos sin
    d.pth = -fma(n_over, (2.0 * a2) * (c * s), VVth * inv_sig);
    return d;
}

__device__ __forceinline__ Y5 rhs_cfg(double a, const Y5& y, double pt, double pph)
{
    double s, c;
    sincos_m(y.th, &s, &c);
    return rhs_cfg_sc(a, y, s, c, pt, pph);
}

template <int MODE>
__device__ __forceinline__ Y5 rhs(double a, const Y5& y, double pt, double pph)
{
    if constexpr (MODE == HADROS_MODE_CFG) return rhs_cfg(a, y, pt, pph);
    else return rhs_ref(a, y, pt, pph);
}

__device__ __forceinline__ Y5 axpy1(const Y5& y, double a1, const Y5& k1)
{
    return {y.r + a1 * k1.r, y.th + a1 * k1.th, y.ph + a1 * k1.ph, y.pr + a1 * k1.pr, y.pth + a1 * k1.pth};
}

// y + a1 k1 by fused multiply-adds (the CFG attempt forms)
__device__ __forceinline__ Y5 axpy1_f(const Y5& y, double a1, const Y5& k1)
{
    return {fma(a1, k1.r, y.r), fma(a1, k1.th, y.th), fma(a1, k1.ph, y.ph), fma(a1, k1.pr, y.pr),
            fma(a1, k1.pth, y.pth)};
}

// Fehlberg 4(5) attempt (tableau of src/kerr_geodesic.cpp:239-308) in CFG
// mode: the stage sums of rkf45 below as explicit fused multiply-adds in the
// reference's left-to-right order, and the tableau's fractions as rounded
// constants: h times a constant instead of the reference's
// (h x numerator) / denominator, a run-time division each.  Written out, not
// contracted by the compiler, so that every kernel rounds an attempt alike.
#define HD_ST3(F) t.F = fma(a32, k2.F, fma(a31, k1.F, y.F))
#define HD_ST4(F) t.F = fma(a43, k3.F, fma(a42, k2.F, fma(a41, k1.F, y.F)))
#define HD_ST5(F) t.F = fma(a54, k4.F, fma(a53, k3.F, fma(a52, k2.F, fma(a51, k1.F, y.F))))
#define HD_ST6(F) t.F = fma(a65, k5.F, fma(a64, k4.F, fma(a63, k3.F, fma(a62, k2.F, fma(a61, k1.F, y.F)))))
#define HD_Y4(F) \
    y4.F = fma(h, fma(-1.0 / 5.0, k5.F, fma(2197.0 / 4104.0, k4.F, fma(1408.0 / 2565.0, k3.F, (25.0 / 216.0) * k1.F))), y.F)
#define HD_Y5(F)                                                                                                 \
    y5.F = fma(h,                                                                                                \
               fma(2.0 / 55.0, k6.F,                                                                             \
                   fma(-9.0 / 50.0, k5.F,                                                                        \
                       fma(28561.0 / 56430.0, k4.F, fma(6656.0 / 12825.0, k3.F, (16.0 / 135.0) * k1.F)))), \
               y.F)
__device__ __forceinline__ void rkf45_cfg(double a, const Y5& y, double pt, double pph, double h, Y5& y4, Y5& y5)
{
    const Y5 k1 = rhs_cfg(a, y, pt, pph);
    const double a21 = h * 0.25;
    const Y5 k2 = rhs_cfg(a, axpy1_f(y, a21, k1), pt, pph);
    const double a31 = h * (3.0 / 32.0), a32 = h * (9.0 / 32.0);
    Y5 t;
    HD_ST3(r); HD_ST3(th); HD_ST3(ph); HD_ST3(pr); HD_ST3(pth);
    const Y5 k3 = rhs_cfg(a, t, pt, pph);
    const double a41 = h * (1932.0 / 2197.0), a42 = h * (-7200.0 / 2197.0), a43 = h * (7296.0 / 2197.0);
    HD_ST4(r); HD_ST4(th); HD_ST4(ph); HD_ST4(pr); HD_ST4(pth);
    const Y5 k4 = rhs_cfg(a, t, pt, pph);
    const double a51 = h * (439.0 / 216.0), a52 = h * -8.0, a53 = h * (3680.0 / 513.0), a54 = h * (-845.0 / 4104.0);
    HD_ST5(r); HD_ST5(th); HD_ST5(ph); HD_ST5(pr); HD_ST5(pth);
    const Y5 k5 = rhs_cfg(a, t, pt, pph);
    const double a61 = h * (-8.0 / 27.0), a62 = h * 2.0, a63 = h * (-3544.0 / 2565.0), a64 = h * (1859.0 / 4104.0),
                 a65 = h * (-11.0 / 40.0);
    HD_ST6(r); HD_ST6(th); HD_ST6(ph); HD_ST6(pr); HD_ST6(pth);
    const Y5 k6 = rhs_cfg(a, t, pt, pph);
    HD_Y4(r); HD_Y4(th); HD_Y4(ph); HD_Y4(pr); HD_Y4(pth);
    HD_Y5(r); HD_Y5(th); HD_Y5(ph); HD_Y5(pr); HD_Y5(pth);
}

// The same attempt with the five later stages' sin/cos by angle addition
// from the attempt's (sincos_shift, theta_j = y.th + d_j with d_j the stage's
// theta offset) and phi's 4th- and 5th-order sums accumulated as the stages
// come (phi feeds no stage).  Returns false when a stage offset leaves
// sincos_shift's range: the attempt is then discarded and redone by
// rkf45_cfg (rkf45_try).
// s0, c0: sin and cos of y.th (sincos_m), which a caller that already has
// them passes in.
__device__ __forceinline__ bool rkf45_cfg_shift_sc(double a, const Y5& y, double s0, double c0, double pt, double pph,
                                                   double h, Y5& y4, Y5& y5)
{
    double sj, cj;
    bool ok = true;
    const Y5 k1 = rhs_cfg_sc(a, y, s0, c0, pt, pph);
    double ph4 = (25.0 / 216.0) * k1.ph, ph5 = (16.0 / 135.0) * k1.ph;
    const double a21 = h * 0.25;
    ok &= sincos_shift(s0, c0, a21 * k1.th, &sj, &cj);
    const Y5 k2 = rhs_cfg_sc(a, axpy1_f(y, a21, k1), sj, cj, pt, pph);
    const double a31 = h * (3.0 / 32.0), a32 = h * (9.0 / 32.0);
    Y5 t;
    HD_ST3(r); HD_ST3(pr); HD_ST3(pth);
    ok &= sincos_shift(s0, c0, fma(a32, k2.th, a31 * k1.th), &sj, &cj);
    const Y5 k3 = rhs_cfg_sc(a, t, sj, cj, pt, pph);
    ph4 = fma(1408.0 / 2565.0, k3.ph, ph4);
    ph5 = fma(6656.0 / 12825.0, k3.ph, ph5);
    const double a41 = h * (1932.0 / 2197.0), a42 = h * (-7200.0 / 2197.0), a43 = h * (7296.0 / 2197.0);
    HD_ST4(r); HD_ST4(pr); HD_ST4(pth);
    ok &= sincos_shift(s0, c0, fma(a43, k3.th, fma(a42, k2.th, a41 * k1.th)), &sj, &cj);
    const Y5 k4 = rhs_cfg_sc(a, t, sj, cj, pt, pph);
    ph4 = fma(2197.0 / 4104.0, k4.ph, ph4);
    ph5 = fma(28561.0 / 56430.0, k4.ph, ph5);
    const double a51 = h * (439.0 / 216.0), a52 = h * -8.0, a53 = h * (3680.0 / 513.0), a54 = h * (-845.0 / 4104.0);
    HD_ST5(r); HD_ST5(pr); HD_ST5(pth);
    ok &= sincos_shift(s0, c0, fma(a54, k4.th, fma(a53, k3.th, fma(a52, k2.th, a51 * k1.th))), &sj, &cj);
    const Y5 k5 = rhs_cfg_sc(a, t, sj, cj, pt, pph);
    ph4 = fma(-1.0 / 5.0, k5.ph, ph4);
    ph5 = fma(-9.0 / 50.0, k5.ph, ph5);
    const double a61 = h * (-8.0 / 27.0), a62 = h * 2.0, a63 = h * (-3544.0 / 2565.0), a64 = h * (1859.0 / 4104.0),
                 a65 = h * (-11.0 / 40.0);
    HD_ST6(r); HD_ST6(pr); HD_ST6(pth);
    ok &= sincos_shift(s0, c0, fma(a65, k5.th, fma(a64, k4.th, fma(a63, k3.th, fma(a62, k2.th, a61 * k1.th)))),
                       &sj, &cj);
    const Y5 k6 = rhs_cfg_sc(a, t, sj, cj, pt, pph);
    ph5 = fma(2.0 / 55.0, k6.ph, ph5);
    HD_Y4(r); HD_Y4(th); HD_Y4(pr); HD_Y4(pth);
    HD_Y5(r); HD_Y5(th); HD_Y5(pr); HD_Y5(pth);
    y4.ph = fma(h, ph4, y.ph);
    y5.ph = fma(h, ph5, y.ph);
    return ok;
}
#undef HD_ST3
#undef HD_ST4
#undef HD_ST5
#undef HD_ST6
#undef HD_Y4
#undef HD_Y5

__device__ __forceinline__ bool rkf45_cfg_shift(double a, const Y5& y, double pt, double pph, double h, Y5& y4,
                                                Y5& y5)
{
    double s0, c0;
    sincos_m(y.th, &s0, &c0);
    return rkf45_cfg_shift_sc(a, y, s0, c0, pt, pph, h, y4, y5);
}

// One RKF45 attempt of a lane.  CFG mode takes the angle-addition form unless
// `direct` is set; when a stage offset leaves its range it sets `direct` and
// returns false (no attempt made): the caller's next pass redoes the attempt
// with rkf45_cfg, which clears `direct`.  Which form a ray's attempt uses thus
// depends on that ray alone, so every kernel traces a ray to the same bits.
template <int MODE>
__device__ __forceinline__ bool rkf45_try(double a, const Y5& y, double pt, double pph, double h, bool& direct,
                                          Y5& y4, Y5& y5);

template <int MODE>
__device__ __forceinline__ void rkf45(double a, const Y5& y, double pt, double pph, double h, Y5& y4, Y5& y5)
{
    if constexpr (MODE == HADROS_MODE_CFG) {
        rkf45_cfg(a, y, pt, pph, h, y4, y5);
        return;
    }
    const Y5 k1 = rhs<MODE>(a, y, pt, pph);
    const double a21 = h * 1.0 / 4.0;
    const Y5 k2 = rhs<MODE>(a, axpy1(y, a21, k1), pt, pph);
    const double a31 = h * 3.0 / 32.0, a32 = h * 9.0 / 32.0;
    Y5 t;
#define HD_ST3(F) t.F = y.F + a31 * k1.F + a32 * k2.F
    HD_ST3(r); HD_ST3(th); HD_ST3(ph); HD_ST3(pr); HD_ST3(pth);
    const Y5 k3 = rhs<MODE>(a, t, pt, pph);
    const double a41 = h * 1932.0 / 2197.0, a42 = h * -7200.0 / 2197.0, a43 = h * 7296.0 / 2197.0;
#define HD_ST4(F) t.F = y.F + a41 * k1.F + a42 * k2.F + a43 * k3.F
    HD_ST4(r); HD_ST4(th); HD_ST4(ph); HD_ST4(pr); HD_ST4(pth);
    const Y5 k4 = rhs<MODE>(a, t, pt, pph);
    const double a51 = h * 439.0 / 216.0, a52 = h * -8.0, a53 = h * 3680.0 / 513.0, a54 = h * -845.0 / 4104.0;
#define HD_ST5(F) t.F = y.F + a51 * k1.F + a52 * k2.F + a53 * k3.F + a54 * k4.F
    HD_ST5(r); HD_ST5(th); HD_ST5(ph); HD_ST5(pr); HD_ST5(pth);
    const Y5 k5 = rhs<MODE>(a, t, pt, pph);
    const double a61 = h * -8.0 / 27.0, a62 = h * 2.0, a63 = h * -3544.0 / 2565.0, a64 = h * 1859.0 / 4104.0,
                 a65 = h * -11.0 / 40.0;
#define HD_ST6(F) t.F = y.F + a61 * k1.F + a62 * k2.F + a63 * k3.F + a64 * k4.F + a65 * k5.F
    HD_ST6(r); HD_ST6(th); HD_ST6(ph); HD_ST6(pr); HD_ST6(pth);
    const Y5 k6 = rhs<MODE>(a, t, pt, pph);
#define HD_Y4(F) y4.F = y.F + h * (25.0 / 216.0 * k1.F + 1408.0 / 2565.0 * k3.F + 2197.0 / 4104.0 * k4.F - 1.0 / 5.0 * k5.F)
#define HD_Y5(F) y5.F = y.F + h * (16.0 / 135.0 * k1.F + 6656.0 / 12825.0 * k3.F + 28561.0 / 56430.0 * k4.F - 9.0 / 50.0 * k5.F + 2.0 / 55.0 * k6.F)
    HD_Y4(r); HD_Y4(th); HD_Y4(ph); HD_Y4(pr); HD_Y4(pth);
    HD_Y5(r); HD_Y5(th); HD_Y5(ph); HD_Y5(pr); HD_Y5(pth);
#undef HD_ST3
#undef HD_ST4
#undef HD_ST5
#undef HD_ST6
#undef HD_Y4
#undef HD_Y5
}

template <int MODE>
__device__ __forceinline__ bool rkf45_try(double a, const Y5& y, double pt, double pph, double h, bool& direct,
                                          Y5& y4, Y5& y5)
{
    if constexpr (MODE == HADROS_MODE_CFG) {
        if (direct) {
            rkf45_cfg(a, y, pt, pph, h, y4, y5);
            direct = false;
            return true;
        }
        if (rkf45_cfg_shift(a, y, pt, pph, h, y4, y5)) return true;
        direct = true;
        return false;
    } else {
        rkf45<MODE>(a, y, pt, pph, h, y4, y5);
        return true;
    }
}

// The same with sin and cos of y.th passed in (sincos_m's, so the results are
// the same bits): a kernel that keeps them from the last accepted point.
template <int MODE>
__device__ __forceinline__ bool rkf45_try_sc(double a, const Y5& y, double s0, double c0, double pt, double pph,
                                             double h, bool& direct, Y5& y4, Y5& y5)
{
    if constexpr (MODE == HADROS_MODE_CFG) {
        if (direct) {
            rkf45_cfg(a, y, pt, pph, h, y4, y5);
            direct = false;
            return true;
        }
        if (rkf45_cfg_shift_sc(a, y, s0, c0, pt, pph, h, y4, y5)) return true;
        direct = true;
        return false;
    } else {
        rkf45<MODE>(a, y, pt, pph, h, y4, y5);
        return true;
    }
}

// KerrGeodesic::step_rk4 at the base step (src/kerr_geodesic.cpp:137-192).
__device__ __noinline__ Y5 rk4_ref(double a, const Y5& y, double pt, double pph, double h)
{
    const Y5 k1 = rhs_ref(a, y, pt, pph);
    const Y5 k2 = rhs_ref(a, axpy1(y, 0.5 * h, k1), pt, pph);
    const Y5 k3 = rhs_ref(a, axpy1(y, 0.5 * h, k2), pt, pph);
    const Y5 k4 = rhs_ref(a, axpy1(y, h, k3), pt, pph);
    Y5 o;
#define HD_RK4(F) o.F = y.F + h * (k1.F + 2.0 * k2.F + 2.0 * k3.F + k4.F) / 6.0
    HD_RK4(r); HD_RK4(th); HD_RK4(ph); HD_RK4(pr); HD_RK4(pth);
#undef HD_RK4
    return o;
}

// error_norm (src/kerr_geodesic.cpp:219-233); std::max semantics keep err on NaN.
__device__ __forceinline__ double err_ref(const Y5& a, const Y5& b)
{
    double e = 0.0;
    e = dmax(e, fabs(a.r - b.r));
    e = dmax(e, fabs(a.th - b.th));
    e = dmax(e, fabs(a.ph - b.ph));
    e = dmax(e, fabs(a.pr - b.pr));
    e = dmax(e, fabs(a.pth - b.pth));
    return e;
}

// CFG mixed error: max |y5-y4| / (tol (1 + max(|y|,|y5|))); NaN forces rejection.
// The largest ratio is selected by cross-multiplication (scales are positive),
// so only the winning ratio is divided out.
__device__ __forceinline__ double err_cfg(const Y5& y, const Y5& y4, const Y5& y5, double tol)
{
    double nb = 0.0, db = 1.0;
    bool bad = false;
    // bitwise ors and selects: no control flow (a short-circuit || here made
    // the compiler branch per component)
#define HD_E(F)                                                                          \
    {                                                                                    \
        const double sc = tol * (1.0 + fmax(fabs(y.F), fabs(y5.F)));                    \
        const double d = fabs(y5.F - y4.F);                                              \
        bad = bad | (d != d) | (sc != sc);                                               \
        const bool gt = d * db > nb * sc;                                                \
        nb = gt ? d : nb;                                                                \
        db = gt ? sc : db;                                                               \
    }
    HD_E(r); HD_E(th); HD_E(ph); HD_E(pr); HD_E(pth);
#undef HD_E
    return bad ? __builtin_nan("") : nb * fast_rcp(db);   // db >= tol > 0
}

constexpr double CFG_H_MIN = 1.0e-9;

// err^(-1/5) for the CFG step controller.  The controller clamps 0.9 err^-0.2
// to [0.1, 5], which only depends on err inside [1e-5, 1e6]: clamp there,
// seed in FP32 and refine with two division-free Newton steps for y^-5 = x,
// y <- y (6 - x y^5) / 5 (quadratic: 1e-6 -> 1e-11 -> rounding).  About
// 25 instructions instead of an FP64 log and exp.
__device__ __forceinline__ double inv_fifth_root(double x)
{
    const double xc = dclamp(x, 1.0e-5, 1.0e6);
    // the hardware log2 and exp2 (no denormal scaling: xc and the seed are
    // normal floats, where the library forms compute the same bits)
    double y = double(__builtin_amdgcn_exp2f(-0.2f * __builtin_amdgcn_logf(float(xc))));
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double y2 = y * y;
        y = y * (6.0 - xc * (y2 * y2 * y)) * 0.2;
    }
    return y;
}

// The CFG controller's step factor: 0.9 err^-1/5 clamped to [0.2, 5] on
// accept (5 for err = 0) and to [0.1, 0.9] on reject (0.1 for a NaN error).
// One err^-1/5 serves both outcomes: a wavefront's lanes accept and reject
// in the same pass, and a branch per outcome evaluated it twice.
__device__ __forceinline__ double cfg_step_factor(double err, bool acc)
{
    const double f = 0.9 * inv_fifth_root(err);
    return acc ? fmin(5.0, fmax(0.2, (err > 0.0) ? f : 5.0)) : fmin(0.9, fmax(0.1, (err == err) ? f : 0.1));
}

// ------------------------------------------------------------------ camera
// The per-point geometry below fixes its fused multiply-adds by hand (no
// HD_CONTRACT): the compiler's choice of which products to fuse depends on the
// surrounding code, and every kernel that traces a ray must round its points
// alike (the batch and plain tracers are bitwise equal).

// Sigma = r^2 + a^2 cos^2, Delta = r^2 - 2 r + a^2
__device__ __forceinline__ void sigma_delta(double a2, double r, double c, double& sig, double& del)
{
    sig = fma(a2 * c, c, r * r);
    del = fma(r, r - 2.0, a2);
}

// src/kerr_camera.cpp:14-29 with the lapse and frame drag of
// src/kerr_metric.cpp:118-144: E = -(p_t + omega p_phi) / alpha, alpha^2 =
// Sigma Delta / A, omega = 2 a r / A, by one reciprocal and one reciprocal
// square root
__device__ __forceinline__ double zamo_energy_sc(double a, double r, double s, double c, double pt, double pph)
{
    const double a2 = a * a;
    double sig, del;
    sigma_delta(a2, r, c, sig, del);
    const double ra = fma(r, r, a2);
    const double inv_A = fast_rcp(fma(ra, ra, -(a2 * del) * (s * s)));
    // 2 (a r), not (2 a) r: the same bits (doubling is exact), and no loop
    // invariant 2a for a caller's loop to hold in registers
    const double omega = 2.0 * (a * r) * inv_A;
    return -fma(omega, pph, pt) * fast_rsqrt(sig * del * inv_A);
}

// with s, c = sincos_m(th)
__device__ __forceinline__ double zamo_energy(double a, double r, double th, double pt, double pph)
{
    double s, c;
    sincos_m(th, &s, &c);
    return zamo_energy_sc(a, r, s, c, pt, pph);
}

// src/kerr_camera.cpp:31-44 (the reference's while loops, bounded)
__device__ __forceinline__ double wrapped_dphi(double pn, double po)
{
    double d = pn - po;
    for (int it = 0; it < 64 && d > PI; ++it) d -= 2.0 * PI;
    for (int it = 0; it < 64 && d < -PI; ++it) d += 2.0 * PI;
    return d;
}

// src/kerr_camera.cpp:46-76 with the spatial metric of src/kerr_metric.cpp:
// 34-75 at the midpoint: g_rr = Sigma / Delta, g_thth = Sigma, g_phph =
// (r^2 + a^2 + 2 a^2 r s^2 / Sigma) s^2, by one reciprocal of Sigma Delta
// s, c: sin and cos of the midpoint's theta.  dl = dl^2 x rsqrt(dl^2), the
// refined hardware estimate (within an ulp); zero, infinite and NaN dl^2 take
// the library square root.
__device__ __forceinline__ double zamo_dl_mid(double a, const Y5& cur, const Y5& prev, double s, double c)
{
    const double r = 0.5 * (cur.r + prev.r);
    const double a2 = a * a;
    double sig, del;
    sigma_delta(a2, r, c, sig, del);
    const double s2 = s * s;
    const double inv_sd = fast_rcp(sig * del);
    const double g_rr = sig * sig * inv_sd;
    const double g_pp = fma(2.0 * a2 * r * s2, del * inv_sd, fma(r, r, a2)) * s2;
    const double dr = cur.r - prev.r;
    const double dth = cur.th - prev.th;
    const double dph = wrapped_dphi(cur.ph, prev.ph);
    const double dl2 = fma(g_rr * dr, dr, fma(sig * dth, dth, g_pp * dph * dph));
    if (dl2 > 0.0 && dl2 < INFINITY) return dl2 * fast_rsqrt(dl2);
    return sqrt(dmax(dl2, 0.0));
}

__device__ __forceinline__ double zamo_dl(double a, const Y5& cur, const Y5& prev)
{
    double s, c;
    sincos_m(0.5 * (cur.th + prev.th), &s, &c);
    return zamo_dl_mid(a, cur, prev, s, c);
}

// The same with sc, cc = sin, cos of cur.th at hand (sincos_m's): the
// midpoint's by angle addition over half the step back, or by sincos_m where
// that half step exceeds sincos_shift's range.  The split tracers (plain and
// batch) use this form, so they round their points alike.
__device__ __forceinline__ double zamo_dl_sc(double a, const Y5& cur, const Y5& prev, double sc, double cc)
{
    double s, c;
    if (!sincos_shift(sc, cc, 0.5 * (prev.th - cur.th), &s, &c)) sincos_m(0.5 * (cur.th + prev.th), &s, &c);
    return zamo_dl_mid(a, cur, prev, s, c);
}

// ------------------------------------------------------------------ medium
__host__ __device__ __forceinline__ double sq(double x) { return x * x; }

// cos of a polar angle (fastmath.h:cos_polar; cheaper than the general cos)
__host__ __device__ __forceinline__ double cos_theta(double th) { return cos_polar(th); }

// The exponential and logarithm of the density and source functions: the
// math library's (the default), or exp_tab and log_tab over the tables a
// kernel staged in LDS (the transfer kernels, where these run once per path
// point).  The exponentials' arguments are at most moderate positive values,
// inside exp_tab's domain.
struct ExpLib {
    __host__ __device__ double operator()(double x) const { return exp(x); }
    __host__ __device__ double ln(double x) const { return log(x); }
};
struct ExpTab {
    const double* T;   // EXP_TAB
    const double* L;   // LOG_TAB
    __device__ double operator()(double x) const { return exp_tab(x, T); }
    __device__ double ln(double x) const { return log_tab(x, L); }
};
// Float-accurate forms (fastmath.h:exp_f32, ln_f32) for the FP32 transfer
// form, whose per-point terms are rounded to float before they are used.
struct ExpF32 {
    __device__ double operator()(double x) const { return exp_f32(x); }
    __device__ double ln(double x) const { return ln_f32(x); }
};

// x^y for x > 0 as exp(y ln x): relative error about |y ln x| ulp, 1e-15 for
// the density radii (r / r0 < 1e3, |y| < 10), and a fraction of pow's cost
template <class Exp = ExpLib>
__host__ __device__ __forceinline__ double pow_pos(double x, double y, Exp ex = Exp{})
{
    return ex(y * ex.ln(x));
}

// TorusProfile::rho (src/torus_profile.cpp:150-272); pow(x, 2.0) is x*x.
// Products of exponentials are one exponential of the summed exponent, the
// power laws share ln(r / r0), and divisions by profile parameters are
// products with host-side reciprocals (relative differences from the
// reference's form of order |exponent| x 1e-16).
template <class Exp = ExpLib>
__host__ __device__ __forceinline__ double rho(const DevMedium& m, double r, double th, Exp ex = Exp{})
{
    HD_CONTRACT
    double shape;
    const double lr = (m.powerlaw_disk || m.envelope) ? ex.ln(r * m.inv_r0) : 0.0;   // ln(r / r0)
    if (m.kind == HADROS_PROFILE_COLLAPSAR_NDAF_LIKE) {
        if (!(r >= m.rmin && r <= m.rmax) || r <= 0.0 || m.r0 <= 0.0) {
            shape = 0.0;
        } else {
            const double mu = cos_theta(th);
            const double vertical = ex(-sq(mu / dmax(m.H, 1.0e-6)));
            const double inner_taper = 1.0 / (1.0 + ex(-(r - m.rmin) / dmax(0.35 * m.sr, 1.0e-6)));
            const double outer_taper = 1.0 / (1.0 + ex((r - m.rmax) / dmax(0.60 * m.sr, 1.0e-6)));
            const double inner_core = 0.45 * ex(-sq((r - 0.75 * m.r0) / dmax(0.75 * m.sr, 1.0e-6)));
            const double radial = pow_pos(dmax(r / m.r0, 1.0e-300), -dmax(m.p, 0.0), ex);
            const double floor_tail = 0.08 * ex(-sq((r - m.r0) / dmax(1.8 * m.sr, 1.0e-6)));
            shape = (radial + inner_core + floor_tail) * vertical * inner_taper * outer_taper;
        }
    } else if (m.powerlaw_disk) {
        if (!(r >= m.rmin && r <= m.rmax) || r <= 0.0 || m.r0 <= 0.0) {
            shape = 0.0;
        } else {
            // radial (r/r0)^-p x vertical ex(-(cos th / H)^2) x outer ex(-(r/rmax)^4)
            const double mu = cos_theta(th) * m.inv_H;
            const double q = r * m.inv_rmax;
            const double inner_taper = 1.0 - ex(-sq((r - m.rmin) * m.inv_sr));
            shape = ex(-m.p * lr - mu * mu - (q * q) * (q * q)) * inner_taper;
        }
    } else {
        const double delta = th - 0.5 * PI;
        if (r <= 4.0 || r >= 18.0 || fabs(delta) >= 0.45) {
            shape = 0.0;
        } else {
            shape = ex(-sq((r - m.r0) * m.inv_sr) - sq(delta * m.inv_H));
        }
    }
    double F = 1.0;
    if (m.funnel) {
        const double north = ex(-sq(th * m.inv_ft));
        const double south = ex(-sq((PI - th) * m.inv_ft));
        F = dclamp(1.0 - m.fd * (north + south), 0.0, 1.0);
    }
    double env = 0.0;
    if (m.envelope && !(m.env0 <= 0.0 || r < m.r0 || r > m.rmax)) env = m.env0 * ex(-m.ealpha * lr);
    return dmax(m.rho0 * shape * F + env, m.floor_);
}

// exp(base) x bipolar_gaussian(th, th0, w) (src/radiative_transfer.cpp:64-76)
// for w > 0, as two exponentials of summed exponents
template <class Exp>
__device__ __forceinline__ double bipolar_times(double th, double th0, double inv_w, double base, Exp ex)
{
    HD_CONTRACT
    const double n = (th - th0) * inv_w;
    const double s = (th - (PI - th0)) * inv_w;
    return ex(base - n * n) + ex(base - s * s);
}

// norm x spatial factor of emissivity_uhe (src/radiative_transfer.cpp:209-395):
// bin-independent, so evaluated once per path point.
// rho_here: the density at (r, th), which the caller has already evaluated
template <class Exp = ExpLib>
__device__ __forceinline__ double source_factor(const DevSource& so, const DevMedium& m, double r, double th,
                                                double rho_here, Exp ex = Exp{})
{
    HD_CONTRACT
    if (so.kind == HADROS_SOURCE_INNER_RING) {
        if (!so.ring_ok) return 0.0;
        const double dr = (r - so.rc) * so.inv_sr;
        const double dt = (th - 0.5 * PI) * so.inv_tw;
        return so.norm * ex(-dr * dr - dt * dt);
    }
    if (r <= 0.0 || so.sr <= 0.0 || so.tw <= 0.0) return 0.0;
    double W = 0.0;
    if (so.kind == HADROS_SOURCE_FUNNEL_WALL) {
        const double dr = (r - so.rc) * so.inv_sr;
        W = bipolar_times(th, so.ft, so.inv_tw, -dr * dr, ex);
    } else if (so.kind == HADROS_SOURCE_JET_BASE) {
        W = bipolar_times(th, 0.0, so.inv_tw, -sq(r * so.inv_rc), ex);
    } else if (so.kind == HADROS_SOURCE_DENSITY_WEIGHTED) {
        const double rh = rho_here;
        const double rho_norm = dmax(so.rho_ref > 0.0 ? so.rho_ref : so.rho_norm_c, 1.0e-300);
        const double dw = pow(dmax(rh / rho_norm, 0.0), so.q);
        const double rw = pow(dmax(r / dmax(so.rc, 1.0e-300), 1.0e-300), -so.s);
        W = dclamp(dw * rw, dmax(so.cmin, 0.0), dmax(so.cmax, dmax(so.cmin, 0.0)));
    } else if (so.kind == HADROS_SOURCE_SHOCK_LAYER) {
        const double dr = so.gdr;     // dmax(dr, 1e-4), clamped on the host
        const double dth = so.gdth;   // dmax(dtheta, 1e-5), clamped on the host
        // rho at (r +- dr, th) and (r, th +- dth): one density evaluation in a
        // rolled loop (four inlined copies would set the kernels' register use)
        double sr = 0.0, st = 0.0;
#pragma unroll 1
        for (int k = 0; k < 4; ++k) {
            const double rr = (k == 0) ? r + dr : (k == 1) ? dmax(r - dr, 1.0e-6) : r;
            const double tt = (k == 2) ? dclamp(th + dth, 0.0, PI) : (k == 3) ? dclamp(th - dth, 0.0, PI) : th;
            const double v = rho(m, rr, tt, ex);
            if (k == 0) sr = v;
            else if (k == 1) sr = sr - v;
            else if (k == 2) st = v;
            else st = st - v;
        }
        const double rho_norm = dmax(so.rho_norm_c, 1.0e-300);
        const double d_r = sr / (2.0 * dr);
        const double d_t = st / (2.0 * dth * dmax(r, 1.0e-6));
        const double grad = sqrt(d_r * d_r + d_t * d_t) / rho_norm;
        const double taper = ex(-sq(r / dmax(so.rc, 1.0e-300)));
        W = dclamp(grad * taper, dmax(so.cmin, 0.0), dmax(so.cmax, dmax(so.cmin, 0.0)));
    }
    return so.norm * W;
}

// ------------------------------------------------------------------ sigma
// LDS-staged table: ln E_k, ln sigma_k, slope_k = d ln sigma / d ln E on
// [k-1, k] (slope_0 unused).  lower_bound over ln E as std::lower_bound
// over E (src/sigma_table.cpp:100-124).
// lower_bound of lx in the ascending table lnE[0..n): the first k with
// lnE[k] >= lx (std::lower_bound of src/sigma_table.cpp:55-58,108-110).  A
// table uniform in ln E (inv_d > 0) starts from the arithmetic index guess
// and moves it at most one node to the exact bound (two LDS reads); other
// tables use a binary search.
template <typename T>
__device__ __forceinline__ int sigma_lower_bound(const T* __restrict__ lnE, int n, T lx, T lo0, T inv_d)
{
    if (inv_d > T(0)) {
        const T t = (lx - lo0) * inv_d;
        const int g = !(t > T(0)) ? 0 : (t >= T(n - 1) ? n - 1 : int(ceil(t)));
        // the host accepts the guess only for tables within d/4 of uniform
        // spacing: then the bound is g - 1, g or g + 1, decided by the two
        // nodes around the guess (read together, no loop)
        const T below = lnE[g > 0 ? g - 1 : 0];
        const T at = lnE[g];
        const bool down = g > 0 && below >= lx;
        const bool up = !down && at < lx;
        return g - int(down) + int(up);
    }
    int lo = 0, cnt = n;
    while (cnt > 0) {
        const int half = cnt >> 1;
        if (lnE[lo + half] < lx) {
            lo += half + 1;
            cnt -= half + 1;
        } else {
            cnt = half;
        }
    }
    return lo;
}

// ln sigma at ln E = lx for a table whose ln E nodes are uniform to within
// 1e-12 of the spacing d (sig_exact): the interval is ceil((lx - lo0) / d),
// clamped to [1, n-1], with no comparison against the nodes.  Near a node the
// rounding of that quotient may pick the neighbouring interval; both give the
// node's value there (the interpolation is continuous), so the result moves by
// rounding only.  Any lx, NaN included, gives an index inside the table.
__device__ __forceinline__ double sigma_ln_exact(const double* __restrict__ lnE, const double* __restrict__ lnS,
                                                 const double* __restrict__ slope, int n, double lx, double lo0,
                                                 double inv_d)
{
    HD_CONTRACT
    const double t = fmin(fmax((lx - lo0) * inv_d, 1.0), double(n - 1));   // NaN -> 1
    const int k = int(ceil(t));
    return lnS[k - 1] + (lx - lnE[k - 1]) * slope[k];
}

__device__ __forceinline__ float sigma_ln_exact_f(const float* __restrict__ lnE, const float* __restrict__ lnS,
                                                  const float* __restrict__ slope, int n, float lx, float lo0,
                                                  float inv_d)
{
    HD_CONTRACT
    const float t = fminf(fmaxf((lx - lo0) * inv_d, 1.0f), float(n - 1));   // NaN -> 1
    const int k = int(ceilf(t));
    return lnS[k - 1] + (lx - lnE[k - 1]) * slope[k];
}

// ln sigma at ln E = lx: log-log interpolation of the table (slope[k] is the
// ln sigma slope of interval [k-1, k]); clamped to the end nodes.
__device__ __forceinline__ double sigma_ln(const double* __restrict__ lnE, const double* __restrict__ lnS,
                                           const double* __restrict__ slope, int n, double lx, double lo0 = 0.0,
                                           double inv_d = 0.0)
{
    HD_CONTRACT
    const int lo = sigma_lower_bound(lnE, n, lx, lo0, inv_d);
    if (lo == 0) return lnS[0];
    if (lo >= n) return lnS[n - 1];
    return lnS[lo - 1] + (lx - lnE[lo - 1]) * slope[lo];
}

__device__ __forceinline__ float sigma_ln_f(const float* __restrict__ lnE, const float* __restrict__ lnS,
                                            const float* __restrict__ slope, int n, float lx, float lo0 = 0.0f,
                                            float inv_d = 0.0f)
{
    HD_CONTRACT
    const int lo = sigma_lower_bound(lnE, n, lx, lo0, inv_d);
    if (lo == 0) return lnS[0];
    if (lo >= n) return lnS[n - 1];
    return lnS[lo - 1] + (lx - lnE[lo - 1]) * slope[lo];
}

// ------------------------------------------------------------------ MeV channel
// MeVPhysicsParams (include/mev_neutrino_physics.hpp:42-80) with the enums as
// ints (HADROS_MEV_*), plus the TorusProfile constants its profiles read
// (rho0, r0, r_min, r_max after the constructor's clamps).
struct DevMev {
    int model, flavor, thermal, ye, spectral, n_bins;
    unsigned proc;   // HADROS_MEV_* process bits
    int pad;
    double enu, norm, sabs, sscat, sscat_e, ns_tau;
    double T0, T_floor, T_power, Ye_torus, Ye_funnel, Ye_envelope, Ye_floor, Ye_ceil, E_min, E_max;
    double rho0, r0, r_min, r_max;
};

constexpr double MEV_HBAR_C = 1.973269804e-11;   // MeV cm
constexpr double MEV_M_E = 0.510998950;          // MeV

// density_shape (src/mev_neutrino_physics.cpp:26-33)
__device__ __forceinline__ double mev_shape(const DevMev& M, double rho)
{
    return dclamp(rho / dmax(M.rho0, 1.0e-300), 0.0, 1.0);
}

// mev_temperature_profile_MeV (src/mev_neutrino_physics.cpp:235-286)
__device__ __forceinline__ double mev_temperature(const DevMev& M, double r, double th, double rho)
{
    HD_CONTRACT
    const double T0 = dmax(M.T0, 1.0e-12);
    const double floor_ = dmax(M.T_floor, 1.0e-12);
    const double shape = mev_shape(M, rho);
    const double eq = (th - 0.5 * PI) / 0.45;
    double T = T0;
    if (M.thermal == HADROS_MEV_T_INNER_HOT_TORUS) {
        T = floor_ + T0 * pow(shape, dmax(M.T_power, 0.0));
    } else if (M.thermal == HADROS_MEV_T_RADIAL_POWERLAW) {
        const double radial = pow(dmax(r / dmax(M.r0, 1.0e-300), 1.0e-300), -M.T_power);
        T = floor_ + T0 * radial * (0.35 + 0.65 * exp(-(eq * eq)));
    } else if (M.thermal == HADROS_MEV_T_TORUS_PLUS_COOL_ENVELOPE) {
        T = floor_ + T0 * pow(shape, dmax(M.T_power, 0.0));
        if (r > M.r_max || r < M.r_min || shape < 1.0e-3) {
            const double radial = pow(dmax(r / dmax(M.r0, 1.0e-300), 1.0e-300), -M.T_power);
            T = floor_ + 0.15 * T0 * radial;
        }
    } else if (M.thermal == HADROS_MEV_T_COLLAPSAR_INNER_HOT) {
        const double q = (r - M.r_min) / dmax(1.15 * M.r0, 1.0e-6);
        const double inner_hot = exp(-(q * q));
        const double radial_cooling = pow(1.0 + dmax(r - M.r_min, 0.0) / dmax(M.r0, 1.0e-6), -0.75);
        T = floor_ + T0 * (0.30 + 0.70 * pow(shape, 0.18)) * (0.45 + 0.55 * exp(-(eq * eq))) *
                         (0.55 + 0.45 * inner_hot) * radial_cooling;
    }
    return dmax(T, floor_);
}

// mev_ye_profile (src/mev_neutrino_physics.cpp:288-336)
__device__ __forceinline__ double mev_ye(const DevMev& M, double r, double th, double rho)
{
    HD_CONTRACT
    const double shape = mev_shape(M, rho);
    double Ye = M.Ye_torus;
    if (M.ye == HADROS_MEV_YE_NEUTRON_RICH_TORUS) {
        const double q = (r - M.r0) / 4.0;
        Ye = M.Ye_torus + 0.05 * exp(-(q * q));
    } else if (M.ye == HADROS_MEV_YE_FUNNEL_PROTON_RICH || M.ye == HADROS_MEV_YE_COLLAPSAR_NEUTRON_RICH) {
        const double a = th / 0.35, b = (PI - th) / 0.35;
        const double pw = dclamp(exp(-(a * a)) + exp(-(b * b)), 0.0, 1.0);
        if (M.ye == HADROS_MEV_YE_FUNNEL_PROTON_RICH) {
            Ye = M.Ye_torus * (1.0 - pw) + M.Ye_funnel * pw;
        } else {
            const double q = (r - M.r0) / dmax(1.2 * M.r0, 1.0e-6);
            const double torus_ye = M.Ye_torus - 0.05 * pow(shape, 0.25) * exp(-(q * q));
            Ye = torus_ye * (1.0 - pw) + M.Ye_funnel * pw;
        }
    } else if (M.ye == HADROS_MEV_YE_TORUS_ENVELOPE_CONTRAST) {
        const bool envelope_like = r > M.r0 && (r > M.r_max || shape < 1.0e-2);
        Ye = envelope_like ? M.Ye_envelope : M.Ye_torus;
        if (r < M.r_min && shape < 1.0e-2) Ye = M.Ye_funnel;
    }
    return dclamp(Ye, dmax(M.Ye_floor, 1.0e-6), dmin(M.Ye_ceil, 1.0 - 1.0e-6));
}

// mev_thermal_spectral_shape without the 1/T^3 (mev_fermi_dirac_weight,
// src/mev_neutrino_physics.cpp:338-372)
__device__ __forceinline__ double mev_fd(double E, double T)
{
    if (E <= 0.0 || T <= 0.0) return 0.0;
    const double x = E / T;
    if (x > 120.0) return 0.0;
    return E * E / (exp(x) + 1.0);
}

// mev_total_emissivity (src/mev_neutrino_physics.cpp:424-593, the urca / pair /
// brems terms and the legacy toy branch)
__device__ __forceinline__ double mev_emissivity(const DevMev& M, double rho, double T, double Ye, double E)
{
    HD_CONTRACT
    if (M.model == HADROS_MEV_MODEL_TOY) {
        if (rho <= 0.0 || T <= 0.0 || E <= 0.0) return 0.0;
        const double nb = rho / M_U_G;
        const double ccw = dclamp(1.0 - Ye, 0.0, 1.0) + 0.5 * dclamp(Ye, 0.0, 1.0);
        const double spec = E * E / (exp(dmin(E / T, 120.0)) + 1.0);
        const double capture_like = nb * ccw * pow(T, 6.0);
        const double pair_like = 0.05 * nb * pow(T, 9.0);
        return M.norm * (capture_like + pair_like) * spec;
    }
    if (T <= 0.0 || E <= 0.0) return 0.0;
    const double x = E / T;
    // thermal spectral shape E^2 / (e^x + 1) / T^3, shared by the three processes
    const double shp = (x > 120.0) ? 0.0 : E * E / (exp(x) + 1.0) / dmax(T * T * T, 1.0e-300);
    const double T10 = T / 10.0;
    const double rho10 = rho / 1.0e10;
    double total = 0.0;
    if ((M.proc & HADROS_MEV_URCA) && rho > 0.0 && M.flavor != HADROS_MEV_NU_X) {
        const double ye = dclamp(Ye, 0.0, 1.0);
        const double ff = (M.flavor == HADROS_MEV_NU_E) ? ye : (1.0 - ye);
        double urca = 1.0e30 * pow(dmax(rho10, 0.0), 1.0) * pow(dmax(T10, 0.0), 6.0) * ff * shp;
        if (M.proc & HADROS_MEV_DEGENERACY) {
            // mev_urca_degeneracy_correction (:415-441) with the zero-temperature
            // Fermi momentum proxy (:380-400)
            const double ne = rho * ye / M_U_G;
            const double pF = (ne <= 0.0) ? 0.0 : MEV_HBAR_C * cbrt(3.0 * PI * PI * ne);
            const double eta = sqrt(pF * pF + MEV_M_E * MEV_M_E) / T;
            double fac = 1.0;
            if (isfinite(eta)) {
                const double shift = 0.25 * tanh((eta - 1.0) / 4.0);
                fac = dclamp((M.flavor == HADROS_MEV_NU_E) ? (1.0 + shift) : (1.0 - shift), 0.5, 1.5);
            }
            urca *= fac;
        }
        total += urca;
    }
    if (M.proc & HADROS_MEV_PAIR)
        total += 3.0e28 * pow(dmax(T10, 0.0), 9.0) * ((M.flavor == HADROS_MEV_NU_X) ? 1.0 : 0.7) * shp;
    if ((M.proc & HADROS_MEV_BREMS) && rho > 0.0)
        total += 1.0e27 * pow(dmax(rho10, 0.0), 2.0) * pow(dmax(T10, 0.0), 5.0) *
                 ((M.flavor == HADROS_MEV_NU_X) ? 1.0 : 0.8) * shp;
    return M.norm * total;
}

// mev_total_opacity_cm_inv: absorption + scattering in cm^-1
// (src/mev_neutrino_physics.cpp:595-769)
__device__ __forceinline__ double mev_opacity(const DevMev& M, double rho, double Ye, double E)
{
    HD_CONTRACT
    if (rho <= 0.0 || E <= 0.0) return 0.0;
    const double nb = rho / M_U_G;
    if (M.model == HADROS_MEV_MODEL_TOY) {
        double k = 0.0;
        if (M.proc & HADROS_MEV_ABSORPTION)
            k = nb * M.sabs * E * E * (dclamp(1.0 - Ye, 0.0, 1.0) + 0.25 * dclamp(Ye, 0.0, 1.0));
        double sc = 0.0;
        if (M.proc & HADROS_MEV_SCATTERING) sc = nb * M.sscat * E * E * (1.0 + 0.5 * dclamp(Ye, 0.0, 1.0));
        return k + sc;
    }
    const double cy = dclamp(Ye, 0.0, 1.0);
    double abs_n = 0.0, abs_p = 0.0, sn = 0.0, sp = 0.0, se = 0.0;
    if (M.proc & HADROS_MEV_ABSORPTION) {
        if ((M.proc & HADROS_MEV_ABS_N) && M.flavor == HADROS_MEV_NU_E) abs_n = nb * (1.0 - cy) * M.sabs * E * E;
        if ((M.proc & HADROS_MEV_ABS_P) && M.flavor == HADROS_MEV_ANTI_NU_E) abs_p = nb * cy * M.sabs * E * E;
    }
    if (M.proc & HADROS_MEV_SCATTERING) {
        if (M.proc & HADROS_MEV_SCAT_N) sn = nb * (1.0 - cy) * M.sscat * E * E;
        if (M.proc & HADROS_MEV_SCAT_P) sp = nb * cy * 0.5 * M.sscat * E * E;
        if (M.proc & HADROS_MEV_SCAT_E) se = rho * cy / M_U_G * M.sscat_e * E * E;
    }
    return (abs_n + abs_p) + (sn + sp + se);
}

// kappa and j of one path point (src/radiative_transfer.cpp:541-600): the
// monochromatic channel at E_obs / g, or the Fermi-Dirac band average of kappa
// and band sum of j
__device__ __forceinline__ void mev_kappa_j(const DevMev& M, double g, double rho, double T, double Ye, double& kappa,
                                            double& j)
{
    HD_CONTRACT
    const double gc = dmax(g, 1.0e-300);
    if (M.spectral == HADROS_MEV_SPECTRAL_FERMI_DIRAC_BAND) {
        const int nbins = M.n_bins > 1 ? M.n_bins : 1;
        const double e_min = dmax(M.E_min, 1.0e-6);
        const double e_max = dmax(M.E_max, e_min);
        const double dE = (e_max - e_min) / double(nbins);
        double ws = 0.0, kw = 0.0, js = 0.0;
#pragma unroll 1
        for (int ib = 0; ib < nbins; ++ib) {
            const double E_local = (e_min + (double(ib) + 0.5) * dE) / gc;
            const double w = mev_fd(E_local, T) * dE;
            if (w <= 0.0 || !isfinite(w)) continue;
            const double lk = mev_opacity(M, rho, Ye, E_local);
            js += mev_emissivity(M, rho, T, Ye, E_local) * dE;
            kw += w * lk;
            ws += w;
        }
        kappa = ws > 0.0 ? kw / ws : 0.0;
        j = js;
    } else {
        const double E = M.enu / gc;
        kappa = mev_opacity(M, rho, Ye, E);
        j = mev_emissivity(M, rho, T, Ye, E);
    }
}

// ------------------------------------------------------------------ transfer
// Per-path-point quantities shared by all bins (integrate_kerr_ray,
// src/radiative_transfer.cpp:501-672): g = 1/z, g^3, rho, n_N dl, emissivity
// spatial factor.  Returns false for points the reference skips (r < r_+ + 1e-3).
struct PointQ {
    double z, lnz, g3Adl, nbar_dl, rho;
};

template <class Exp = ExpLib>
__device__ __forceinline__ bool point_q(const DevRt& P, double r, double th, double dl, double z, PointQ& q,
                                        Exp ex = Exp{})
{
    HD_CONTRACT
    if (r < P.r_h_rt + 1.0e-3) return false;
    const double zc = dmax(z, 1.0e-300);
    const double g = fast_rcp(zc);
    const double dl_cm = dl * P.rg_cm;
    q.rho = rho(P.med, r, th, ex);
    const double A = source_factor(P.src, P.med, r, th, q.rho, ex);
    q.z = zc;
    q.lnz = ex.ln(zc);
    q.g3Adl = g * g * g * A * dl_cm;
    q.nbar_dl = q.rho * (1.0 / M_U_G) * dl_cm;
    return true;
}

// e^x rounds to +0 in FP64 for every x below this (e^-750 < 2^-1081, under half
// the smallest subnormal 2^-1074): emission terms behind such a column are
// skipped, which leaves the FP64 intensity bit for bit unchanged
constexpr double EXP64_ZERO_BELOW = -750.0;
// log-domain FP32 emission terms below e^-1000 leave the FP64 intensity
// unchanged (zero on their own, below FP32 resolution beside any larger term)
constexpr float EXP32_LOG_NEGLIGIBLE = -1000.0f;

// The per-(point, bin) terms of the transfer kernels' rolled bin loops, for
// a point with PointQ q and bin energy E = E_mid z, lnEb = ln E_mid + ln z
// (FMA contraction allowed, as in the loops).
//
// Optical depth increment n_N sigma(E) dl where the point absorbs and E lies
// inside the table, else 0.  The lookup runs unconditionally (its index stays
// inside the table) and the mask selects, so the loop body has no branch.
// FP64: sigma by the table's exact-uniform or bracketing lookup, e^x by
// exp_tab with T = 2^(j/64).
__device__ __forceinline__ double bin_dtau64(bool exact, bool absorbs, double E, double lnEb, double nbar_dl,
                                             const double* lnE, const double* lnS, const double* slope, int n,
                                             double sig_lo, double sig_inv_d, double e_lo, double e_hi,
                                             const double* T)
{
    HD_CONTRACT
    const double ls = exact ? sigma_ln_exact(lnE, lnS, slope, n, lnEb, sig_lo, sig_inv_d)
                            : sigma_ln(lnE, lnS, slope, n, lnEb, sig_lo, sig_inv_d);
    const double sig_dl = nbar_dl * exp_tab_nz(fmax(ls, EXP_ZERO_BELOW), T);
    return (absorbs && E >= e_lo && E <= e_hi) ? sig_dl : 0.0;
}
// The same for a power-law table (DevRt::sig_pow): wdl = n_N dl z^s staged per
// point, sig_b = sigma(E_mid) of the bin, E = E_mid z for the table-range test
__device__ __forceinline__ double bin_dtau64_pow(bool absorbs, double E, double wdl, double sig_b, double e_lo,
                                                 double e_hi)
{
    return (absorbs && E >= e_lo && E <= e_hi) ? wdl * sig_b : 0.0;
}
// FP32 form: e^(ln(n_N dl) + ln sigma), lN = ln(n_N dl)
__device__ __forceinline__ float bin_dtau32(bool exact, bool absorbs, double E, float lnEb, float lN,
                                            const float* lnEf, const float* lnSf, const float* slopef, int n,
                                            float sig_lo, float sig_inv_d, double e_lo, double e_hi)
{
    HD_CONTRACT
    const float ls = exact ? sigma_ln_exact_f(lnEf, lnSf, slopef, n, lnEb, sig_lo, sig_inv_d)
                           : sigma_ln_f(lnEf, lnSf, slopef, n, lnEb, sig_lo, sig_inv_d);
    return (absorbs && E >= e_lo && E <= e_hi) ? __expf(lN + ls) : 0.0f;
}
// Exponent of the spectral factor S(E) = e^(-gam ln E - E invcut); the FP64
// emission term is g3Adl e^(arg - tau before the point)
__device__ __forceinline__ double bin_emis_arg(double gam, double invcut, double E, double lnEb)
{
    HD_CONTRACT
    return -gam * lnEb - E * invcut;
}
// FP32 log-form emission term before absorption: ln(g3Adl) + ln S(E)
__device__ __forceinline__ float bin_emis_log32(float lP, float gam, float invcut, double E, float lnEb)
{
    HD_CONTRACT
    return lP - gam * lnEb - float(E) * invcut;
}

// FP64 transfer update of every bin for one path point:
//   I_b   += g^3 j(E_b) e^{-tau_b} dl       (emission uses tau before the point)
//   tau_b += n_N sigma(E_b) dl              if rho > 0 and E_b inside the table
// with E_b = E_mid,b z and j = A S(E_b), S = exp(-gam ln E - E invcut).
template <int NBMAX>
__device__ __forceinline__ void point_rt64(const DevRt& P, const PointQ& q, const double* lnE, const double* lnS,
                                           const double* slope, double* tau, double* I)
{
#pragma unroll
    for (int b = 0; b < NBMAX; ++b) {
        if (b >= P.nb) continue;
        const double E = P.emid[b] * q.z;
        const double lnEb = P.lnemid[b] + q.lnz;
        if (P.spec_ok && q.g3Adl != 0.0) {
            const double arg = -P.gam * lnEb - E * P.invcut - tau[b];
            if (!(arg < EXP64_ZERO_BELOW)) I[b] += q.g3Adl * exp(arg);   // else the term is exactly 0
        }
        if (q.rho > 0.0 && E >= P.e_lo && E <= P.e_hi) {
            const double ls = sigma_ln(lnE, lnS, slope, P.n_sigma, lnEb, P.sig_lo, P.sig_inv_d);
            tau[b] += q.nbar_dl * exp(ls);
        }
    }
}

// Kahan-compensated FP32 accumulator.
struct Kahan {
    float s, c;
    __device__ __forceinline__ void add(float x)
    {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

// I accumulator with a per-bin power-of-e offset so that FP32 holds values
// like 1e-60: value = (s) * e^{off}.
struct ScaledKahan {
    Kahan k;
    float off;   // 0 until the first non-zero term
    int set;
    __device__ __forceinline__ void add_log(float L)
    {
        if (!set) {
            off = ceilf(L);
            set = 1;
        } else if (L - off > 40.0f) {
            const float noff = ceilf(L);
            const float sc = __expf(off - noff);
            k.s *= sc;
            k.c *= sc;
            off = noff;
        }
        k.add(__expf(L - off));
    }
};


// FP32-Kahan transfer of one path point (the precision form): dtau and the
// emission term in FP32 from FP64 per-point quantities; I kept in log form
// with a per-bin power-of-e offset.
template <int NBMAX>
__device__ __forceinline__ void point_rt32(const DevRt& P, const PointQ& q, const float* lnEf, const float* lnSf,
                                           const float* slopef, int n, Kahan* ktau, ScaledKahan* kI)
{
    const float lnz = float(q.lnz);
    const float lP = (q.g3Adl > 0.0) ? float(log(q.g3Adl)) : 0.0f;
    const float lN = (q.nbar_dl > 0.0) ? float(log(q.nbar_dl)) : 0.0f;
    const float gam = float(P.gam), invcut = float(P.invcut);
#pragma unroll
    for (int b = 0; b < NBMAX; ++b) {
        if (b >= P.nb) continue;
        const double E = P.emid[b] * q.z;
        const float lnEb = float(P.lnemid[b]) + lnz;
        if (P.spec_ok && q.g3Adl > 0.0) {
            const float L = lP - gam * lnEb - float(E) * invcut - ktau[b].s;
            if (!(L < EXP32_LOG_NEGLIGIBLE)) kI[b].add_log(L);
        }
        if (q.rho > 0.0 && q.nbar_dl > 0.0 && E >= P.e_lo && E <= P.e_hi)
            ktau[b].add(__expf(lN + sigma_ln_f(lnEf, lnSf, slopef, n, lnEb, float(P.sig_lo), float(P.sig_inv_d))));
    }
}

}  // namespace hd
