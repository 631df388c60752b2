#ifndef _GNU_SOURCE
#define _GNU_SOURCE
#endif
// hadros_oracle.cpp -- TEST INFRASTRUCTURE ONLY (see hadros_oracle.h).
//
// CPU FP64 restatement of the reference's per-pixel path.  REF mode keeps the
// reference's operation order so that, built with g++ -O3 -ffp-contract=off,
// it reproduces the compiled reference bit for bit.  CFG mode is the
// error-controlled integrator that the GPU library implements as well.
//
// Build: oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off -fPIC -shared).

#include "hadros_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <math.h>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double PI = 3.141592653589793238462643383279502884;  // include/constants.hpp:7
constexpr double C_CGS = 2.99792458e10;                        // include/constants.hpp:10
constexpr double G_CGS = 6.67430e-8;                           // :11
constexpr double MSUN_G = 1.98847e33;                          // :12
constexpr double M_U_G = 1.66053906660e-24;                    // :13

inline double max_(double a, double b) { return (a < b) ? b : a; }           // std::max
inline double min_(double a, double b) { return (b < a) ? b : a; }           // std::min
inline double clamp_(double v, double lo, double hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// ---------------------------------------------------------------- metric
// src/kerr_metric.cpp:10-32
inline double sigma_(double a, double r, double th) { return r * r + a * a * std::cos(th) * std::cos(th); }
inline double delta_(double a, double r) { return r * r - 2.0 * r + a * a; }
inline double bigA_(double a, double r, double th)
{
    const double s2 = std::sin(th) * std::sin(th);
    return (r * r + a * a) * (r * r + a * a) - a * a * delta_(a, r) * s2;
}

// src/kerr_metric.cpp:77-116 -> g^tt, g^tphi, g^rr, g^thth, g^phph
inline void inv_metric5(double a, double r, double th, double gi[5])
{
    const double sig = sigma_(a, r, th);
    const double del = delta_(a, r);
    const double s = std::sin(th);
    const double s2 = s * s;
    const double bigA = bigA_(a, r, th);
    gi[0] = -bigA / (sig * del);
    gi[1] = -2.0 * a * r / (sig * del);
    gi[2] = del / sig;
    gi[3] = 1.0 / sig;
    gi[4] = (del - a * a * s2) / (sig * del * s2);
}

// src/kerr_metric.cpp:34-75 -> g_tt, g_tphi, g_rr, g_thth, g_phph
inline void metric5(double a, double r, double th, double g[5])
{
    const double sig = sigma_(a, r, th);
    const double del = delta_(a, r);
    const double s = std::sin(th);
    const double s2 = s * s;
    g[0] = -(1.0 - 2.0 * r / sig);
    g[1] = -2.0 * a * r * s2 / sig;
    g[2] = sig / del;
    g[3] = sig;
    g[4] = (r * r + a * a + 2.0 * a * a * r * s2 / sig) * s2;
}

// src/kerr_metric.cpp:118-144
inline double lapse_(double a, double r, double th)
{
    const double sig = sigma_(a, r, th);
    const double del = delta_(a, r);
    const double bigA = bigA_(a, r, th);
    return std::sqrt(sig * del / bigA);
}
inline double omega_(double a, double r, double th)
{
    const double s2 = std::sin(th) * std::sin(th);
    return 2.0 * a * r / ((r * r + a * a) * (r * r + a * a) - a * a * delta_(a, r) * s2);
}

// ---------------------------------------------------------------- rhs
// State layout: 0 t, 1 r, 2 theta, 3 phi, 4 p_t, 5 p_r, 6 p_theta, 7 p_phi
// (include/geodesic_state.hpp:4-17).

// KerrGeodesic::rhs with dg_inv_dr / dg_inv_dtheta (src/kerr_geodesic.cpp:47-135).
// Each central difference of the reference evaluates the full inverse metric at
// (r +- eps, th) or (r, th +- eps) and keeps one component; the four
// neighbour metrics are therefore computed once here -- same values, same
// order of the accumulation (mu-major over the 16 (mu, nu) pairs, zero
// components contribute +-0 and leave the sums unchanged).
void rhs_ref(double a, const double y[8], double dy[8])
{
    const double r = y[1], th = y[2];
    double gi[5];
    inv_metric5(a, r, th, gi);
    const double p0 = y[4], p1 = y[5], p2 = y[6], p3 = y[7];

    dy[0] = gi[0] * p0 + gi[1] * p3;   // dydl.t  (g^t0 p0 + g^t3 p3)
    dy[1] = gi[2] * p1;
    dy[2] = gi[3] * p2;
    dy[3] = gi[1] * p0 + gi[4] * p3;   // g^30 p0 + g^33 p3
    dy[4] = 0.0;
    dy[7] = 0.0;

    const double er = 1.0e-5 * max_(1.0, std::abs(r));
    const double et = 1.0e-5;
    double gp[5], gm[5], tp[5], tm[5];
    inv_metric5(a, r + er, th, gp);
    inv_metric5(a, r - er, th, gm);
    inv_metric5(a, r, th + et, tp);
    inv_metric5(a, r, th - et, tm);

    double d_r[5], d_t[5];
    for (int c = 0; c < 5; ++c) {
        d_r[c] = (gp[c] - gm[c]) / (2.0 * er);
        d_t[c] = (tp[c] - tm[c]) / (2.0 * et);
    }
    // (mu,nu) order: (0,0) (0,3) (1,1) (2,2) (3,0) (3,3)
    double pr = 0.0, pth = 0.0;
    pr -= 0.5 * d_r[0] * p0 * p0;   pth -= 0.5 * d_t[0] * p0 * p0;
    pr -= 0.5 * d_r[1] * p0 * p3;   pth -= 0.5 * d_t[1] * p0 * p3;
    pr -= 0.5 * d_r[2] * p1 * p1;   pth -= 0.5 * d_t[2] * p1 * p1;
    pr -= 0.5 * d_r[3] * p2 * p2;   pth -= 0.5 * d_t[3] * p2 * p2;
    pr -= 0.5 * d_r[1] * p3 * p0;   pth -= 0.5 * d_t[1] * p3 * p0;
    pr -= 0.5 * d_r[4] * p3 * p3;   pth -= 0.5 * d_t[4] * p3 * p3;
    dy[5] = pr;
    dy[6] = pth;
}

// Analytic Hamiltonian derivative (CFG mode).  With W = (r^2+a^2) p_t + a p_phi,
// V = p_phi / sin(th) + a sin(th) p_t,
//   2 Sigma H = Delta p_r^2 + p_th^2 - W^2/Delta + V^2 =: N,
// which equals g^{mu nu} p_mu p_nu Sigma of src/kerr_metric.cpp:77-116.
void rhs_cfg(double a, const double y[8], double dy[8])
{
    const double r = y[1], th = y[2];
    const double pt = y[4], pr = y[5], pth = y[6], pph = y[7];
    const double s = std::sin(th), c = std::cos(th);
    const double r2 = r * r, a2 = a * a;
    const double sig = r2 + a2 * c * c;
    const double del = r2 - 2.0 * r + a2;
    const double ra = r2 + a2;
    const double W = ra * pt + a * pph;
    const double V = pph / s + a * s * pt;
    const double inv_sig = 1.0 / sig;
    const double inv_del = 1.0 / del;
    const double Wd = W * inv_del;
    const double N = del * pr * pr + pth * pth - W * Wd + V * V;
    const double ddel = 2.0 * r - 2.0;
    const double dNr = ddel * pr * pr - 4.0 * r * pt * Wd + Wd * Wd * ddel;
    const double Vth = c * (a * pt - pph / (s * s));
    const double dNth = 2.0 * V * Vth;
    const double half_inv_sig = 0.5 * inv_sig;
    const double n_over = N * half_inv_sig * inv_sig;   // N / (2 Sigma^2)
    dy[0] = (a * s * V - ra * Wd) * inv_sig;
    dy[1] = del * pr * inv_sig;
    dy[2] = pth * inv_sig;
    dy[3] = (V / s - a * Wd) * inv_sig;
    dy[4] = 0.0;
    dy[5] = -(dNr * half_inv_sig - n_over * (2.0 * r));
    dy[6] = -(dNth * half_inv_sig - n_over * (-2.0 * a2 * c * s));
    dy[7] = 0.0;
}

inline void rhs(double a, int mode, const double y[8], double dy[8])
{
    if (mode == HADROS_MODE_CFG) rhs_cfg(a, y, dy); else rhs_ref(a, y, dy);
}

// ---------------------------------------------------------------- RKF45
// Stage combination of src/kerr_geodesic.cpp:196-217: y0 + a1 k1 + ... evaluated
// left to right (terms with a zero coefficient add +0 and are dropped).
struct RKF {
    double y4[8], y5[8];
};

void rkf45(double a, int mode, const double y[8], double h, RKF& o)
{
    double k1[8], k2[8], k3[8], k4[8], k5[8], k6[8], t[8];
    rhs(a, mode, y, k1);
    const double a21 = h * 1.0 / 4.0;
    for (int c = 0; c < 8; ++c) t[c] = y[c] + a21 * k1[c];
    rhs(a, mode, t, k2);
    const double a31 = h * 3.0 / 32.0, a32 = h * 9.0 / 32.0;
    for (int c = 0; c < 8; ++c) t[c] = y[c] + a31 * k1[c] + a32 * k2[c];
    rhs(a, mode, t, k3);
    const double a41 = h * 1932.0 / 2197.0, a42 = h * -7200.0 / 2197.0, a43 = h * 7296.0 / 2197.0;
    for (int c = 0; c < 8; ++c) t[c] = y[c] + a41 * k1[c] + a42 * k2[c] + a43 * k3[c];
    rhs(a, mode, t, k4);
    const double a51 = h * 439.0 / 216.0, a52 = h * -8.0, a53 = h * 3680.0 / 513.0, a54 = h * -845.0 / 4104.0;
    for (int c = 0; c < 8; ++c) t[c] = y[c] + a51 * k1[c] + a52 * k2[c] + a53 * k3[c] + a54 * k4[c];
    rhs(a, mode, t, k5);
    const double a61 = h * -8.0 / 27.0, a62 = h * 2.0, a63 = h * -3544.0 / 2565.0, a64 = h * 1859.0 / 4104.0,
                 a65 = h * -11.0 / 40.0;
    for (int c = 0; c < 8; ++c) t[c] = y[c] + a61 * k1[c] + a62 * k2[c] + a63 * k3[c] + a64 * k4[c] + a65 * k5[c];
    rhs(a, mode, t, k6);
    // src/kerr_geodesic.cpp:288-308
    for (int c = 0; c < 8; ++c) {
        o.y4[c] = y[c] + h * (25.0 / 216.0 * k1[c] + 1408.0 / 2565.0 * k3[c] + 2197.0 / 4104.0 * k4[c]
                              - 1.0 / 5.0 * k5[c]);
        o.y5[c] = y[c] + h * (16.0 / 135.0 * k1[c] + 6656.0 / 12825.0 * k3[c] + 28561.0 / 56430.0 * k4[c]
                              - 9.0 / 50.0 * k5[c] + 2.0 / 55.0 * k6[c]);
    }
}

// KerrGeodesic::step_rk4 (src/kerr_geodesic.cpp:137-192)
void rk4_ref(double a, double h, double y[8])
{
    double k1[8], k2[8], k3[8], k4[8], t[8];
    rhs_ref(a, y, k1);
    for (int c = 0; c < 8; ++c) t[c] = y[c] + 0.5 * h * k1[c];
    rhs_ref(a, t, k2);
    for (int c = 0; c < 8; ++c) t[c] = y[c] + 0.5 * h * k2[c];
    rhs_ref(a, t, k3);
    for (int c = 0; c < 8; ++c) t[c] = y[c] + h * k3[c];
    rhs_ref(a, t, k4);
    for (int c = 0; c < 8; ++c) y[c] += h * (k1[c] + 2.0 * k2[c] + 2.0 * k3[c] + k4[c]) / 6.0;
}

// error_norm of src/kerr_geodesic.cpp:219-233 (std::max keeps err on NaN)
inline double err_ref(const RKF& o)
{
    double err = 0.0;
    const int comp[6] = {1, 2, 3, 5, 6, 7};
    for (int q = 0; q < 6; ++q) err = max_(err, std::abs(o.y4[comp[q]] - o.y5[comp[q]]));
    return err;
}

// Mixed error of CFG mode: max |y5-y4| / (tol (1 + max(|y|, |y5|))) over the
// evolving components r, theta, phi, p_r, p_theta.
inline double err_cfg(const double y[8], const RKF& o, double tol)
{
    double err = 0.0;
    const int comp[5] = {1, 2, 3, 5, 6};
    for (int q = 0; q < 5; ++q) {
        const int c = comp[q];
        const double sc = tol * (1.0 + std::fmax(std::abs(y[c]), std::abs(o.y5[c])));
        const double e = std::abs(o.y5[c] - o.y4[c]) / sc;
        if (!(e <= err)) err = e;   // NaN propagates (forces rejection)
    }
    return err;
}

// KerrGeodesic::step_adaptive (src/kerr_geodesic.cpp:194-329).  Returns the
// number of RKF45 attempts made.
int step_ref(double a, double h0, double tol, double y[8])
{
    double h = h0;
    const double h_min = h0 * 1.0e-5;
    RKF o;
    for (int attempt = 0; attempt < 50; ++attempt) {
        rkf45(a, HADROS_MODE_REF, y, h, o);
        const double err = err_ref(o);
        if (err < tol || h <= h_min) {
            std::memcpy(y, o.y5, sizeof(o.y5));
            return attempt + 1;
        }
        const double safety = 0.8;
        const double factor = safety * std::pow(tol / max_(err, 1.0e-30), 0.25);
        h *= clamp_(factor, 0.1, 5.0);
        if (h < h_min) h = h_min;
    }
    rk4_ref(a, h0, y);
    return 50;
}

// CFG step control: accept if err <= 1 (or h at its floor), then grow/shrink
// by 0.9 err^-1/5 within [0.2, 5] (accept) or [0.1, 0.9] (reject).
constexpr double CFG_H_MIN = 1.0e-9;

struct CfgStepper {
    double h;
};

// One attempt; returns true if accepted (y updated).
bool attempt_cfg(double a, double tol, CfgStepper& st, double y[8])
{
    RKF o;
    rkf45(a, HADROS_MODE_CFG, y, st.h, o);
    const double err = err_cfg(y, o, tol);
    if (err <= 1.0 || st.h <= CFG_H_MIN) {
        std::memcpy(y, o.y5, sizeof(o.y5));
        double fac = (err > 0.0) ? 0.9 * std::pow(err, -0.2) : 5.0;
        fac = std::fmin(5.0, std::fmax(0.2, fac));
        st.h = std::fmax(st.h * fac, CFG_H_MIN);
        return true;
    }
    double fac = (err == err) ? 0.9 * std::pow(err, -0.2) : 0.1;
    fac = std::fmin(0.9, std::fmax(0.1, fac));
    st.h = std::fmax(st.h * fac, CFG_H_MIN);
    return false;
}

// ---------------------------------------------------------------- camera
// src/kerr_camera.cpp:14-29
inline double zamo_energy(double a, double r, double th, double pt, double pph)
{
    const double alpha = lapse_(a, r, th);
    const double omega = omega_(a, r, th);
    return -(pt + omega * pph) / alpha;
}

// src/kerr_camera.cpp:31-44
inline double wrapped_dphi(double pn, double po)
{
    double d = pn - po;
    while (d > PI) d -= 2.0 * PI;
    while (d < -PI) d += 2.0 * PI;
    return d;
}

// src/kerr_camera.cpp:46-76
inline double zamo_dl(double a, const double* cur, const double* prev)
{
    const double r_mid = 0.5 * (cur[1] + prev[1]);
    const double th_mid = 0.5 * (cur[2] + prev[2]);
    double g[5];
    metric5(a, r_mid, th_mid, g);
    const double dr = cur[1] - prev[1];
    const double dth = cur[2] - prev[2];
    const double dph = wrapped_dphi(cur[3], prev[3]);
    const double dl2 = g[2] * dr * dr + g[3] * dth * dth + g[4] * dph * dph;
    return std::sqrt(max_(dl2, 0.0));
}

inline double fov_rad(const hadros_camera* c) { return c->fov_deg * PI / 180.0; }   // :78-81

inline double pix_u(const hadros_camera* c, int i)
{
    return (2.0 * (i + 0.5) / c->nx - 1.0) * std::tan(0.5 * fov_rad(c));
}
inline double pix_v(const hadros_camera* c, int j)
{
    return (2.0 * (j + 0.5) / c->ny - 1.0) * std::tan(0.5 * fov_rad(c));
}

// src/kerr_camera.cpp:110-186
void initial_state(const hadros_camera* c, int i, int j, double y[8])
{
    const double a = c->a_spin;
    const double u = pix_u(c, i);
    const double v = pix_v(c, j);
    const double norm = std::sqrt(1.0 + u * u + v * v);
    const double n_r = -1.0 / norm;
    const double n_th = v / norm;
    const double n_ph = u / norm;
    double g[5];
    metric5(a, c->r_obs, c->theta_obs_rad, g);
    const double alpha = lapse_(a, c->r_obs, c->theta_obs_rad);
    const double omega = omega_(a, c->r_obs, c->theta_obs_rad);
    const double pt_c = 1.0 / alpha;
    const double pr_c = n_r / std::sqrt(g[2]);
    const double pth_c = n_th / std::sqrt(g[3]);
    const double pph_c = n_ph / std::sqrt(g[4]) + omega * pt_c;
    y[0] = 0.0;
    y[1] = c->r_obs;
    y[2] = c->theta_obs_rad;
    y[3] = 0.0;
    y[4] = g[0] * pt_c + g[1] * pph_c;
    y[5] = g[2] * pr_c;
    y[6] = g[3] * pth_c;
    y[7] = g[1] * pt_c + g[4] * pph_c;
}

// ---------------------------------------------------------------- medium
// src/torus_profile.cpp:12-42 (constructor clamps) and :93-272 (shapes).
struct Medium {
    int kind;
    double rho0, r0, sr, H, p, fd, ft, env0, ealpha, rmin, rmax, floor_;
    explicit Medium(const hadros_medium* m)
        : kind(m->kind), rho0(m->rho0), r0(m->r0), sr(m->sigma_r), H(m->h_over_r), p(m->radial_power),
          fd(clamp_(m->funnel_depletion, 0.0, 1.0)), ft(max_(m->funnel_theta_rad, 1.0e-6)),
          env0(max_(m->envelope_rho0, 0.0)), ealpha(m->envelope_alpha), rmin(m->r_min),
          rmax(max_(m->r_max, m->r_min)), floor_(max_(m->rho_floor, 0.0))
    {
    }
    bool powerlaw_disk() const
    {
        return kind == HADROS_PROFILE_POWERLAW || kind == HADROS_PROFILE_POWERLAW_FUNNEL ||
               kind == HADROS_PROFILE_POWERLAW_ENVELOPE || kind == HADROS_PROFILE_POWERLAW_FUNNEL_ENVELOPE;
    }
    bool funnel() const
    {
        return kind == HADROS_PROFILE_GAUSSIAN_FUNNEL || kind == HADROS_PROFILE_POWERLAW_FUNNEL ||
               kind == HADROS_PROFILE_POWERLAW_FUNNEL_ENVELOPE;
    }
    bool envelope() const
    {
        return kind == HADROS_PROFILE_GAUSSIAN_ENVELOPE || kind == HADROS_PROFILE_POWERLAW_ENVELOPE ||
               kind == HADROS_PROFILE_POWERLAW_FUNNEL_ENVELOPE;
    }
    bool in_torus(double r) const { return r >= rmin && r <= rmax; }

    double collapsar(double r, double th) const   // :150-179
    {
        if (!in_torus(r) || r <= 0.0 || r0 <= 0.0) return 0.0;
        const double mu = std::cos(th);
        const double vertical = std::exp(-std::pow(mu / max_(H, 1.0e-6), 2.0));
        const double inner_taper = 1.0 / (1.0 + std::exp(-(r - rmin) / max_(0.35 * sr, 1.0e-6)));
        const double outer_taper = 1.0 / (1.0 + std::exp((r - rmax) / max_(0.60 * sr, 1.0e-6)));
        const double inner_core = 0.45 * std::exp(-std::pow((r - 0.75 * r0) / max_(0.75 * sr, 1.0e-6), 2.0));
        const double radial = std::pow(max_(r / r0, 1.0e-300), -max_(p, 0.0));
        const double floor_tail = 0.08 * std::exp(-std::pow((r - r0) / max_(1.8 * sr, 1.0e-6), 2.0));
        return (radial + inner_core + floor_tail) * vertical * inner_taper * outer_taper;
    }
    double gaussian(double r, double th) const   // :181-196
    {
        const double delta = th - 0.5 * PI;
        if (r <= 4.0 || r >= 18.0 || std::abs(delta) >= 0.45) return 0.0;
        const double radial = std::exp(-std::pow((r - r0) / sr, 2.0));
        const double vertical = std::exp(-std::pow(delta / H, 2.0));
        return radial * vertical;
    }
    double powerlaw(double r, double th) const   // :198-217
    {
        if (!in_torus(r) || r <= 0.0 || r0 <= 0.0) return 0.0;
        const double vertical = std::exp(-std::pow(std::cos(th) / H, 2.0));
        const double inner_taper = 1.0 - std::exp(-std::pow((r - rmin) / sr, 2.0));
        const double outer_taper = std::exp(-std::pow(r / rmax, 4.0));
        const double radial = std::pow(r / r0, -p);
        return radial * vertical * inner_taper * outer_taper;
    }
    double funnel_factor(double th) const   // :219-236
    {
        if (!funnel()) return 1.0;
        const double north = std::exp(-std::pow(th / ft, 2.0));
        const double south = std::exp(-std::pow((PI - th) / ft, 2.0));
        return clamp_(1.0 - fd * (north + south), 0.0, 1.0);
    }
    double envelope_rho(double r) const   // :238-248
    {
        if (!envelope() || env0 <= 0.0 || r < r0 || r > rmax) return 0.0;
        return env0 * std::pow(r / r0, -ealpha);
    }
    double disk_shape(double r, double th) const   // :250-259
    {
        if (kind == HADROS_PROFILE_COLLAPSAR_NDAF_LIKE) return collapsar(r, th);
        return powerlaw_disk() ? powerlaw(r, th) : gaussian(r, th);
    }
    double rho(double r, double th) const   // :261-272
    {
        const double disk = rho0 * disk_shape(r, th) * funnel_factor(th);
        return max_(disk + envelope_rho(r), floor_);
    }
};

// ---------------------------------------------------------------- source / spectrum
// src/radiative_transfer.cpp:64-76
inline double bipolar(double th, double th0, double w)
{
    if (w <= 0.0) return 0.0;
    const double north = th - th0;
    const double south = th - (PI - th0);
    return std::exp(-(north / w) * (north / w)) + std::exp(-(south / w) * (south / w));
}

// src/radiative_transfer.cpp:172-207
double spectral_weight(const hadros_scene* s, double E)
{
    if (E <= 0.0) return 0.0;
    const double safe_E = max_(E, 1.0e-300);
    const hadros_spectrum& sp = s->spectrum;
    if (sp.kind == HADROS_SPECTRUM_MONOCHROMATIC) {
        if (s->source.emax_GeV <= 0.0) return 0.0;
        return std::pow(safe_E, -s->source.powerlaw) * std::exp(-safe_E / s->source.emax_GeV);
    }
    if (sp.kind == HADROS_SPECTRUM_POWERLAW) return std::pow(safe_E, -sp.gamma);
    if (sp.kind == HADROS_SPECTRUM_POWERLAW_CUTOFF) {
        if (sp.ecut_GeV <= 0.0) return 0.0;
        return std::pow(safe_E, -sp.gamma) * std::exp(-safe_E / sp.ecut_GeV);
    }
    return 0.0;
}

// src/radiative_transfer.cpp:209-333 (rho_norm is the per-scene constant
// torus.rho(r_center, pi/2) the reference re-evaluates at every call)
double source_weight(const hadros_scene* s, const Medium& med, double rho_norm_c, double r, double th)
{
    const hadros_source& so = s->source;
    if (r <= 0.0 || so.sigma_r <= 0.0 || so.theta_width_rad <= 0.0) return 0.0;
    switch (so.kind) {
    case HADROS_SOURCE_INNER_RING: {
        const double dr = (r - so.r_center) / so.sigma_r;
        const double dt = (th - 0.5 * PI) / so.theta_width_rad;
        return std::exp(-dr * dr) * std::exp(-dt * dt);
    }
    case HADROS_SOURCE_FUNNEL_WALL: {
        const double dr = (r - so.r_center) / so.sigma_r;
        return std::exp(-dr * dr) * bipolar(th, so.funnel_theta_rad, so.theta_width_rad);
    }
    case HADROS_SOURCE_JET_BASE: {
        const double radial = std::exp(-std::pow(r / so.r_center, 2.0));
        const double polar = bipolar(th, 0.0, so.theta_width_rad);
        return radial * polar;
    }
    case HADROS_SOURCE_DENSITY_WEIGHTED: {
        const double rho = med.rho(r, th);
        const double rho_norm = max_(so.rho_ref > 0.0 ? so.rho_ref : rho_norm_c, 1.0e-300);
        const double dw = std::pow(max_(rho / rho_norm, 0.0), so.density_q);
        const double rw = std::pow(max_(r / max_(so.r_center, 1.0e-300), 1.0e-300), -so.radial_s);
        return clamp_(dw * rw, max_(so.cutoff_min, 0.0), max_(so.cutoff_max, max_(so.cutoff_min, 0.0)));
    }
    case HADROS_SOURCE_SHOCK_LAYER: {
        const double dr = max_(so.gradient_dr, 1.0e-4);
        const double dth = max_(so.gradient_dtheta_rad, 1.0e-5);
        const double r_minus = max_(r - dr, 1.0e-6);
        const double th_minus = clamp_(th - dth, 0.0, PI);
        const double th_plus = clamp_(th + dth, 0.0, PI);
        const double rp = med.rho(r + dr, th);
        const double rm = med.rho(r_minus, th);
        const double tp = med.rho(r, th_plus);
        const double tm = med.rho(r, th_minus);
        const double rho_norm = max_(rho_norm_c, 1.0e-300);
        const double d_r = (rp - rm) / (2.0 * dr);
        const double d_t = (tp - tm) / (2.0 * dth * max_(r, 1.0e-6));
        const double grad = std::sqrt(d_r * d_r + d_t * d_t) / rho_norm;
        const double taper = std::exp(-std::pow(r / max_(so.r_center, 1.0e-300), 2.0));
        return clamp_(grad * taper, max_(so.cutoff_min, 0.0), max_(so.cutoff_max, max_(so.cutoff_min, 0.0)));
    }
    default:
        return 0.0;
    }
}

// ---------------------------------------------------------------- sigma
// src/sigma_table.cpp:80-124 with the table logs hoisted (log of a table node
// is the same double at every call).
struct Sigma {
    const double* e;
    const double* s;
    int n;
    std::vector<double> le, ls;
    Sigma(const double* e_, const double* s_, int n_) : e(e_), s(s_), n(n_), le(n_), ls(n_)
    {
        for (int k = 0; k < n; ++k) {
            le[k] = std::log(e[k]);
            ls[k] = std::log(s[k]);
        }
    }
    double emin() const { return e[0]; }
    double emax() const { return e[n - 1]; }
    double eval(double x) const
    {
        if (!(x > 0.0)) return std::numeric_limits<double>::quiet_NaN();
        if (x < e[0] || x > e[n - 1]) return std::numeric_limits<double>::quiet_NaN();
        const double* it = std::lower_bound(e, e + n, x);
        if (it == e) return s[0];
        if (it == e + n) return s[n - 1];
        const int i = int(it - e);
        const double lx = std::log(x);
        const double t = (lx - le[i - 1]) / (le[i] - le[i - 1]);
        return std::exp(ls[i - 1] + t * (ls[i] - ls[i - 1]));
    }
};

// ---------------------------------------------------------------- transfer
struct RtCtx {
    const hadros_scene* s;
    Medium med;
    Sigma sig;
    double rg_cm;
    double r_h_rt;
    double rho_norm_c;
    int nb;            // integrated bins (>= 0); -1 = empty spectral range
    bool mono;
    double e_mid[HADROS_MAX_BINS], de[HADROS_MAX_BINS], w[HADROS_MAX_BINS];
    bool ring_ok;      // emissivity_collapsar_ring parameter guard (:343-349)

    RtCtx(const hadros_scene* s_, const double* se, const double* ss, int ns, double a_bh)
        : s(s_), med(&s_->medium), sig(se, ss, ns)
    {
        rg_cm = G_CGS * s->m_bh_msun * MSUN_G / (C_CGS * C_CGS);   // include/constants.hpp:17-20
        const double a = clamp_(a_bh, 0.0, 0.999999);              // src/radiative_transfer.cpp:485-489
        r_h_rt = 1.0 + std::sqrt(1.0 - a * a);
        rho_norm_c = med.rho(s->source.r_center, 0.5 * PI);
        mono = s->spectrum.kind == HADROS_SPECTRUM_MONOCHROMATIC;
        nb = or_bins(s, se, ns, e_mid, de, w);
        const hadros_source& so = s->source;
        ring_ok = !(so.r_center <= 0.0 || so.sigma_r <= 0.0 || so.theta_width_rad <= 0.0 || so.emax_GeV <= 0.0);
    }

    // One path point of integrate_kerr_ray (:501-672) for every bin.
    void point(double r, double th, double dl, double z, double* tau, double* I) const
    {
        if (r < r_h_rt + 1.0e-3) return;
        const double g = 1.0 / max_(z, 1.0e-300);
        const double dl_cm = dl * rg_cm;
        const double rho = med.rho(r, th);
        // spatial factor of emissivity_uhe (:369-395); emissivity_collapsar_ring
        // multiplies norm * radial * vertical before the spectral factor.
        double A;
        const hadros_source& so = s->source;
        if (so.kind == HADROS_SOURCE_INNER_RING) {
            if (ring_ok) {
                const double dr = (r - so.r_center) / so.sigma_r;
                const double dt = (th - 0.5 * PI) / so.theta_width_rad;
                A = so.norm * std::exp(-dr * dr) * std::exp(-dt * dt);
            } else {
                A = 0.0;
            }
        } else {
            A = so.norm * source_weight(s, med, rho_norm_c, r, th);
        }
        const double g3 = std::pow(g, 3.0);
        const double nbar = rho / M_U_G;
        for (int b = 0; b < nb; ++b) {
            const double E = e_mid[b] / max_(g, 1.0e-300);
            double j = 0.0;
            if (E > 0.0 && (so.kind != HADROS_SOURCE_INNER_RING || ring_ok)) j = A * spectral_weight(s, E);
            I[b] += g3 * j * std::exp(-tau[b]) * dl_cm;
            if (rho <= 0.0 || E < sig.emin() || E > sig.emax()) continue;
            const double sc = sig.eval(E);
            tau[b] += nbar * sc * dl_cm;
        }
    }

    // aggregation of integrate_kerr_ray_spectral (:711-769)
    void finish(const double* tau, const double* I, double agg[3], double* bt, double* bp, double* bi) const
    {
        if (nb < 0) {
            agg[0] = 0.0; agg[1] = 1.0; agg[2] = 0.0;
            return;
        }
        for (int b = 0; b < nb; ++b) {
            if (bt) bt[b] = tau[b];
            if (bp) bp[b] = std::exp(-tau[b]);
            if (bi) bi[b] = I[b];
        }
        if (mono) {
            agg[0] = tau[0];
            agg[1] = std::exp(-tau[0]);
            agg[2] = I[0];
            return;
        }
        double ws = 0.0, ts = 0.0, ps = 0.0, is = 0.0;
        for (int b = 0; b < nb; ++b) {
            ws += w[b];
            ts += w[b] * tau[b];
            ps += w[b] * std::exp(-tau[b]);
            is += I[b] * de[b];
        }
        if (ws > 0.0) {
            agg[0] = ts / ws; agg[1] = ps / ws; agg[2] = is;
        } else {
            agg[0] = 0.0; agg[1] = 1.0; agg[2] = 0.0;
        }
    }
};

// ---------------------------------------------------------------- tracing
// The loop of KerrCamera::trace_pixel (src/kerr_camera.cpp:188-283); `emit`
// receives every PathPoint and the state it was taken from.  REF: stop at r <= r_+ + 1e-3; CFG: r < 1.01 r_+.
template <class Emit>
void trace(const hadros_camera* c, int i, int j, Emit&& emit, int& captured, int& exhausted, int64_t& attempts)
{
    const double a = c->a_spin;
    double y[8], yp[8];
    initial_state(c, i, j, y);
    std::memcpy(yp, y, sizeof(y));
    const double energy_obs = zamo_energy(a, c->r_obs, c->theta_obs_rad, y[4], y[7]);
    const double r_h = 1.0 + std::sqrt(1.0 - a * a);
    const bool cfg = c->mode == HADROS_MODE_CFG;
    const double r_stop = cfg ? 1.01 * r_h : r_h + 1.0e-3;
    captured = 0;
    exhausted = 1;
    attempts = 0;
    CfgStepper st{c->h};
    const int64_t max_attempts = 8 * int64_t(c->max_steps) + 64;
    for (int step = 0; step < c->max_steps; ++step) {
        if (cfg ? (y[1] < r_stop) : (y[1] <= r_stop)) {
            captured = 1;
            exhausted = 0;
            break;
        }
        if (cfg ? (y[1] > c->r_max && step > 10) : (y[1] >= c->r_max && step > 10)) {
            exhausted = 0;
            break;
        }
        // g++ -O3 turns the reference's sin/cos pairs of one argument in
        // trace_pixel (src/kerr_camera.cpp:236-243) into sincos(), whose
        // last bit differs from sin/cos for ~0.1% of inputs; x, y, z follow.
        double sth, cth, sph, cph;
        ::sincos(y[2], &sth, &cth);
        ::sincos(y[3], &sph, &cph);
        double pt[7];
        pt[0] = y[1];
        pt[1] = y[2];
        pt[2] = y[1] * sth * cph;
        pt[3] = y[1] * sth * sph;
        pt[4] = y[1] * cth;
        pt[5] = (step == 0) ? 0.0 : zamo_dl(a, y, yp);
        pt[6] = zamo_energy(a, y[1], y[2], y[4], y[7]) / max_(energy_obs, 1.0e-300);
        emit(pt, y);
        std::memcpy(yp, y, sizeof(y));
        if (cfg) {
            bool ok = false;
            while (!ok) {
                ++attempts;
                ok = attempt_cfg(a, c->tol, st, y);
                if (!ok && attempts >= max_attempts) break;
            }
            if (!ok) break;   // exhausted stays 1
        } else {
            attempts += step_ref(a, c->h, c->tol, y);
        }
    }
}

int check_scene(const hadros_scene* s, int n_sig)
{
    if (!s || n_sig < 2 || n_sig > HADROS_MAX_SIGMA) return HADROS_ERR_ARG;
    if (s->spectrum.kind != HADROS_SPECTRUM_MONOCHROMATIC && std::max(s->spectrum.n_bins, 1) > HADROS_MAX_BINS)
        return HADROS_ERR_ARG;
    return HADROS_OK;
}

}  // namespace

// ================================================================ C API
extern "C" {

void or_metric(double a, double r, double th, double out[5]) { metric5(a, r, th, out); }
void or_inverse_metric(double a, double r, double th, double out[5]) { inv_metric5(a, r, th, out); }
double or_lapse(double a, double r, double th) { return lapse_(a, r, th); }
double or_omega(double a, double r, double th) { return omega_(a, r, th); }
double or_horizon(double a) { return 1.0 + std::sqrt(1.0 - a * a); }

void or_rhs(double a, int mode, const double y[8], double dy[8]) { rhs(a, mode, y, dy); }

// src/kerr_geodesic.cpp:17-45 (all 16 terms, zero components included)
double or_hamiltonian(double a, const double y[8])
{
    double gi[5];
    inv_metric5(a, y[1], y[2], gi);
    double G[4][4] = {{gi[0], 0, 0, gi[1]}, {0, gi[2], 0, 0}, {0, 0, gi[3], 0}, {gi[1], 0, 0, gi[4]}};
    const double p[4] = {y[4], y[5], y[6], y[7]};
    double H = 0.0;
    for (int mu = 0; mu < 4; ++mu)
        for (int nu = 0; nu < 4; ++nu) H += 0.5 * G[mu][nu] * p[mu] * p[nu];
    return H;
}

void or_step_ref(double a, double h, double tol, double y[8]) { step_ref(a, h, tol, y); }

void or_step_rk4(double a, double h, double y[8]) { rk4_ref(a, h, y); }

void or_rkf45_attempt(double a, int mode, double h, const double y[8], double y4[8], double y5[8])
{
    RKF o;
    rkf45(a, mode, y, h, o);
    std::memcpy(y4, o.y4, sizeof(o.y4));
    std::memcpy(y5, o.y5, sizeof(o.y5));
}

int64_t or_trace_states(const hadros_camera* cam, int i, int j, double* states, int64_t cap)
{
    int64_t n = 0;
    int cp = 0, ex = 0;
    int64_t at = 0;
    trace(cam, i, j,
          [&](const double*, const double* y) {
              if (states && n < cap) std::memcpy(states + 8 * n, y, 8 * sizeof(double));
              ++n;
          },
          cp, ex, at);
    return n;
}

void or_initial_state(const hadros_camera* cam, int i, int j, double y[8]) { initial_state(cam, i, j, y); }

int64_t or_trace_pixel(const hadros_camera* cam, int i, int j, double* pts, int64_t cap, int* captured,
                       int* exhausted, int64_t* attempts)
{
    int64_t n = 0;
    int cp = 0, ex = 0;
    int64_t at = 0;
    trace(cam, i, j,
          [&](const double* p, const double*) {
              if (pts && n < cap) std::memcpy(pts + 7 * n, p, 7 * sizeof(double));
              ++n;
          },
          cp, ex, at);
    if (captured) *captured = cp;
    if (exhausted) *exhausted = ex;
    if (attempts) *attempts = at;
    return n;
}

double or_sigma(const double* e, const double* s, int n, double x)
{
    Sigma sg(e, s, n);
    return sg.eval(x);
}

double or_rho(const hadros_medium* m, double r, double th) { return Medium(m).rho(r, th); }

double or_spectral_weight(const hadros_scene* s, double e) { return spectral_weight(s, e); }

double or_source_weight(const hadros_scene* s, double r, double th)
{
    Medium med(&s->medium);
    return source_weight(s, med, med.rho(s->source.r_center, 0.5 * PI), r, th);
}

// src/radiative_transfer.cpp:335-395
double or_emissivity(const hadros_scene* s, double r, double th, double E)
{
    if (E <= 0.0) return 0.0;
    const hadros_source& so = s->source;
    if (so.kind == HADROS_SOURCE_INNER_RING) {
        if (so.r_center <= 0.0 || so.sigma_r <= 0.0 || so.theta_width_rad <= 0.0 || so.emax_GeV <= 0.0) return 0.0;
        const double dr = (r - so.r_center) / so.sigma_r;
        const double dt = (th - 0.5 * PI) / so.theta_width_rad;
        const double radial = std::exp(-dr * dr);
        const double vertical = std::exp(-dt * dt);
        return so.norm * radial * vertical * spectral_weight(s, E);
    }
    return so.norm * or_source_weight(s, r, th) * spectral_weight(s, E);
}

// src/radiative_transfer.cpp:698-761
int or_bins(const hadros_scene* s, const double* sig_e, int n_sig, double* e_mid, double* de, double* w)
{
    if (s->spectrum.kind == HADROS_SPECTRUM_MONOCHROMATIC) {
        e_mid[0] = s->enu_obs_GeV;
        de[0] = 1.0;
        w[0] = 1.0;
        return 1;
    }
    const int n_bins = std::max(s->spectrum.n_bins, 1);
    const double e_min = max_(s->spectrum.e_min_GeV, sig_e[0]);
    const double e_max = min_(s->spectrum.e_max_GeV, sig_e[n_sig - 1]);
    if (e_min <= 0.0 || e_max <= e_min) return -1;
    const double log_min = std::log(e_min);
    const double log_max = std::log(e_max);
    const double dlog = (log_max - log_min) / static_cast<double>(n_bins);
    int nb = 0;
    for (int i = 0; i < n_bins && i < HADROS_MAX_BINS; ++i) {
        const double log_left = log_min + dlog * static_cast<double>(i);
        const double log_right = log_left + dlog;
        const double E_mid = std::exp(0.5 * (log_left + log_right));
        const double dE = std::exp(log_right) - std::exp(log_left);
        const double wi = spectral_weight(s, E_mid) * dE;
        if (wi <= 0.0 || !std::isfinite(wi)) continue;
        e_mid[nb] = E_mid;
        de[nb] = dE;
        w[nb] = wi;
        ++nb;
    }
    return nb;
}

void or_integrate_ray(const hadros_scene* s, const double* sig_e, const double* sig_s, int n_sig, double a_bh,
                      const double* pts, int64_t n, double out_agg[3], double* bin_tau, double* bin_p,
                      double* bin_i)
{
    RtCtx ctx(s, sig_e, sig_s, n_sig, a_bh);
    double tau[HADROS_MAX_BINS] = {0}, I[HADROS_MAX_BINS] = {0};
    for (int64_t k = 0; k < n; ++k) {
        const double* p = pts + 7 * k;
        ctx.point(p[0], p[1], p[5], p[6], tau, I);
    }
    ctx.finish(tau, I, out_agg, bin_tau, bin_p, bin_i);
}

int or_trace_image_range(const hadros_camera* cam, const hadros_scene* s, const double* sig_e,
                         const double* sig_s, int n_sig, int nthreads, int64_t k0, int64_t k1,
                         hadros_image_out* out, hadros_stats* st)
{
    if (!cam || !out || cam->nx <= 0 || cam->ny <= 0) return HADROS_ERR_ARG;
    const int rc = check_scene(s, n_sig);
    if (rc) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    RtCtx ctx(s, sig_e, sig_s, n_sig, cam->a_spin);
    const int nbo = ctx.nb > 0 ? ctx.nb : 0;
    int64_t tot_pts = 0, tot_att = 0;
    int n_cap = 0, n_ex = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic) reduction(+ : tot_pts, tot_att, n_cap, n_ex)
    for (int64_t k = k0; k < k1; ++k) {
        const int i = int(k / cam->ny), j = int(k % cam->ny);
        double tau[HADROS_MAX_BINS] = {0}, I[HADROS_MAX_BINS] = {0};
        int64_t np = 0;
        int cp = 0, ex = 0;
        int64_t at = 0;
        trace(cam, i, j,
              [&](const double* p, const double*) {
                  ctx.point(p[0], p[1], p[5], p[6], tau, I);
                  ++np;
              },
              cp, ex, at);
        double agg[3];
        ctx.finish(tau, I, agg, out->bin_tau ? out->bin_tau + k * nbo : nullptr,
                   out->bin_p ? out->bin_p + k * nbo : nullptr, out->bin_i ? out->bin_i + k * nbo : nullptr);
        if (out->tau) out->tau[k] = agg[0];
        if (out->p_surv) out->p_surv[k] = agg[1];
        if (out->intensity) out->intensity[k] = agg[2];
        if (out->captured) out->captured[k] = cp;
        if (out->npoints) out->npoints[k] = int32_t(np);
        if (out->flags) out->flags[k] = ex ? HADROS_FLAG_EXHAUSTED : 0;
        if (out->alpha) out->alpha[k] = pix_u(cam, i);
        if (out->beta) out->beta[k] = pix_v(cam, j);
        tot_pts += np;
        tot_att += at;
        n_cap += cp;
        n_ex += ex;
    }
    if (st) {
        st->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        st->kernel_ms = st->total_ms;
        st->total_points = tot_pts;
        st->total_attempts = tot_att;
        st->n_captured = n_cap;
        st->n_exhausted = n_ex;
        st->n_bins = ctx.nb;
        st->block_size = 0;
    }
    return HADROS_OK;
}

int or_trace_image(const hadros_camera* cam, const hadros_scene* s, const double* sig_e, const double* sig_s,
                   int n_sig, int nthreads, hadros_image_out* out, hadros_stats* st)
{
    if (!cam) return HADROS_ERR_ARG;
    return or_trace_image_range(cam, s, sig_e, sig_s, n_sig, nthreads, 0, int64_t(cam->nx) * cam->ny, out, st);
}

}  // extern "C"

// ---------------------------------------------------------------- MeV channel
// src/mev_neutrino_physics.cpp restated with the reference's operation order
// and libm calls (std::pow, std::exp, std::cbrt, std::tanh).
namespace {

const double MEV_HBAR_C = 1.973269804e-11;   // :13
const double MEV_M_E = 0.510998950;          // :14

double mev_clamp_ye(double Ye) { return clamp_(Ye, 0.0, 1.0); }                        // :17-20
double mev_safe_pow(double x, double p) { return std::pow(max_(x, 0.0), p); }           // :22-25
double mev_density_shape(double rho, double rho0)                                       // :26-33
{
    return clamp_(rho / max_(rho0, 1.0e-300), 0.0, 1.0);
}

struct MevP {
    const hadros_mev* m;
    double rho0, r0, r_min, r_max;   // TorusProfile accessors (src/torus_profile.cpp:12-42)
    bool has(unsigned bit) const { return (m->processes & bit) != 0; }
};

double mev_temperature(const MevP& P, double r, double th, double rho)   // :235-286
{
    const hadros_mev& m = *P.m;
    const double T0 = max_(m.T0_MeV, 1.0e-12);
    const double floor_ = max_(m.T_floor_MeV, 1.0e-12);
    const double shape = mev_density_shape(rho, P.rho0);
    const double radial = std::pow(max_(r / max_(P.r0, 1.0e-300), 1.0e-300), -m.T_power);
    const double equatorial = std::exp(-std::pow((th - 0.5 * PI) / 0.45, 2.0));
    const bool outside_torus = r > P.r_max || r < P.r_min;
    double T = T0;
    if (m.thermal_profile == HADROS_MEV_T_CONSTANT) {
        T = T0;
    } else if (m.thermal_profile == HADROS_MEV_T_INNER_HOT_TORUS) {
        T = floor_ + T0 * std::pow(shape, max_(m.T_power, 0.0));
    } else if (m.thermal_profile == HADROS_MEV_T_RADIAL_POWERLAW) {
        T = floor_ + T0 * radial * (0.35 + 0.65 * equatorial);
    } else if (m.thermal_profile == HADROS_MEV_T_TORUS_PLUS_COOL_ENVELOPE) {
        T = floor_ + T0 * std::pow(shape, max_(m.T_power, 0.0));
        if (outside_torus || shape < 1.0e-3) T = floor_ + 0.15 * T0 * radial;
    } else if (m.thermal_profile == HADROS_MEV_T_COLLAPSAR_INNER_HOT) {
        const double inner_hot = std::exp(-std::pow((r - P.r_min) / max_(1.15 * P.r0, 1.0e-6), 2.0));
        const double radial_cooling = std::pow(1.0 + max_(r - P.r_min, 0.0) / max_(P.r0, 1.0e-6), -0.75);
        T = floor_ + T0 * (0.30 + 0.70 * std::pow(shape, 0.18)) * (0.45 + 0.55 * equatorial) *
                         (0.55 + 0.45 * inner_hot) * radial_cooling;
    }
    return max_(T, floor_);
}

double mev_ye(const MevP& P, double r, double th, double rho)   // :288-336
{
    const hadros_mev& m = *P.m;
    const double shape = mev_density_shape(rho, P.rho0);
    const double polar = std::exp(-std::pow(th / 0.35, 2.0)) + std::exp(-std::pow((PI - th) / 0.35, 2.0));
    const bool envelope_like = r > P.r0 && (r > P.r_max || shape < 1.0e-2);
    double Ye = m.Ye_torus;
    if (m.ye_profile == HADROS_MEV_YE_CONSTANT) {
        Ye = m.Ye_torus;
    } else if (m.ye_profile == HADROS_MEV_YE_NEUTRON_RICH_TORUS) {
        Ye = m.Ye_torus + 0.05 * std::exp(-std::pow((r - P.r0) / 4.0, 2.0));
    } else if (m.ye_profile == HADROS_MEV_YE_FUNNEL_PROTON_RICH) {
        Ye = m.Ye_torus * (1.0 - clamp_(polar, 0.0, 1.0)) + m.Ye_funnel * clamp_(polar, 0.0, 1.0);
    } else if (m.ye_profile == HADROS_MEV_YE_TORUS_ENVELOPE_CONTRAST) {
        Ye = envelope_like ? m.Ye_envelope : m.Ye_torus;
        if (r < P.r_min && shape < 1.0e-2) Ye = m.Ye_funnel;
    } else if (m.ye_profile == HADROS_MEV_YE_COLLAPSAR_NEUTRON_RICH) {
        const double polar_weight = clamp_(polar, 0.0, 1.0);
        const double dense_weight = std::pow(shape, 0.25);
        const double radial_neutronization = std::exp(-std::pow((r - P.r0) / max_(1.2 * P.r0, 1.0e-6), 2.0));
        const double torus_ye = m.Ye_torus - 0.05 * dense_weight * radial_neutronization;
        Ye = torus_ye * (1.0 - polar_weight) + m.Ye_funnel * polar_weight;
    }
    return clamp_(Ye, max_(m.Ye_floor, 1.0e-6), min_(m.Ye_ceil, 1.0 - 1.0e-6));
}

double mev_thermal_shape(double E, double T)   // :338-358
{
    if (E <= 0.0 || T <= 0.0) return 0.0;
    const double x = E / T;
    if (x > 120.0) return 0.0;
    const double T3 = max_(T * T * T, 1.0e-300);
    return E * E / (std::exp(x) + 1.0) / T3;
}

double mev_fd_weight(double E, double T)   // :360-372
{
    if (E <= 0.0 || T <= 0.0) return 0.0;
    const double x = E / T;
    if (x > 120.0) return 0.0;
    return E * E / (std::exp(x) + 1.0);
}

double mev_ne(double rho, double Ye) { return rho <= 0.0 ? 0.0 : rho * mev_clamp_ye(Ye) / M_U_G; }   // :374-385

double mev_urca_deg(double rho, double Ye, double T, int flavor)   // :387-441
{
    if (flavor == HADROS_MEV_NU_X) return 1.0;
    const double ne = mev_ne(rho, Ye);
    const double pF = ne <= 0.0 ? 0.0 : MEV_HBAR_C * std::cbrt(3.0 * PI * PI * ne);
    const double mu = std::sqrt(pF * pF + MEV_M_E * MEV_M_E);
    const double eta = T <= 0.0 ? 0.0 : mu / T;
    if (!std::isfinite(eta)) return 1.0;
    const double shift = 0.25 * std::tanh((eta - 1.0) / 4.0);
    const double factor = flavor == HADROS_MEV_NU_E ? (1.0 + shift) : (1.0 - shift);
    return clamp_(factor, 0.5, 1.5);
}

double mev_emissivity(const MevP& P, double rho, double T, double Ye, double E)   // :443-593
{
    const hadros_mev& m = *P.m;
    if (m.model == HADROS_MEV_MODEL_TOY) {
        if (rho <= 0.0 || T <= 0.0 || E <= 0.0) return 0.0;
        const double nb = rho / M_U_G;
        const double neutron_fraction = clamp_(1.0 - Ye, 0.0, 1.0);
        const double proton_fraction = clamp_(Ye, 0.0, 1.0);
        const double ccw = neutron_fraction + 0.5 * proton_fraction;
        const double thermal_spectrum = E * E / (std::exp(min_(E / T, 120.0)) + 1.0);
        const double capture_like = nb * ccw * std::pow(T, 6.0);
        const double pair_like = 0.05 * nb * std::pow(T, 9.0);
        return m.norm * (capture_like + pair_like) * thermal_spectrum;
    }
    double total = 0.0;
    if (P.has(HADROS_MEV_URCA)) {
        double urca = 0.0;
        if (!(rho <= 0.0 || T <= 0.0 || E <= 0.0) && m.flavor != HADROS_MEV_NU_X) {
            const double ye = mev_clamp_ye(Ye);
            const double ff = m.flavor == HADROS_MEV_NU_E ? ye : (1.0 - ye);
            urca = 1.0e30 * mev_safe_pow(rho / 1.0e10, 1.0) * mev_safe_pow(T / 10.0, 6.0) * ff * mev_thermal_shape(E, T);
        }
        if (P.has(HADROS_MEV_DEGENERACY)) urca *= mev_urca_deg(rho, Ye, T, m.flavor);
        total += urca;
    }
    if (P.has(HADROS_MEV_PAIR)) {
        if (!(T <= 0.0 || E <= 0.0))
            total += 3.0e28 * mev_safe_pow(T / 10.0, 9.0) * (m.flavor == HADROS_MEV_NU_X ? 1.0 : 0.7) *
                     mev_thermal_shape(E, T);
        else
            total += 0.0;
    }
    if (P.has(HADROS_MEV_BREMS)) {
        if (!(rho <= 0.0 || T <= 0.0 || E <= 0.0))
            total += 1.0e27 * mev_safe_pow(rho / 1.0e10, 2.0) * mev_safe_pow(T / 10.0, 5.0) *
                     (m.flavor == HADROS_MEV_NU_X ? 1.0 : 0.8) * mev_thermal_shape(E, T);
        else
            total += 0.0;
    }
    return m.norm * total;
}

double mev_opacity(const MevP& P, double rho, double Ye, double E)   // :595-769
{
    const hadros_mev& m = *P.m;
    const bool absn = P.has(HADROS_MEV_ABSORPTION), scat = P.has(HADROS_MEV_SCATTERING);
    double absorption, scattering;
    if (m.model == HADROS_MEV_MODEL_TOY) {
        absorption = 0.0;
        if (!(!absn || rho <= 0.0 || E <= 0.0)) {
            const double nb = rho / M_U_G;
            const double target_fraction = clamp_(1.0 - Ye, 0.0, 1.0) + 0.25 * clamp_(Ye, 0.0, 1.0);
            absorption = nb * m.sigma_abs0_cm2 * E * E * target_fraction;
        }
        scattering = 0.0;
        if (!(!scat || rho <= 0.0 || E <= 0.0)) {
            const double nb = rho / M_U_G;
            const double composition_factor = 1.0 + 0.5 * clamp_(Ye, 0.0, 1.0);
            scattering = nb * m.sigma_scat0_cm2 * E * E * composition_factor;
        }
        return absorption + scattering;
    }
    const bool ok = rho > 0.0 && E > 0.0;
    const double nb = rho / M_U_G;
    double an = 0.0, ap = 0.0, sn = 0.0, sp = 0.0, se = 0.0;
    if (absn && P.has(HADROS_MEV_ABS_N) && m.flavor == HADROS_MEV_NU_E && ok)
        an = nb * (1.0 - mev_clamp_ye(Ye)) * m.sigma_abs0_cm2 * E * E;
    if (absn && P.has(HADROS_MEV_ABS_P) && m.flavor == HADROS_MEV_ANTI_NU_E && ok)
        ap = nb * mev_clamp_ye(Ye) * m.sigma_abs0_cm2 * E * E;
    if (scat && P.has(HADROS_MEV_SCAT_N) && ok) sn = nb * (1.0 - mev_clamp_ye(Ye)) * m.sigma_scat0_cm2 * E * E;
    if (scat && P.has(HADROS_MEV_SCAT_P) && ok) sp = nb * mev_clamp_ye(Ye) * 0.5 * m.sigma_scat0_cm2 * E * E;
    if (scat && P.has(HADROS_MEV_SCAT_E) && ok) se = mev_ne(rho, Ye) * m.sigma_scat_e0_cm2 * E * E;
    absorption = an + ap;
    scattering = sn + sp + se;
    return absorption + scattering;
}

// kappa and j at one point (src/radiative_transfer.cpp:541-600)
void mev_kappa_j(const MevP& P, double g, double rho, double T, double Ye, double& kappa, double& j)
{
    const hadros_mev& m = *P.m;
    kappa = 0.0;
    j = 0.0;
    if (m.spectral_mode == HADROS_MEV_SPECTRAL_FERMI_DIRAC_BAND) {
        const int n_bins = std::max(m.n_bins, 1);
        const double e_min = max_(m.E_min_MeV, 1.0e-6);
        const double e_max = max_(m.E_max_MeV, e_min);
        const double dE = (e_max - e_min) / static_cast<double>(n_bins);
        double weight_sum = 0.0, kappa_weighted = 0.0;
        for (int ib = 0; ib < n_bins; ++ib) {
            const double E_mid = e_min + (static_cast<double>(ib) + 0.5) * dE;
            const double E_local = E_mid / max_(g, 1.0e-300);
            const double w = mev_fd_weight(E_local, T) * dE;
            if (w <= 0.0 || !std::isfinite(w)) continue;
            const double local_kappa = mev_opacity(P, rho, Ye, E_local);
            j += mev_emissivity(P, rho, T, Ye, E_local) * dE;
            kappa_weighted += w * local_kappa;
            weight_sum += w;
        }
        kappa = weight_sum > 0.0 ? kappa_weighted / weight_sum : 0.0;
    } else {
        const double E = m.enu_obs_MeV / max_(g, 1.0e-300);
        kappa = mev_opacity(P, rho, Ye, E);
        j = mev_emissivity(P, rho, T, Ye, E);
    }
}

// The MeV accumulators of integrate_kerr_ray (:477-485, :500-640, :763-769).
struct MevRay {
    double tau = 0.0, I = 0.0, leak_sum = 0.0, leak_w = 0.0, r_ns = -1.0;
    void point(const MevP& P, const Medium& med, double r_h_rt, double rg_cm, double r, double th, double dl, double z)
    {
        if (r < r_h_rt + 1.0e-3) return;
        const double g = 1.0 / max_(z, 1.0e-300);
        const double dl_cm = dl * rg_cm;
        const double rho = med.rho(r, th);
        const double T = mev_temperature(P, r, th, rho);
        const double Ye = mev_ye(P, r, th, rho);
        double kappa, j;
        mev_kappa_j(P, g, rho, T, Ye, kappa, j);
        const double dtau = kappa * dl_cm;
        if (kappa > 0.0 && dtau > 1.0e-12) {
            const double att = std::exp(-min_(dtau, 700.0));
            I = I * att + std::pow(g, 3.0) * (j / kappa) * (1.0 - att);
        } else {
            I += std::pow(g, 3.0) * j * dl_cm;
        }
        leak_sum += j * dl_cm;
        leak_w += j * dl_cm;
        tau += dtau;
        if (r_ns < 0.0 && tau >= P.m->neutrinosphere_tau) r_ns = r;
    }
    void finish(double out[5]) const
    {
        out[0] = tau;
        out[1] = std::exp(-tau);
        out[2] = I;
        out[3] = r_ns;
        out[4] = leak_w > 0.0 ? leak_sum / leak_w : 1.0;
    }
};

MevP mev_params(const hadros_mev* m, const Medium& med) { return MevP{m, med.rho0, med.r0, med.rmin, med.rmax}; }

}  // namespace

extern "C" {

void or_mev_point(const hadros_scene* s, const hadros_mev* m, double r, double th, double g, double out[4])
{
    const Medium med(&s->medium);
    const MevP P = mev_params(m, med);
    const double rho = med.rho(r, th);
    const double T = mev_temperature(P, r, th, rho);
    const double Ye = mev_ye(P, r, th, rho);
    double kappa, j;
    mev_kappa_j(P, g, rho, T, Ye, kappa, j);
    out[0] = T;
    out[1] = Ye;
    out[2] = kappa;
    out[3] = j;
}

int or_rt_from_cache_mev(const hadros_cache_view* c, const hadros_scene* s, const double* sig_e, const double* sig_s,
                         int n_sig, const hadros_mev* m, int nthreads, double* mev_out)
{
    if (!c || !m || !mev_out) return HADROS_ERR_ARG;
    const int rc = check_scene(s, n_sig);
    if (rc) return rc;
    const RtCtx ctx(s, sig_e, sig_s, n_sig, c->a_bh);
    const MevP P = mev_params(m, ctx.med);
    // integrate_kerr_ray_spectral takes the MeV fields from its first integrated
    // bin (:741-760): none without one (or for an empty range)
    const bool have = ctx.mono || ctx.nb > 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic)
    for (int64_t k = 0; k < c->n_rays; ++k) {
        MevRay ray;
        if (have)
            for (int64_t q = c->offset[k]; q < c->offset[k + 1]; ++q)
                ray.point(P, ctx.med, ctx.r_h_rt, ctx.rg_cm, c->r[q], c->theta[q], c->dl[q], c->redshift[q]);
        ray.finish(mev_out + 5 * k);
    }
    return HADROS_OK;
}

int or_rt_from_cache(const hadros_cache_view* c, const hadros_scene* s, const double* sig_e, const double* sig_s,
                     int n_sig, int nthreads, hadros_image_out* out)
{
    if (!c || !out) return HADROS_ERR_ARG;
    const int rc = check_scene(s, n_sig);
    if (rc) return rc;
    RtCtx ctx(s, sig_e, sig_s, n_sig, c->a_bh);
    const int nbo = ctx.nb > 0 ? ctx.nb : 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic)
    for (int64_t k = 0; k < c->n_rays; ++k) {
        double tau[HADROS_MAX_BINS] = {0}, I[HADROS_MAX_BINS] = {0};
        double agg[3] = {0.0, 1.0, 0.0};
        const int64_t b = c->offset[k], e = c->offset[k + 1];
        double* bt = out->bin_tau ? out->bin_tau + k * nbo : nullptr;
        double* bp = out->bin_p ? out->bin_p + k * nbo : nullptr;
        double* bi = out->bin_i ? out->bin_i + k * nbo : nullptr;
        if (e > b) {
            for (int64_t q = b; q < e; ++q) ctx.point(c->r[q], c->theta[q], c->dl[q], c->redshift[q], tau, I);
            ctx.finish(tau, I, agg, bt, bp, bi);
        } else {
            // empty ray: results keep RTResult defaults (:611-618)
            for (int q = 0; q < nbo; ++q) {
                if (bt) bt[q] = 0.0;
                if (bp) bp[q] = 1.0;
                if (bi) bi[q] = 0.0;
            }
        }
        if (out->tau) out->tau[k] = agg[0];
        if (out->p_surv) out->p_surv[k] = agg[1];
        if (out->intensity) out->intensity[k] = agg[2];
    }
    return HADROS_OK;
}

// optical_depth::tau_along_ray (src/optical_depth.cpp:56-91) and
// survival_probability (:50-53): every point counts (no horizon cut), the
// local energy is E_inf times the point's redshift factor, points whose
// energy leaves the table are skipped, and dl is scaled by r_g of 3 M_sun
// whatever the scene's mass (:83-85).  out: tau, P per ray (either NULL).
int or_tau_along_ray(const hadros_cache_view* c, const hadros_medium* m, const double* sig_e, const double* sig_s,
                     int n_sig, double enu_inf, double* tau_out, double* p_out)
{
    if (!c || !m || n_sig < 2) return HADROS_ERR_ARG;
    const Medium med(m);
    const Sigma sig(sig_e, sig_s, n_sig);
    const double rg3 = G_CGS * 3.0 * MSUN_G / (C_CGS * C_CGS);   // constants::rg_cm(3.0)
    for (int64_t k = 0; k < c->n_rays; ++k) {
        double tau = 0.0;
        for (int64_t q = c->offset[k]; q < c->offset[k + 1]; ++q) {
            const double rho = med.rho(c->r[q], c->theta[q]);
            const double nb = rho / M_U_G;
            const double E = enu_inf * c->redshift[q];
            if (E < sig.emin() || E > sig.emax()) continue;
            const double sc = sig.eval(E);
            const double dl_cm = c->dl[q] * rg3;
            tau += nb * sc * dl_cm;
        }
        if (tau_out) tau_out[k] = tau;
        if (p_out) p_out[k] = std::exp(-tau);
    }
    return HADROS_OK;
}

}  // extern "C"

// apps/extract_opacity_surface.cpp:127-206 (grid :127-129, tau :165-174,
// crossing :176-190, classify_tau :55-68)
int or_opacity_surface(const hadros_medium* m, const double* sig_e, const double* sig_s, int n_sig,
                       double enu_GeV, double m_bh_msun, double tau_surface, int nr, int ntheta,
                       double* theta, double* r_tau, double* tau_total, double* cls, int* crossing)
{
    if (ntheta < 2 || nr < 4 || m->r_max <= m->r_min) return 1;
    if (enu_GeV < sig_e[0] || enu_GeV > sig_e[n_sig - 1]) return 2;
    const Medium med(m);
    const double sigma_cm2 = Sigma(sig_e, sig_s, n_sig).eval(enu_GeV);
    const double rg_cm = G_CGS * m_bh_msun * MSUN_G / (C_CGS * C_CGS);
    const double dr = (m->r_max - m->r_min) / static_cast<double>(nr - 1);
    std::vector<double> rho(static_cast<size_t>(nr)), tau(static_cast<size_t>(nr));
    for (int it = 0; it < ntheta; ++it) {
        const double th = PI * static_cast<double>(it) / static_cast<double>(ntheta - 1);
        for (int ir = 0; ir < nr; ++ir) rho[size_t(ir)] = med.rho(m->r_min + dr * static_cast<double>(ir), th);
        tau[size_t(nr - 1)] = 0.0;
        for (int ir = nr - 2; ir >= 0; --ir) {
            const double rho_mid = 0.5 * (rho[size_t(ir)] + rho[size_t(ir + 1)]);
            const double nb = rho_mid / M_U_G;
            const double dtau = nb * sigma_cm2 * dr * rg_cm;
            tau[size_t(ir)] = tau[size_t(ir + 1)] + dtau;
        }
        const double tt = tau[0];
        double rt = -1.0;
        int cf = 0;
        if (tt >= tau_surface) {
            for (int ir = 0; ir < nr - 1; ++ir) {
                const double t0 = tau[size_t(ir)], t1 = tau[size_t(ir + 1)];
                if (t0 >= tau_surface && t1 <= tau_surface) {
                    const double r0 = m->r_min + dr * static_cast<double>(ir);
                    const double frac = (tau_surface - t0) / std::max(t1 - t0, -1.0e-300);
                    rt = r0 + std::clamp(frac, 0.0, 1.0) * dr;
                    cf = 1;
                    break;
                }
            }
        }
        theta[it] = th;
        r_tau[it] = rt;
        tau_total[it] = tt;
        cls[it] = tt < 0.1 ? 0.0 : tt < 1.0 ? 1.0 : tt < 3.0 ? 2.0 : 3.0;
        crossing[it] = cf;
    }
    return 0;
}
