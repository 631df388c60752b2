// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// extern "C" entry points onto the reference's own C++ classes (compiled from
// the reference's sources by `make -C oracle ref`), used only by
// tests/golden/make_golden.py to write golden vectors and by CPU tests that
// compare this repo's oracle against the reference when /root/reference is
// present.  Nothing here restates the physics: every call goes to the
// reference's code.

#include "kerr_camera.hpp"
#include "kerr_geodesic.hpp"
#include "kerr_metric.hpp"
#include "radiative_transfer.hpp"
#include "sigma_table.hpp"
#include "torus_profile.hpp"

#include <cstring>
#include <string>

namespace {
GeodesicState to_state(const double* y)
{
    GeodesicState s;
    s.t = y[0]; s.r = y[1]; s.theta = y[2]; s.phi = y[3];
    s.pt = y[4]; s.pr = y[5]; s.ptheta = y[6]; s.pphi = y[7];
    return s;
}
void from_state(const GeodesicState& s, double* y)
{
    y[0] = s.t; y[1] = s.r; y[2] = s.theta; y[3] = s.phi;
    y[4] = s.pt; y[5] = s.pr; y[6] = s.ptheta; y[7] = s.pphi;
}
}  // namespace

extern "C" {

void ref_metric(double a, double r, double th, double* g16, double* gi16)
{
    KerrMetric m(a);
    double g[4][4], gi[4][4];
    m.metric(r, th, g);
    m.inverse_metric(r, th, gi);
    std::memcpy(g16, g, sizeof(g));
    std::memcpy(gi16, gi, sizeof(gi));
}

void ref_lapse_omega(double a, double r, double th, double* out3)
{
    KerrMetric m(a);
    out3[0] = m.lapse(r, th);
    out3[1] = m.omega_frame_drag(r, th);
    out3[2] = m.horizon_radius();
}

void ref_rhs(double a, const double* y, double* dy)
{
    KerrGeodesic g{KerrMetric(a)};
    from_state(g.rhs(to_state(y)), dy);
}

double ref_hamiltonian(double a, const double* y)
{
    KerrGeodesic g{KerrMetric(a)};
    return g.hamiltonian(to_state(y));
}

void ref_step_adaptive(double a, double h, double tol, double* y)
{
    KerrGeodesic g(KerrMetric(a), h, tol);
    GeodesicState s = to_state(y);
    g.step_adaptive(s);
    from_state(s, y);
}

// KerrCamera::trace_pixel; pts receives up to cap points of 7 doubles.
long ref_trace_pixel(double a, double r_obs, double th_rad, double fov_deg, int nx, int ny, double r_max,
                     double h, int i, int j, double* pts, long cap, int* captured, double* alpha_beta)
{
    KerrCamera cam(a, r_obs, th_rad, fov_deg, nx, ny, r_max, h);
    RayPath ray = cam.trace_pixel(i, j);
    const long n = long(ray.points.size());
    for (long k = 0; k < n && k < cap; ++k) {
        const PathPoint& p = ray.points[k];
        double* o = pts + 7 * k;
        o[0] = p.r_rg; o[1] = p.theta; o[2] = p.x_rg; o[3] = p.y_rg; o[4] = p.z_rg;
        o[5] = p.dl_rg; o[6] = p.redshift_factor;
    }
    *captured = ray.captured ? 1 : 0;
    alpha_beta[0] = ray.alpha_rg;
    alpha_beta[1] = ray.beta_rg;
    return n;
}

int ref_sigma(const char* path, const double* e, double* out, long n)
{
    try {
        SigmaTable t(path);
        for (long k = 0; k < n; ++k) {
            try {
                out[k] = t.sigma_cm2(e[k]);
            } catch (...) {
                out[k] = __builtin_nan("");
            }
        }
    } catch (...) {
        return 1;
    }
    return 0;
}

// torus params in TorusProfile constructor order (include/torus_profile.hpp:19-33)
static TorusProfile make_torus(const char* profile, const double* p)
{
    return TorusProfile(p[0], p[1], p[2], p[3], profile, p[4], p[5], p[6], p[7], p[8], p[9], p[10], p[11]);
}

int ref_rho(const char* profile, const double* tp, const double* r, const double* th, double* out, long n)
{
    try {
        TorusProfile t = make_torus(profile, tp);
        for (long k = 0; k < n; ++k) out[k] = t.rho(r[k], th[k]);
    } catch (...) {
        return 1;
    }
    return 0;
}

// source params in UHESourceParams order (include/radiative_transfer.hpp:47-64),
// without model/model_name: r_center, sigma_r, theta_width_rad, powerlaw,
// emax, norm, funnel_theta_rad, q, s, rho_ref, cutoff_min, cutoff_max,
// gradient_dr, gradient_dtheta_rad.  spectral: gamma, ecut, e_min, e_max, n_bins.
static void fill_source(UHESourceParams& s, const char* model, const double* sp)
{
    s.model_name = model;
    s.model = radiative_transfer::parse_uhe_source_model(model);
    s.r_center_rg = sp[0]; s.sigma_r_rg = sp[1]; s.theta_width_rad = sp[2]; s.powerlaw = sp[3];
    s.emax_GeV = sp[4]; s.norm = sp[5]; s.funnel_theta_rad = sp[6]; s.density_power_q = sp[7];
    s.radial_power_s = sp[8]; s.rho_ref_gcm3 = sp[9]; s.cutoff_min = sp[10]; s.cutoff_max = sp[11];
    s.gradient_dr_rg = sp[12]; s.gradient_dtheta_rad = sp[13];
}
static void fill_spectral(UHESpectralParams& s, const char* model, const double* sp)
{
    s.model_name = model;
    s.model = radiative_transfer::parse_uhe_spectral_model(model);
    s.gamma = sp[0]; s.ecut_GeV = sp[1]; s.e_min_GeV = sp[2]; s.e_max_GeV = sp[3]; s.n_bins = int(sp[4]);
}

int ref_emissivity(const char* profile, const double* tp, const char* source, const double* sp,
                   const char* spectral, const double* spp, const double* r, const double* th, const double* e,
                   double* out_j, double* out_w, double* out_s, long n)
{
    try {
        TorusProfile t = make_torus(profile, tp);
        UHESourceParams so;
        fill_source(so, source, sp);
        UHESpectralParams spec;
        fill_spectral(spec, spectral, spp);
        for (long k = 0; k < n; ++k) {
            out_j[k] = radiative_transfer::emissivity_uhe(r[k], th[k], e[k], t, so, spec);
            out_w[k] = radiative_transfer::uhe_source_spatial_weight(r[k], th[k], t, so);
            out_s[k] = radiative_transfer::uhe_spectral_weight(e[k], so, spec);
        }
    } catch (...) {
        return 1;
    }
    return 0;
}

// integrate_kerr_ray_spectral over one ray of 7-double points; out: tau, P, I
// and the MeV fields tau_mev, P_mev, I_mev, r_ns, leakage.
int ref_integrate(const char* sigma_path, const char* profile, const double* tp, const char* source,
                  const double* sp, const char* spectral, const double* spp, double enu, double mbh, double a_bh,
                  const double* pts, long n, double* out8)
{
    try {
        SigmaTable sig(sigma_path);
        TorusProfile t = make_torus(profile, tp);
        UHESourceParams so;
        fill_source(so, source, sp);
        UHESpectralParams spec;
        fill_spectral(spec, spectral, spp);
        RayPath ray;
        ray.a_bh = a_bh;
        ray.points.resize(size_t(n));
        for (long k = 0; k < n; ++k) {
            const double* o = pts + 7 * k;
            PathPoint& p = ray.points[k];
            p.r_rg = o[0]; p.theta = o[1]; p.x_rg = o[2]; p.y_rg = o[3]; p.z_rg = o[4];
            p.dl_rg = o[5]; p.redshift_factor = o[6];
        }
        MeVThermalParams mev;
        RTResult rt = radiative_transfer::integrate_kerr_ray_spectral(ray, enu, mbh, t, sig, so, mev, spec);
        out8[0] = rt.tau; out8[1] = rt.P_surv; out8[2] = rt.I_obs; out8[3] = rt.tau_mev;
        out8[4] = rt.P_surv_mev; out8[5] = rt.I_obs_mev; out8[6] = rt.r_neutrinosphere_rg;
        out8[7] = rt.leakage_factor;
    } catch (...) {
        return 1;
    }
    return 0;
}

}  // extern "C"
