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
#include "mev_neutrino_physics.hpp"
#include "optical_depth.hpp"
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

void ref_step_rk4(double a, double h, double* y)
{
    KerrGeodesic g(KerrMetric(a), h);
    GeodesicState s = to_state(y);
    g.step_rk4(s);
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

// optical_depth::tau_along_ray and survival_probability on rays given as SoA
// (r, theta, dl, redshift) ranges [off[k], off[k+1]); out2 = tau, P per ray.
int ref_tau_along_ray(const char* sigma_path, const char* profile, const double* tp, double enu,
                      const double* r, const double* th, const double* dl, const double* z, const long* off,
                      long n_rays, double* out2)
{
    try {
        SigmaTable sig(sigma_path);
        TorusProfile t = make_torus(profile, tp);
        for (long k = 0; k < n_rays; ++k) {
            RayPath ray;
            for (long q = off[k]; q < off[k + 1]; ++q) {
                PathPoint p{};
                p.r_rg = r[q]; p.theta = th[q]; p.dl_rg = dl[q]; p.redshift_factor = z[q];
                ray.points.push_back(p);
            }
            const double tau = optical_depth::tau_along_ray(ray, enu, t, sig);
            out2[2 * k] = tau;
            out2[2 * k + 1] = optical_depth::survival_probability(tau);
        }
    } catch (...) {
        return 1;
    }
    return 0;
}

// MeV parameters: names of model, flavor, thermal profile, Ye profile and
// spectral mode, then mp = enu, norm, sigma_abs0, sigma_scat0, sigma_scat_e0,
// neutrinosphere_tau, T0, T_floor, T_power, Ye_torus, Ye_funnel, Ye_envelope,
// Ye_floor, Ye_ceil, E_min, E_max, n_bins, process bits (HADROS_MEV_* order:
// urca, pair, brems, absorption, scattering, degeneracy, abs_n, abs_p,
// scat_n, scat_p, scat_e).
static void fill_mev(MeVThermalParams& m, const char* const* names, const double* mp)
{
    m.model_name = names[0];
    m.model = mev_neutrino::parse_mev_model(m.model_name);
    m.flavor_name = names[1];
    m.flavor = mev_neutrino::parse_mev_flavor(m.flavor_name);
    m.thermal_profile_name = names[2];
    m.thermal_profile = mev_neutrino::parse_mev_thermal_profile(m.thermal_profile_name);
    m.ye_profile_name = names[3];
    m.ye_profile = mev_neutrino::parse_mev_ye_profile(m.ye_profile_name);
    m.spectral_mode_name = names[4];
    m.spectral_mode = mev_neutrino::parse_mev_spectral_mode(m.spectral_mode_name);
    m.Enu_obs_MeV = mp[0]; m.norm = mp[1]; m.sigma_abs0_cm2 = mp[2]; m.sigma_scat0_cm2 = mp[3];
    m.neutrinosphere_tau = mp[5]; m.T0_MeV = mp[6]; m.T_floor_MeV = mp[7]; m.T_power = mp[8];
    m.Ye_torus = mp[9]; m.Ye_funnel = mp[10]; m.Ye_envelope = mp[11]; m.Ye_floor = mp[12]; m.Ye_ceil = mp[13];
    m.E_min_MeV = mp[14]; m.E_max_MeV = mp[15]; m.n_bins = int(mp[16]);
    const unsigned bits = unsigned(mp[17]);
    m.include_urca = bits & 1; m.include_pair = bits & 2; m.include_brems = bits & 4;
    m.include_absorption = bits & 8; m.include_scattering = bits & 16; m.use_degeneracy_correction = bits & 32;
    m.include_abs_n = bits & 64; m.include_abs_p = bits & 128; m.include_scat_n = bits & 256;
    m.include_scat_p = bits & 512; m.include_scat_e = bits & 1024;
}

// the same parameters as the physics functions take them (field copy)
static mev_neutrino::MeVPhysicsParams physics_params(const MeVThermalParams& m, double sigma_scat_e0)
{
    mev_neutrino::MeVPhysicsParams p;
    p.model = m.model; p.flavor = m.flavor; p.include_urca = m.include_urca; p.include_pair = m.include_pair;
    p.include_brems = m.include_brems; p.include_absorption = m.include_absorption;
    p.include_scattering = m.include_scattering; p.use_degeneracy_correction = m.use_degeneracy_correction;
    p.include_abs_n = m.include_abs_n; p.include_abs_p = m.include_abs_p; p.include_scat_n = m.include_scat_n;
    p.include_scat_p = m.include_scat_p; p.include_scat_e = m.include_scat_e; p.norm = m.norm;
    p.sigma_abs0_cm2 = m.sigma_abs0_cm2; p.sigma_scat0_cm2 = m.sigma_scat0_cm2; p.sigma_scat_e0_cm2 = sigma_scat_e0;
    p.thermal_profile = m.thermal_profile; p.ye_profile = m.ye_profile; p.spectral_mode = m.spectral_mode;
    p.T0_MeV = m.T0_MeV; p.T_floor_MeV = m.T_floor_MeV; p.T_power = m.T_power; p.Ye_torus = m.Ye_torus;
    p.Ye_funnel = m.Ye_funnel; p.Ye_envelope = m.Ye_envelope; p.Ye_floor = m.Ye_floor; p.Ye_ceil = m.Ye_ceil;
    p.E_min_MeV = m.E_min_MeV; p.E_max_MeV = m.E_max_MeV; p.n_bins = m.n_bins;
    return p;
}

// per point: T, Ye, total opacity and total emissivity at local energy E[k]
int ref_mev_point(const char* profile, const double* tp, const char* const* names, const double* mp,
                  const double* r, const double* th, const double* E, double* out4, long n)
{
    try {
        TorusProfile t = make_torus(profile, tp);
        MeVThermalParams m;
        fill_mev(m, names, mp);
        const mev_neutrino::MeVPhysicsParams P = physics_params(m, mp[4]);
        for (long k = 0; k < n; ++k) {
            const double rho = t.rho(r[k], th[k]);
            const double T = mev_neutrino::mev_temperature_profile_MeV(r[k], th[k], rho, t.rho0_gcm3(), t.r0_rg(),
                                                                        t.r_min_rg(), t.r_max_rg(), P);
            const double Ye = mev_neutrino::mev_ye_profile(r[k], th[k], rho, t.rho0_gcm3(), t.r0_rg(), t.r_min_rg(),
                                                           t.r_max_rg(), P);
            out4[4 * k] = T;
            out4[4 * k + 1] = Ye;
            out4[4 * k + 2] = mev_neutrino::mev_total_opacity_cm_inv(rho, Ye, E[k], P);
            out4[4 * k + 3] = mev_neutrino::mev_total_emissivity(rho, T, Ye, E[k], P);
        }
    } catch (...) {
        return 1;
    }
    return 0;
}

// integrate_kerr_ray_spectral with MeV parameters; out8 as ref_integrate
int ref_integrate_mev(const char* sigma_path, const char* profile, const double* tp, const char* source,
                      const double* sp, const char* spectral, const double* spp, double enu, double mbh, double a_bh,
                      const char* const* mev_names, const double* mp, const double* pts, long n, double* out8)
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
        fill_mev(mev, mev_names, mp);
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
