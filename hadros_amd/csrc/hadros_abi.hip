// hadros_abi.hip -- host side of the C-ABI declared in include/hadros.h.
//
// Parameter normalisation (the reference's constructor clamps and per-scene
// constants) happens here on the host, with the host libm, so that the values
// every lane starts from are the ones the reference computes; the per-ray
// work runs in the kernels of tracer.hip / replay.hip.
#include "kernels.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <mutex>
#include <sstream>
#include <utility>
#include <string>
#include <vector>

using namespace hd;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPC(expr)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return fail(HADROS_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

constexpr double C_CGS = 2.99792458e10;   // include/constants.hpp:10-12
constexpr double G_CGS = 6.67430e-8;
constexpr double MSUN_G = 1.98847e33;

int g_device_ok = -1;   // -1 unknown, 0 missing, 1 ok
std::string g_device_msg;

int ensure_device()
{
    if (g_device_ok == 1) return HADROS_OK;
    if (g_device_ok == 0) return fail(HADROS_ERR_DEVICE, g_device_msg);
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        g_device_ok = 0;
        g_device_msg = "no HIP device available (hadros requires an MI355X / gfx950 GPU; there is no CPU path)";
        return fail(HADROS_ERR_DEVICE, g_device_msg);
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        g_device_ok = 0;
        g_device_msg = std::string("device is not gfx950: ") + prop.gcnArchName;
        return fail(HADROS_ERR_DEVICE, g_device_msg);
    }
    g_device_ok = 1;
    return HADROS_OK;
}

// ---------------------------------------------------------------- params
DevMedium make_medium(const hadros_medium& m)
{
    DevMedium d{};
    d.kind = m.kind;
    d.powerlaw_disk = (m.kind == HADROS_PROFILE_POWERLAW || m.kind == HADROS_PROFILE_POWERLAW_FUNNEL ||
                       m.kind == HADROS_PROFILE_POWERLAW_ENVELOPE || m.kind == HADROS_PROFILE_POWERLAW_FUNNEL_ENVELOPE);
    d.funnel = (m.kind == HADROS_PROFILE_GAUSSIAN_FUNNEL || m.kind == HADROS_PROFILE_POWERLAW_FUNNEL ||
                m.kind == HADROS_PROFILE_POWERLAW_FUNNEL_ENVELOPE);
    d.envelope = (m.kind == HADROS_PROFILE_GAUSSIAN_ENVELOPE || m.kind == HADROS_PROFILE_POWERLAW_ENVELOPE ||
                  m.kind == HADROS_PROFILE_POWERLAW_FUNNEL_ENVELOPE);
    // TorusProfile constructor clamps (src/torus_profile.cpp:12-42)
    d.rho0 = m.rho0;
    d.r0 = m.r0;
    d.sr = m.sigma_r;
    d.H = m.h_over_r;
    d.p = m.radial_power;
    d.fd = dclamp(m.funnel_depletion, 0.0, 1.0);
    d.ft = dmax(m.funnel_theta_rad, 1.0e-6);
    d.env0 = dmax(m.envelope_rho0, 0.0);
    d.ealpha = m.envelope_alpha;
    d.rmin = m.r_min;
    d.rmax = dmax(m.r_max, m.r_min);
    d.floor_ = dmax(m.rho_floor, 0.0);
    d.inv_r0 = 1.0 / d.r0;
    d.inv_sr = 1.0 / d.sr;
    d.inv_H = 1.0 / d.H;
    d.inv_rmax = 1.0 / d.rmax;
    d.inv_ft = 1.0 / d.ft;
    return d;
}

int check_kinds(const hadros_scene& s)
{
    if (s.medium.kind < 0 || s.medium.kind > 7) return fail(HADROS_ERR_KIND, "unknown density profile kind");
    if (s.source.kind < 0 || s.source.kind > 4) return fail(HADROS_ERR_KIND, "unknown UHE source kind");
    if (s.spectrum.kind < 0 || s.spectrum.kind > 2) return fail(HADROS_ERR_KIND, "unknown UHE spectral kind");
    return HADROS_OK;
}

// uhe_spectral_weight (src/radiative_transfer.cpp:172-207), host side for the bin weights
double spectral_weight_host(const hadros_scene& s, double E)
{
    if (E <= 0.0) return 0.0;
    const double safe_E = dmax(E, 1.0e-300);
    if (s.spectrum.kind == HADROS_SPECTRUM_MONOCHROMATIC) {
        if (s.source.emax_GeV <= 0.0) return 0.0;
        return std::pow(safe_E, -s.source.powerlaw) * std::exp(-safe_E / s.source.emax_GeV);
    }
    if (s.spectrum.kind == HADROS_SPECTRUM_POWERLAW) return std::pow(safe_E, -s.spectrum.gamma);
    if (s.spectrum.kind == HADROS_SPECTRUM_POWERLAW_CUTOFF) {
        if (s.spectrum.ecut_GeV <= 0.0) return 0.0;
        return std::pow(safe_E, -s.spectrum.gamma) * std::exp(-safe_E / s.spectrum.ecut_GeV);
    }
    return 0.0;
}

int check_sigma(const double* e, const double* s, int n)
{
    if (!e || !s) return fail(HADROS_ERR_ARG, "sigma table pointer is NULL");
    if (n < 2) return fail(HADROS_ERR_TABLE, "Sigma table must contain at least two valid data lines.");
    if (n > HADROS_MAX_SIGMA) return fail(HADROS_ERR_ARG, "sigma table longer than HADROS_MAX_SIGMA");
    for (int k = 0; k < n; ++k)
        if (!(e[k] > 0.0) || !(s[k] > 0.0)) return fail(HADROS_ERR_TABLE, "sigma table has non-positive entries");
    for (int k = 1; k < n; ++k)
        if (e[k] <= e[k - 1]) return fail(HADROS_ERR_TABLE, "Energy grid in sigma table must be strictly increasing.");
    return HADROS_OK;
}

// Index guess of the device sigma lookup: a table whose every ln E node lies
// within d/4 of the uniform grid lo0 + k d gets inv_d = 1/d (the guess
// ceil((ln E - lo0)/d) is then at most one node from the lower bound);
// other tables get inv_d = 0 (binary search).
// exact = 1 when every node lies within 1e-12 d of that grid (log-spaced
// tables): the lookup then takes the guess as the interval, with no compare.
void sigma_guess(const double* se, int n, double& lo0, double& inv_d, int* exact = nullptr)
{
    const double l0 = std::log(se[0]), l1 = std::log(se[n - 1]);
    const double d = (l1 - l0) / double(n - 1);
    bool uniform = d > 0.0;
    double dev = 0.0;
    for (int k = 0; uniform && k < n; ++k) {
        const double e = std::fabs(std::log(se[k]) - (l0 + d * double(k)));
        uniform = e <= 0.25 * d;
        dev = std::max(dev, e);
    }
    lo0 = l0;
    inv_d = uniform ? 1.0 / d : 0.0;
    if (exact) *exact = uniform && dev <= 1.0e-12 * d;
}

// A table whose every node lies within 1e-13 (in ln sigma) of the line through
// its end nodes in ln E - ln sigma is one power law sigma_0 (E / E_0)^s, and its
// log-log interpolation is that power law: sigma(E z) = sigma(E) z^s.  The
// transfer kernel then takes z^s once per path point and one product per
// (point, bin) instead of a table lookup and an exponential.
void sigma_powerlaw(const double* se, const double* ss, int n, int& pow_ok, double& s)
{
    const double l0 = std::log(se[0]), s0 = std::log(ss[0]);
    s = (std::log(ss[n - 1]) - s0) / (std::log(se[n - 1]) - l0);
    bool ok = std::isfinite(s);
    for (int k = 0; ok && k < n; ++k) ok = std::fabs(std::log(ss[k]) - (s0 + s * (std::log(se[k]) - l0))) <= 1.0e-13;
    pow_ok = ok;
}

// sigma at E_mid of every bin of a power-law table (sigma_powerlaw)
void sigma_bins(DevRt& P, const double* se, const double* ss)
{
    if (!P.sig_pow) return;
    const double l0 = std::log(se[0]), s0 = std::log(ss[0]);
    for (int i = 0; i < P.nb; ++i) P.sigb[i] = std::exp(s0 + P.sig_s * (P.lnemid[i] - l0));
}

// DevRt of one scene: integrate_kerr_ray_spectral's bins (src/radiative_transfer.cpp:687-761),
// spectral factor, r_g, horizon of the transfer, sigma range.
int make_rt(const hadros_scene& s, const double* se, const double* ss, int n, double a_bh, DevRt& P)
{
    int rc = check_kinds(s);
    if (rc) return rc;
    rc = check_sigma(se, ss, n);
    if (rc) return rc;
    P = DevRt{};
    P.med = make_medium(s.medium);
    const hadros_source& so = s.source;
    P.src.kind = so.kind;
    P.src.ring_ok = !(so.r_center <= 0.0 || so.sigma_r <= 0.0 || so.theta_width_rad <= 0.0 || so.emax_GeV <= 0.0);
    P.src.rc = so.r_center;
    P.src.sr = so.sigma_r;
    P.src.tw = so.theta_width_rad;
    P.src.norm = so.norm;
    P.src.ft = so.funnel_theta_rad;
    P.src.q = so.density_q;
    P.src.s = so.radial_s;
    P.src.rho_ref = so.rho_ref;
    P.src.cmin = so.cutoff_min;
    P.src.cmax = so.cutoff_max;
    // the shock layer's finite-difference steps, clamped here as the
    // source term clamps them (dmax(dr, 1e-4), dmax(dtheta, 1e-5)): the
    // kernels read them as they are
    P.src.gdr = dmax(so.gradient_dr, 1.0e-4);
    P.src.gdth = dmax(so.gradient_dtheta_rad, 1.0e-5);
    P.src.inv_rc = 1.0 / so.r_center;
    P.src.inv_sr = 1.0 / so.sigma_r;
    P.src.inv_tw = 1.0 / so.theta_width_rad;
    P.src.rho_norm_c = rho(P.med, so.r_center, 0.5 * hd::PI);
    P.n_sigma = n;
    P.e_lo = se[0];
    P.e_hi = se[n - 1];
    sigma_guess(se, n, P.sig_lo, P.sig_inv_d, &P.sig_exact);
    sigma_powerlaw(se, ss, n, P.sig_pow, P.sig_s);
    P.rg_cm = G_CGS * s.m_bh_msun * MSUN_G / (C_CGS * C_CGS);
    const double a = dclamp(a_bh, 0.0, 0.999999);
    P.r_h_rt = 1.0 + std::sqrt(1.0 - a * a);
    const hadros_spectrum& sp = s.spectrum;
    P.mono = sp.kind == HADROS_SPECTRUM_MONOCHROMATIC;
    if (P.mono) {
        P.gam = so.powerlaw;
        P.spec_ok = so.emax_GeV > 0.0;
        P.invcut = P.spec_ok ? 1.0 / so.emax_GeV : 0.0;
        P.nb = 1;
        P.emid[0] = s.enu_obs_GeV;
        P.lnemid[0] = std::log(s.enu_obs_GeV);
        P.de[0] = 1.0;
        P.w[0] = 1.0;
        sigma_bins(P, se, ss);
        return HADROS_OK;
    }
    P.gam = sp.gamma;
    if (sp.kind == HADROS_SPECTRUM_POWERLAW) {
        P.spec_ok = 1;
        P.invcut = 0.0;
    } else {
        P.spec_ok = sp.ecut_GeV > 0.0;
        P.invcut = P.spec_ok ? 1.0 / sp.ecut_GeV : 0.0;
    }
    const int n_bins = std::max(sp.n_bins, 1);
    if (n_bins > HADROS_MAX_BINS) return fail(HADROS_ERR_ARG, "n_bins exceeds HADROS_MAX_BINS (16)");
    const double e_min = dmax(sp.e_min_GeV, se[0]);
    const double e_max = dmin(sp.e_max_GeV, se[n - 1]);
    P.nb = 0;
    if (e_min <= 0.0 || e_max <= e_min) return HADROS_OK;   // RTResult{} for every ray
    const double log_min = std::log(e_min);
    const double log_max = std::log(e_max);
    const double dlog = (log_max - log_min) / static_cast<double>(n_bins);
    for (int i = 0; i < n_bins; ++i) {
        const double log_left = log_min + dlog * static_cast<double>(i);
        const double log_right = log_left + dlog;
        const double E_mid = std::exp(0.5 * (log_left + log_right));
        const double dE = std::exp(log_right) - std::exp(log_left);
        const double w = spectral_weight_host(s, E_mid) * dE;
        if (w <= 0.0 || !std::isfinite(w)) continue;
        P.emid[P.nb] = E_mid;
        P.lnemid[P.nb] = std::log(E_mid);
        P.de[P.nb] = dE;
        P.w[P.nb] = w;
        ++P.nb;
    }
    sigma_bins(P, se, ss);
    return HADROS_OK;
}

// Entry k of the device sigma table (3 n doubles): ln E, then ln sigma, then
// the log-log slope of each interval ending at node m (0 for m = 0).
double sigma_table_entry(const double* se, const double* ss, int n, int k)
{
    if (k < n) return std::log(se[k]);
    if (k < 2 * n) return std::log(ss[k - n]);
    const int m = k - 2 * n;
    if (m == 0) return 0.0;
    return (std::log(ss[m]) - std::log(ss[m - 1])) / (std::log(se[m]) - std::log(se[m - 1]));
}

void sigma_tables(const double* se, const double* ss, int n, std::vector<double>& d, std::vector<float>& f)
{
    d.resize(3 * size_t(n));
    f.resize(3 * size_t(n));
    for (int k = 0; k < 3 * n; ++k) {
        d[size_t(k)] = sigma_table_entry(se, ss, n, k);
        f[size_t(k)] = float(d[size_t(k)]);
    }
}

// KerrMetric on the host (src/kerr_metric.cpp:10-144) for the observer constants.
void host_metric(double a, double r, double th, double g[5], double& alpha, double& omega)
{
    const double c = std::cos(th);
    const double sig = r * r + a * a * c * c;
    const double del = r * r - 2.0 * r + a * a;
    const double s = std::sin(th);
    const double s2 = s * s;
    const double bigA = (r * r + a * a) * (r * r + a * a) - a * a * del * s2;
    g[0] = -(1.0 - 2.0 * r / sig);
    g[1] = -2.0 * a * r * s2 / sig;
    g[2] = sig / del;
    g[3] = sig;
    g[4] = (r * r + a * a + 2.0 * a * a * r * s2 / sig) * s2;
    alpha = std::sqrt(sig * del / bigA);
    omega = 2.0 * a * r / bigA;
}

int make_cam(const hadros_camera& c, DevCam& d)
{
    if (c.nx <= 0 || c.ny <= 0) return fail(HADROS_ERR_ARG, "camera nx, ny must be positive");
    if (c.mode != HADROS_MODE_REF && c.mode != HADROS_MODE_CFG) return fail(HADROS_ERR_ARG, "unknown integrator mode");
    if (!(c.h > 0.0) || !(c.tol > 0.0) || c.max_steps <= 0)
        return fail(HADROS_ERR_ARG, "camera h, tol and max_steps must be positive");
    d = DevCam{};
    d.a = c.a_spin;
    d.r_obs = c.r_obs;
    d.th_obs = c.theta_obs_rad;
    const double fov_rad = c.fov_deg * hd::PI / 180.0;   // src/kerr_camera.cpp:78-81
    d.tanhalf = std::tan(0.5 * fov_rad);
    d.r_max = c.r_max;
    d.h = c.h;
    d.tol = c.tol;
    d.nx = c.nx;
    d.ny = c.ny;
    d.mode = c.mode;
    d.max_steps = c.max_steps;
    double g[5];
    host_metric(c.a_spin, c.r_obs, c.theta_obs_rad, g, d.alpha, d.omega);
    d.g_tt = g[0];
    d.g_tph = g[1];
    d.g_rr = g[2];
    d.g_thth = g[3];
    d.g_phph = g[4];
    d.r_h = 1.0 + std::sqrt(1.0 - c.a_spin * c.a_spin);
    d.sub_stride = 1;
    d.sub_offset = 0;
    d.sub_nx = c.nx;
    d.sub_ny = c.ny;
    return HADROS_OK;
}

template <class T>
int dalloc(T** p, size_t n)
{
    *p = nullptr;
    if (n == 0) return HADROS_OK;
    HIPC(hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)));
    return HADROS_OK;
}

template <class T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

// error reporting and device check shared with the other host units
int hd::set_error(int code, const char* msg) { return fail(code, msg); }
int hd::device_ready() { return ensure_device(); }

// ================================================================ plans
// MeV channel of a plan: device parameters and the five per-ray output columns
struct MevState {
    bool on = false;
    long n = 0;
    DevMev* d_mev = nullptr;
    double* d_out[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // tau, P, I, r_ns, leakage
};

namespace {

int check_mev(const hadros_mev& m)
{
    if (m.model < 0 || m.model > 1) return fail(HADROS_ERR_KIND, "unknown MeV model");
    if (m.flavor < 0 || m.flavor > 2) return fail(HADROS_ERR_KIND, "unknown MeV flavor");
    if (m.thermal_profile < 0 || m.thermal_profile > 4) return fail(HADROS_ERR_KIND, "unknown MeV thermal profile");
    if (m.ye_profile < 0 || m.ye_profile > 4) return fail(HADROS_ERR_KIND, "unknown MeV Ye profile");
    if (m.spectral_mode < 0 || m.spectral_mode > 1) return fail(HADROS_ERR_KIND, "unknown MeV spectral mode");
    return HADROS_OK;
}

// make_mev_physics_params (src/radiative_transfer.cpp:14-53) plus the
// TorusProfile accessors the profiles read (src/torus_profile.cpp:12-42)
DevMev make_mev(const hadros_mev& m, const DevMedium& med)
{
    DevMev d{};
    d.model = m.model;
    d.flavor = m.flavor;
    d.thermal = m.thermal_profile;
    d.ye = m.ye_profile;
    d.spectral = m.spectral_mode;
    d.n_bins = m.n_bins;
    d.proc = m.processes;
    d.enu = m.enu_obs_MeV;
    d.norm = m.norm;
    d.sabs = m.sigma_abs0_cm2;
    d.sscat = m.sigma_scat0_cm2;
    d.sscat_e = m.sigma_scat_e0_cm2;
    d.ns_tau = m.neutrinosphere_tau;
    d.T0 = m.T0_MeV;
    d.T_floor = m.T_floor_MeV;
    d.T_power = m.T_power;
    d.Ye_torus = m.Ye_torus;
    d.Ye_funnel = m.Ye_funnel;
    d.Ye_envelope = m.Ye_envelope;
    d.Ye_floor = m.Ye_floor;
    d.Ye_ceil = m.Ye_ceil;
    d.E_min = m.E_min_MeV;
    d.E_max = m.E_max_MeV;
    d.rho0 = med.rho0;
    d.r0 = med.r0;
    d.r_min = med.rmin;
    d.r_max = med.rmax;
    return d;
}

void mev_free(MevState& M)
{
    dfree(M.d_mev);
    for (auto& p : M.d_out) dfree(p);
    M.on = false;
}

int mev_set(MevState& M, const hadros_mev* m, const DevMedium& med, long n)
{
    if (!m || !m->enabled) {
        M.on = false;
        return HADROS_OK;
    }
    int rc = check_mev(*m);
    if (rc) return rc;
    const DevMev d = make_mev(*m, med);
    if (!M.d_mev && (rc = dalloc(&M.d_mev, 1))) return rc;
    for (auto& p : M.d_out)
        if (!p && (rc = dalloc(&p, size_t(std::max(n, 1L))))) return rc;
    HIPC(hipMemcpy(M.d_mev, &d, sizeof(DevMev), hipMemcpyHostToDevice));
    M.n = n;
    M.on = true;
    return HADROS_OK;
}

void mev_apply(const MevState& M, ReplayArgs& R)
{
    R.mev = M.on ? M.d_mev : nullptr;
    R.tau_m = M.on ? M.d_out[0] : nullptr;
    R.p_m = M.on ? M.d_out[1] : nullptr;
    R.I_m = M.on ? M.d_out[2] : nullptr;
    R.rns_m = M.on ? M.d_out[3] : nullptr;
    R.leak_m = M.on ? M.d_out[4] : nullptr;
}

// the last run's columns, or the reference's "not computed" values
int mev_fetch(const MevState& M, bool ran_with_mev, long n, hadros_mev_out* o)
{
    double* dst[5] = {o->tau, o->p_surv, o->intensity, o->r_neutrinosphere, o->leakage};
    const double dflt[5] = {0.0, 1.0, 0.0, -1.0, 1.0};
    for (int c = 0; c < 5; ++c) {
        if (!dst[c] || n <= 0) continue;
        if (ran_with_mev) HIPC(hipMemcpy(dst[c], M.d_out[c], size_t(n) * sizeof(double), hipMemcpyDeviceToHost));
        else std::fill(dst[c], dst[c] + n, dflt[c]);
    }
    return HADROS_OK;
}

}  // namespace

struct hadros_tracer {
    TraceArgs args{};
    int precision = 0;
    int block = 0;
    long npix = 0;
    // split form (geodesic kernel + transfer over the chunk pool)
    int path = HADROS_PATH_AUTO;
    bool split = false;        // effective choice of the last run
    PointsArgs pargs{};
    ReplayArgs rargs{};
    long capacity = 0;         // chunks
    double* d_pool = nullptr;
    int *d_next = nullptr, *d_nst = nullptr;
    unsigned long long* d_ctr = nullptr;   // the pool's chunk counter: d_cnt + 4
    double *d_sig = nullptr;
    float* d_sigf = nullptr;
    double *d_tau = nullptr, *d_p = nullptr, *d_I = nullptr, *d_alpha = nullptr, *d_beta = nullptr;
    double *d_bt = nullptr, *d_bp = nullptr, *d_bi = nullptr;
    int *d_cap = nullptr, *d_np = nullptr, *d_flags = nullptr;
    unsigned long long* d_cnt = nullptr;   // [5]: 4 run counters, then the chunk counter
    DevRt* d_rt = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // split form over segments of tile rows: segment 0 runs on the caller's
    // stream, segment i > 0 on seg_s[i]; each segment's transfer follows its
    // own geodesic launch, so finished segments' transfer runs while the
    // longest rays (a few tiles near the critical curve, which set the
    // geodesic kernel's length) are still being traced
    static constexpr int NSEG = HADROS_TRACER_MAX_SEGMENTS;
    bool settled = false;                // tracer_settle has checked the last run's pool
    int seg_req = 0;                     // hadros_tracer_set_segments: 0 = automatic
    int n_seg = 1;                       // segments of the last split run
    hipStream_t seg_s[NSEG] = {};        // [1, NSEG): default priority, created on first use
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_geo[NSEG] = {}, ev_end[NSEG] = {};
    hipStream_t stream = nullptr;   // stream of the last run (a pool regrowth re-runs on it)
    bool pool_too_big = false;  // AUTO: the pool the rays need exceeds SPLIT_POOL_MAX_BYTES, run fused
    bool ran = false;
    MevState mev;
    bool ran_mev = false;       // the last run computed the MeV columns
};

struct hadros_replay {
    ReplayArgs args{};
    int precision = 0;
    long n_rays = 0, n_points = 0;
    double* d_sig = nullptr;
    float* d_sigf = nullptr;
    long* d_off = nullptr;
    double *d_r = nullptr, *d_th = nullptr, *d_dl = nullptr, *d_z = nullptr;
    double *d_tau = nullptr, *d_p = nullptr, *d_I = nullptr, *d_bt = nullptr, *d_bp = nullptr, *d_bi = nullptr;
    unsigned long long* d_cnt = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ran = false;
    MevState mev;
    bool ran_mev = false;
};

struct hadros_batch {
    BatchArgs args{};
    int mode = 0, precision = 0, nb_max = 0, block = 128, spc = 8, cu = 0;
    long n_rays = 0;
    std::vector<long> pix0;
    std::vector<int> nbc;
    DevCam* d_cams = nullptr;
    long* d_pix0 = nullptr;
    DevRt* d_rts = nullptr;
    double* d_sig = nullptr;
    float* d_sigf = nullptr;
    unsigned long long *d_queue = nullptr, *d_cnt = nullptr;
    double *d_tau = nullptr, *d_p = nullptr, *d_I = nullptr, *d_bt = nullptr, *d_bp = nullptr, *d_bi = nullptr;
    int *d_cap = nullptr, *d_np = nullptr, *d_flags = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ran = false;
    // split form: sub-batches of `sub` rays, each traced into the chunk pool by
    // k_batch_points and then integrated by k_replay
    int path = HADROS_PATH_AUTO;
    bool split = false;
    long sub = 0, capacity = 0, last_limit = 0;
    double pool_max = 0.0;          // bytes the pool may take
    double* d_pool = nullptr;
    int *d_next = nullptr, *d_nst = nullptr;
    unsigned long long* d_ctr = nullptr;   // [n_ctr] chunk counter of each sub-batch
    int n_ctr = 0;
    // with more rays than one sub-batch holds, the pool is two halves: the
    // replay of a sub-batch (stream s2) overlaps the next sub-batch's
    // geodesic kernel (the run's stream) in the other half
    int halves = 1;
    hipStream_t s2 = nullptr;
    hipEvent_t evg[2] = {nullptr, nullptr};   // geodesic kernel done, per half
    hipEvent_t evr[2] = {nullptr, nullptr};   // replay done (half free again)
};

namespace {
int batch_settle(hadros_batch* b);
}

extern "C" {

const char* hadros_last_error(void) { return g_err.c_str(); }
int hadros_abi_version(void) { return HADROS_ABI_VERSION; }

int hadros_device_info(char* name, int name_cap, int* cu_count, int* gcn_arch)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(HADROS_ERR_DEVICE, "no HIP device available");
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    HIPC(hipGetDeviceProperties(&prop, dev));
    if (name && name_cap > 0) std::snprintf(name, size_t(name_cap), "%s (%s)", prop.name, prop.gcnArchName);
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (gcn_arch) *gcn_arch = std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 950 : 0;
    return HADROS_OK;
}

void hadros_camera_defaults(hadros_camera* c)
{
    if (!c) return;
    c->a_spin = 0.9;
    c->r_obs = 60.0;
    c->theta_obs_rad = M_PI * 74.0 / 180.0;
    c->fov_deg = 45.0;
    c->nx = 100;
    c->ny = 100;
    c->r_max = 120.0;
    c->h = 0.001;
    c->tol = 1.0e-6;
    c->mode = HADROS_MODE_REF;
    c->max_steps = 200000;
}

void hadros_scene_defaults(hadros_scene* s)
{
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    hadros_medium& m = s->medium;
    m.kind = HADROS_PROFILE_GAUSSIAN;
    m.rho0 = 1.0e10;
    m.r0 = 10.0;
    m.sigma_r = 5.0;
    m.h_over_r = 0.01;
    m.radial_power = 2.0;
    m.funnel_depletion = 0.0;
    m.funnel_theta_rad = 15.0 * M_PI / 180.0;
    m.envelope_rho0 = 0.0;
    m.envelope_alpha = 2.5;
    m.r_min = 4.0;
    m.r_max = 60.0;
    m.rho_floor = 1.0e-99;
    hadros_source& so = s->source;
    so.kind = HADROS_SOURCE_INNER_RING;
    so.r_center = 3.5;
    so.sigma_r = 1.0;
    so.theta_width_rad = 8.0 * 3.141592653589793238462643383279502884 / 180.0;
    so.powerlaw = 2.0;
    so.emax_GeV = 1.0e12;
    so.norm = 1.0;
    so.funnel_theta_rad = 20.0 * 3.141592653589793238462643383279502884 / 180.0;
    so.density_q = 1.0;
    so.radial_s = 2.0;
    so.rho_ref = -1.0;
    so.cutoff_min = 0.0;
    so.cutoff_max = 1.0e2;
    so.gradient_dr = 0.1;
    so.gradient_dtheta_rad = 1.0 * 3.141592653589793238462643383279502884 / 180.0;
    hadros_spectrum& sp = s->spectrum;
    sp.kind = HADROS_SPECTRUM_MONOCHROMATIC;
    sp.n_bins = 8;
    sp.gamma = 2.0;
    sp.ecut_GeV = 1.0e12;
    sp.e_min_GeV = 1.0e5;
    sp.e_max_GeV = 1.0e12;
    s->enu_obs_GeV = 1.0e9;
    s->m_bh_msun = 3.0;
}

int hadros_parse_profile(const char* name, int32_t* kind)
{
    if (!name || !kind) return fail(HADROS_ERR_ARG, "NULL argument");
    const std::string n(name);
    int k = -1;
    if (n == "gaussian" || n == "torus" || n == "gaussian_torus") k = 0;
    else if (n == "powerlaw" || n == "powerlaw_disk") k = 1;
    else if (n == "funnel" || n == "gaussian_funnel") k = 2;
    else if (n == "powerlaw_funnel") k = 3;
    else if (n == "gaussian_envelope") k = 4;
    else if (n == "powerlaw_envelope") k = 5;
    else if (n == "powerlaw_funnel_envelope") k = 6;
    else if (n == "collapsar_ndaf_like" || n == "collapsar_ndaf") k = 7;
    if (k < 0)
        return fail(HADROS_ERR_KIND, "Unknown density profile '" + n +
                                         "'. Use gaussian, powerlaw, gaussian_funnel, powerlaw_funnel, "
                                         "gaussian_envelope, powerlaw_envelope, powerlaw_funnel_envelope, "
                                         "or collapsar_ndaf_like.");
    *kind = k;
    return HADROS_OK;
}

int hadros_parse_source(const char* name, int32_t* kind)
{
    if (!name || !kind) return fail(HADROS_ERR_ARG, "NULL argument");
    const std::string n(name);
    int k = -1;
    if (n == "inner_ring" || n == "INNER_RING" || n == "ring") k = 0;
    else if (n == "funnel_wall" || n == "FUNNEL_WALL") k = 1;
    else if (n == "jet_base" || n == "JET_BASE") k = 2;
    else if (n == "shock_layer" || n == "SHOCK_LAYER") k = 3;
    else if (n == "density_weighted" || n == "DENSITY_WEIGHTED") k = 4;
    if (k < 0)
        return fail(HADROS_ERR_KIND, "Unknown UHE source model '" + n +
                                         "'. Use inner_ring, funnel_wall, jet_base, shock_layer, or density_weighted.");
    *kind = k;
    return HADROS_OK;
}

int hadros_parse_spectrum(const char* name, int32_t* kind)
{
    if (!name || !kind) return fail(HADROS_ERR_ARG, "NULL argument");
    const std::string n(name);
    int k = -1;
    if (n == "monochromatic" || n == "MONOCHROMATIC" || n == "mono") k = 0;
    else if (n == "powerlaw" || n == "POWERLAW" || n == "power_law") k = 1;
    else if (n == "powerlaw_cutoff" || n == "POWERLAW_CUTOFF" || n == "power_law_cutoff") k = 2;
    if (k < 0)
        return fail(HADROS_ERR_KIND,
                    "Unknown UHE spectral model '" + n + "'. Use monochromatic, powerlaw, or powerlaw_cutoff.");
    *kind = k;
    return HADROS_OK;
}

void hadros_mev_defaults(hadros_mev* m)
{
    if (!m) return;
    *m = hadros_mev{};
    m->enabled = 1;
    m->model = HADROS_MEV_MODEL_PHYSICAL;
    m->flavor = HADROS_MEV_ANTI_NU_E;
    m->thermal_profile = HADROS_MEV_T_INNER_HOT_TORUS;
    m->ye_profile = HADROS_MEV_YE_NEUTRON_RICH_TORUS;
    m->spectral_mode = HADROS_MEV_SPECTRAL_MONOCHROMATIC;
    m->n_bins = 8;
    m->processes = HADROS_MEV_URCA | HADROS_MEV_PAIR | HADROS_MEV_BREMS | HADROS_MEV_ABSORPTION |
                   HADROS_MEV_SCATTERING | HADROS_MEV_ABS_N | HADROS_MEV_ABS_P | HADROS_MEV_SCAT_N |
                   HADROS_MEV_SCAT_P | HADROS_MEV_SCAT_E;
    m->enu_obs_MeV = 10.0;
    m->norm = 1.0;
    m->sigma_abs0_cm2 = 9.6e-44;
    m->sigma_scat0_cm2 = 1.7e-44;
    m->sigma_scat_e0_cm2 = 3.4e-45;
    m->neutrinosphere_tau = 2.0 / 3.0;
    m->T0_MeV = 6.0;
    m->T_floor_MeV = 0.1;
    m->T_power = 0.2;
    m->Ye_torus = 0.25;
    m->Ye_funnel = 0.55;
    m->Ye_envelope = 0.45;
    m->Ye_floor = 0.01;
    m->Ye_ceil = 0.60;
    m->E_min_MeV = 3.0;
    m->E_max_MeV = 50.0;
}

namespace {
int parse_names(const char* name, int32_t* kind, std::initializer_list<std::pair<const char*, int>> table,
                const std::string& what)
{
    if (!name || !kind) return fail(HADROS_ERR_ARG, "NULL argument");
    for (const auto& e : table)
        if (std::strcmp(name, e.first) == 0) {
            *kind = e.second;
            return HADROS_OK;
        }
    return fail(HADROS_ERR_KIND, "Unknown " + what.substr(0, what.find('|')) + " '" + name + "'. " +
                                     what.substr(what.find('|') + 1));
}
}  // namespace

// parse_mev_* (src/mev_neutrino_physics.cpp:36-233): names, aliases, messages
int hadros_parse_mev_model(const char* name, int32_t* kind)
{
    return parse_names(name, kind, {{"toy", 0}, {"TOY", 0}, {"physical", 1}, {"PHYSICAL", 1}, {"phys", 1}},
                       "MEV_MODEL|Use toy or physical.");
}
int hadros_parse_mev_flavor(const char* name, int32_t* kind)
{
    return parse_names(name, kind,
                       {{"nu_e", 0}, {"nue", 0}, {"NU_E", 0}, {"anti_nu_e", 1}, {"antinue", 1}, {"anti-nu-e", 1},
                        {"ANTI_NU_E", 1}, {"nu_x", 2}, {"nux", 2}, {"NU_X", 2}},
                       "MEV_FLAVOR|Use nu_e, anti_nu_e, or nu_x.");
}
int hadros_parse_mev_thermal(const char* name, int32_t* kind)
{
    return parse_names(name, kind,
                       {{"constant", 0}, {"CONSTANT", 0}, {"inner_hot_torus", 1}, {"INNER_HOT_TORUS", 1},
                        {"radial_powerlaw", 2}, {"RADIAL_POWERLAW", 2}, {"torus_plus_cool_envelope", 3},
                        {"TORUS_PLUS_COOL_ENVELOPE", 3}, {"collapsar_inner_hot", 4}, {"COLLAPSAR_INNER_HOT", 4}},
                       "MEV_THERMAL_PROFILE|Use constant, inner_hot_torus, radial_powerlaw, "
                       "torus_plus_cool_envelope, or collapsar_inner_hot.");
}
int hadros_parse_mev_ye(const char* name, int32_t* kind)
{
    return parse_names(name, kind,
                       {{"constant", 0}, {"CONSTANT", 0}, {"neutron_rich_torus", 1}, {"NEUTRON_RICH_TORUS", 1},
                        {"funnel_proton_rich", 2}, {"FUNNEL_PROTON_RICH", 2}, {"torus_envelope_contrast", 3},
                        {"TORUS_ENVELOPE_CONTRAST", 3}, {"collapsar_neutron_rich", 4},
                        {"COLLAPSAR_NEUTRON_RICH", 4}},
                       "MEV_YE_PROFILE|Use constant, neutron_rich_torus, funnel_proton_rich, "
                       "torus_envelope_contrast, or collapsar_neutron_rich.");
}
int hadros_parse_mev_spectral(const char* name, int32_t* kind)
{
    return parse_names(name, kind,
                       {{"monochromatic", 0}, {"MONOCHROMATIC", 0}, {"mono", 0}, {"fermi_dirac_band", 1},
                        {"FERMI_DIRAC_BAND", 1}, {"fd_band", 1}},
                       "MEV_SPECTRAL_MODE|Use monochromatic or fermi_dirac_band.");
}

const char* hadros_profile_name(int32_t k)
{
    static const char* names[] = {"gaussian", "powerlaw", "gaussian_funnel", "powerlaw_funnel",
                                  "gaussian_envelope", "powerlaw_envelope", "powerlaw_funnel_envelope",
                                  "collapsar_ndaf_like"};
    return (k >= 0 && k < 8) ? names[k] : "unknown";
}
const char* hadros_source_name(int32_t k)
{
    static const char* names[] = {"inner_ring", "funnel_wall", "jet_base", "shock_layer", "density_weighted"};
    return (k >= 0 && k < 5) ? names[k] : "unknown";
}
const char* hadros_spectrum_name(int32_t k)
{
    static const char* names[] = {"monochromatic", "powerlaw", "powerlaw_cutoff"};
    return (k >= 0 && k < 3) ? names[k] : "unknown";
}

// SigmaTable::SigmaTable (src/sigma_table.cpp:9-53)
int hadros_sigma_load(const char* path, double* e, double* sg2, double* scm2, int cap, int* n)
{
    if (!path || !n) return fail(HADROS_ERR_ARG, "NULL argument");
    std::ifstream in(path);
    if (!in) return fail(HADROS_ERR_IO, std::string("Could not open sigma table: ") + path);
    std::string line;
    int cnt = 0;
    std::vector<double> E, G, S;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        double a, b, c;
        if (!(iss >> a >> b >> c)) continue;
        if (a <= 0.0 || b <= 0.0 || c <= 0.0) continue;
        E.push_back(a);
        G.push_back(b);
        S.push_back(c);
        ++cnt;
    }
    if (cnt < 2) return fail(HADROS_ERR_TABLE, "Sigma table must contain at least two valid data lines.");
    for (int k = 1; k < cnt; ++k)
        if (E[k] <= E[k - 1]) return fail(HADROS_ERR_TABLE, "Energy grid in sigma table must be strictly increasing.");
    for (int k = 0; k < cnt && k < cap; ++k) {
        if (e) e[k] = E[k];
        if (sg2) sg2[k] = G[k];
        if (scm2) scm2[k] = S[k];
    }
    *n = cnt;
    return HADROS_OK;
}

int hadros_sigma_check(const double* e, const double* s, int n) { return check_sigma(e, s, n); }

// ---------------------------------------------------------------- tracer plan
int hadros_tracer_create(const hadros_camera* cam, const hadros_scene* scene, const double* se, const double* ss,
                         int n_sigma, int precision, hadros_tracer** out)
{
    if (!cam || !scene || !out) return fail(HADROS_ERR_ARG, "NULL argument");
    if (precision != HADROS_PREC_FP64 && precision != HADROS_PREC_FP32K) return fail(HADROS_ERR_ARG, "unknown precision");
    int rc = ensure_device();
    if (rc) return rc;
    hadros_tracer* t = new hadros_tracer();
    rc = make_cam(*cam, t->args.cam);
    if (!rc) rc = make_rt(*scene, se, ss, n_sigma, cam->a_spin, t->args.rt);
    if (rc) {
        delete t;
        return rc;
    }
    t->precision = precision;
    t->block = 0;   // automatic: 128 for the fused kernel, 64 for the split geodesic kernel
    t->npix = long(cam->nx) * cam->ny;
    std::vector<double> sd;
    std::vector<float> sf;
    sigma_tables(se, ss, n_sigma, sd, sf);
    const size_t nb = size_t(std::max(t->args.rt.nb, 0));
    const size_t np = size_t(t->npix);
    if ((rc = dalloc(&t->d_sig, sd.size())) || (rc = dalloc(&t->d_sigf, sf.size())) || (rc = dalloc(&t->d_tau, np)) ||
        (rc = dalloc(&t->d_p, np)) || (rc = dalloc(&t->d_I, np)) || (rc = dalloc(&t->d_alpha, np)) ||
        (rc = dalloc(&t->d_beta, np)) || (rc = dalloc(&t->d_bt, np * nb)) || (rc = dalloc(&t->d_bp, np * nb)) ||
        (rc = dalloc(&t->d_bi, np * nb)) || (rc = dalloc(&t->d_cap, np)) || (rc = dalloc(&t->d_np, np)) ||
        (rc = dalloc(&t->d_flags, np)) || (rc = dalloc(&t->d_cnt, 5)) || (rc = dalloc(&t->d_rt, 1))) {
        hadros_tracer_destroy(t);
        return rc;
    }
    HIPC(hipMemcpy(t->d_rt, &t->args.rt, sizeof(DevRt), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(t->d_sig, sd.data(), sd.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(t->d_sigf, sf.data(), sf.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPC(hipEventCreate(&t->ev0));
    HIPC(hipEventCreate(&t->ev1));
    TraceArgs& A = t->args;
    A.sig = t->d_sig;
    A.sigf = t->d_sigf;
    A.rtg = t->d_rt;
    A.out = TraceOut{t->d_tau, t->d_p, t->d_I, t->d_alpha, t->d_beta, t->d_bt, t->d_bp, t->d_bi,
                     t->d_cap, t->d_np, t->d_flags, t->d_cnt};
    PointsArgs& Q = t->pargs;
    Q.cam = A.cam;
    Q.out = A.out;
    ReplayArgs& R = t->rargs;
    R.rt = A.rt;
    R.sig = t->d_sig;
    R.sigf = t->d_sigf;
    R.n_rays = int(t->npix);
    R.tau = t->d_tau;
    R.p = t->d_p;
    R.I = t->d_I;
    R.btau = t->d_bt;
    R.bp = t->d_bp;
    R.bi = t->d_bi;
    *out = t;
    return HADROS_OK;
}

namespace {

// pool for the split form: a first chunk per pixel plus `extra` chunks
int tracer_pool(hadros_tracer* t, long extra)
{
    dfree(t->d_pool);
    dfree(t->d_next);
    t->capacity = 0;
    const long cap = t->npix + extra;
    int rc = 0;
    if ((rc = dalloc(&t->d_pool, size_t(cap) * CHUNK_DOUBLES)) || (rc = dalloc(&t->d_next, size_t(cap)))) return rc;
    if (!t->d_nst && (rc = dalloc(&t->d_nst, size_t(t->npix)))) return rc;
    t->d_ctr = t->d_cnt + 4;
    t->capacity = cap;
    t->pargs.cp = ChunkPool{t->d_pool, t->d_next, reinterpret_cast<unsigned*>(t->d_ctr), t->npix, cap};
    t->pargs.n_st = t->d_nst;
    t->rargs.pool = t->d_pool;
    t->rargs.next = t->d_next;
    t->rargs.n_st = t->d_nst;
    return HADROS_OK;
}

// largest chunk pool of the split form; bigger images use the fused kernel
constexpr double SPLIT_POOL_MAX_BYTES = 8.0e9;

bool want_split(const hadros_tracer* t)
{
    if (t->path == HADROS_PATH_FUSED) return false;
    if (t->path == HADROS_PATH_SPLIT || t->mev.on) return true;   // the MeV channel lives in the transfer kernel
    if (t->pool_too_big) return false;
    return t->args.cam.mode == HADROS_MODE_CFG &&
           3.0 * double(t->npix) * CHUNK_DOUBLES * sizeof(double) <= SPLIT_POOL_MAX_BYTES;
}

int tracer_launch(hadros_tracer* t, hipStream_t s)
{
    if (t->mev.on && t->path == HADROS_PATH_FUSED)
        return fail(HADROS_ERR_ARG, "the MeV channel runs in the split form: HADROS_PATH_FUSED cannot compute it");
    // the run counters and the chunk counter, cleared by one fill kernel (no
    // memset nodes if the caller captures the stream)
    HIPC(launch_zero_u64(t->d_cnt, 5, s));
    t->stream = s;
    t->settled = false;
    t->split = want_split(t);
    mev_apply(t->mev, t->rargs);
    t->ran_mev = t->split && t->mev.on;
    if (t->split && !t->d_pool) {
        const int rc = tracer_pool(t, 2 * t->npix);
        if (rc) return rc;
    }
    HIPC(hipEventRecord(t->ev0, s));
    if (t->split) {
        const DevCam& C = t->pargs.cam;
        const int tiles_y = (C.ny + 7) / 8, tile_rows = (C.nx + 7) / 8;
        // band i covers tile rows [rows[i], rows[i + 1])
        static_assert(hadros_tracer::NSEG >= 4, "the automatic bands need four streams");
        int rows[hadros_tracer::NSEG + 1];
        int nseg;
        if (!t->seg_req && tile_rows >= 16) {
            // automatic: the image's middle rows hold the rays that pass near
            // the photon orbit (the longest, for a camera aimed at the hole),
            // so two narrow middle bands finish last, and the outer bands'
            // transfer (most of the image's points) runs under their tail
            nseg = 4;
            rows[0] = 0;
            rows[1] = (3 * tile_rows) / 8;
            rows[2] = tile_rows / 2;
            rows[3] = (5 * tile_rows) / 8;
            rows[4] = tile_rows;
        } else {
            // n equal bands (automatic on small cameras: of at least 4 tile rows)
            nseg = t->seg_req ? std::min(t->seg_req, tile_rows)
                              : std::max(1, std::min(hadros_tracer::NSEG, tile_rows / 4));
            for (int i = 0; i <= nseg; ++i) rows[i] = int(long(tile_rows) * i / nseg);
        }
        t->n_seg = nseg;
        if (nseg > 1) {
            if (!t->ev_fork) HIPC(hipEventCreateWithFlags(&t->ev_fork, hipEventDisableTiming));
            for (int i = 1; i < nseg; ++i)
                if (!t->seg_s[i]) HIPC(hipStreamCreateWithFlags(&t->seg_s[i], hipStreamNonBlocking));
            HIPC(hipEventRecord(t->ev_fork, s));
        }
        // all geodesic launches first, so that the segments start together;
        // then each segment's transfer behind its own geodesic launch
        for (int i = 0; i < nseg; ++i) {
            if (!t->ev_geo[i]) {
                HIPC(hipEventCreate(&t->ev_geo[i]));
                HIPC(hipEventCreate(&t->ev_end[i]));
            }
            hipStream_t si = (i == 0) ? s : t->seg_s[i];
            if (i > 0) HIPC(hipStreamWaitEvent(si, t->ev_fork, 0));
            const int r0 = rows[i], r1 = rows[i + 1];
            PointsArgs Q = t->pargs;
            Q.tile0 = r0 * tiles_y;
            Q.tile1 = r1 * tiles_y;
            HIPC(launch_trace_points(Q, t->block ? t->block : 64, si));
            HIPC(hipEventRecord(t->ev_geo[i], si));
        }
        for (int i = 0; i < nseg; ++i) {
            hipStream_t si = (i == 0) ? s : t->seg_s[i];
            const int r0 = rows[i], r1 = rows[i + 1];
            // pixels of tile rows [r0, r1): rows [8 r0, 8 r1) of the image
            ReplayArgs R = t->rargs;
            R.ray0 = long(std::min(8 * r0, C.nx)) * C.ny;
            R.n_rays = int(long(std::min(8 * r1, C.nx)) * C.ny - R.ray0);
            HIPC(launch_replay(R, t->precision, si));
            HIPC(hipEventRecord(t->ev_end[i], si));
            if (i > 0) HIPC(hipStreamWaitEvent(s, t->ev_end[i], 0));
        }
    } else {
        HIPC(launch_trace_rt(t->args, t->precision, t->block ? t->block : 128, s));
    }
    HIPC(hipEventRecord(t->ev1, s));
    t->ran = true;
    return HADROS_OK;
}

// after a split run: if rays ran out of chunks, grow the pool to what they
// asked for (or give up on the split form) and run again
int tracer_settle(hadros_tracer* t)
{
    // the fetch and timing calls after one run settle it once (one counter copy)
    if (t->settled) return HADROS_OK;
    HIPC(hipEventSynchronize(t->ev1));
    // the counter holds every extra chunk the rays needed, so one retry suffices;
    // the loop bound only guards against a pool that cannot be allocated
    for (int attempt = 0; attempt < 3 && t->split; ++attempt) {
        unsigned long long req = 0;
        HIPC(hipMemcpy(&req, t->d_ctr, sizeof(req), hipMemcpyDeviceToHost));
        const long need = t->npix + long(unsigned(req));
        if (need <= t->capacity) {
            t->settled = true;
            return HADROS_OK;
        }
        const long grown = long(double(need) * 1.25);
        const double bytes = double(grown) * CHUNK_DOUBLES * sizeof(double);
        int rc = 0;
        if (t->path == HADROS_PATH_SPLIT || t->mev.on || bytes <= SPLIT_POOL_MAX_BYTES) {
            if ((rc = tracer_pool(t, grown - t->npix))) return rc;
        } else {
            // AUTO only: this camera runs fused from now on; the caller's path
            // setting stays as it was (a later MeV run still takes the split form)
            dfree(t->d_pool);
            dfree(t->d_next);
            t->capacity = 0;
            t->pool_too_big = true;
        }
        if ((rc = tracer_launch(t, t->stream))) return rc;
        HIPC(hipEventSynchronize(t->ev1));
    }
    if (t->split) {
        unsigned long long req = 0;
        HIPC(hipMemcpy(&req, t->d_ctr, sizeof(req), hipMemcpyDeviceToHost));
        if (t->npix + long(unsigned(req)) > t->capacity) return fail(HADROS_ERR_MAX_POINTS, "split tracer: chunk pool too small");
    }
    t->settled = true;
    return HADROS_OK;
}

}  // namespace

int hadros_tracer_set_block(hadros_tracer* t, int block)
{
    if (!t) return fail(HADROS_ERR_ARG, "NULL tracer");
    if (block != 0 && (block < 64 || block > 256 || block % 64))
        return fail(HADROS_ERR_ARG, "block size must be 0 (automatic), 64, 128, 192 or 256");
    t->block = block;
    return HADROS_OK;
}

int hadros_tracer_set_segments(hadros_tracer* t, int n_segments)
{
    if (!t) return fail(HADROS_ERR_ARG, "NULL tracer");
    if (n_segments < 0 || n_segments > HADROS_TRACER_MAX_SEGMENTS)
        return fail(HADROS_ERR_ARG, "segments must be 0 (automatic) or 1 to HADROS_TRACER_MAX_SEGMENTS");
    t->seg_req = n_segments;
    return HADROS_OK;
}

int hadros_tracer_run(hadros_tracer* t, void* stream)
{
    if (!t) return fail(HADROS_ERR_ARG, "NULL tracer");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    // counters cleared by a fill kernel (no memset nodes if the caller captures the stream)
    return tracer_launch(t, s);
}

int hadros_tracer_set_path(hadros_tracer* t, int path)
{
    if (!t) return fail(HADROS_ERR_ARG, "NULL tracer");
    if (path != HADROS_PATH_AUTO && path != HADROS_PATH_FUSED && path != HADROS_PATH_SPLIT)
        return fail(HADROS_ERR_ARG, "path must be HADROS_PATH_AUTO, _FUSED or _SPLIT");
    t->path = path;
    return HADROS_OK;
}

int hadros_tracer_last_ms(hadros_tracer* t, double* ms)
{
    if (!t || !ms || !t->ran) return fail(HADROS_ERR_ARG, "tracer has not run");
    const int rc = tracer_settle(t);
    if (rc) return rc;
    float f = 0.0f;
    HIPC(hipEventElapsedTime(&f, t->ev0, t->ev1));
    *ms = f;
    return HADROS_OK;
}

int hadros_tracer_set_mev(hadros_tracer* t, const hadros_mev* mev)
{
    if (!t) return fail(HADROS_ERR_ARG, "NULL tracer");
    if (t->ran) HIPC(hipEventSynchronize(t->ev1));   // the parameters may still be in use
    return mev_set(t->mev, mev, t->args.rt.med, t->npix);
}

int hadros_tracer_fetch_mev(hadros_tracer* t, hadros_mev_out* o)
{
    if (!t || !t->ran || !o) return fail(HADROS_ERR_ARG, "tracer has not run");
    const int rc = tracer_settle(t);
    if (rc) return rc;
    return mev_fetch(t->mev, t->ran_mev, t->npix, o);
}

int hadros_tracer_kernel_ms(hadros_tracer* t, double* geodesic_ms, double* transfer_ms)
{
    if (!t || !t->ran) return fail(HADROS_ERR_ARG, "tracer has not run");
    const int rc = tracer_settle(t);
    if (rc) return rc;
    float g = 0.0f, r = 0.0f;
    if (t->split) {
        // the segments overlap: the geodesic span runs from the start to the
        // last segment's geodesic end, the transfer span from the first
        // segment's transfer start to the end
        float tot = 0.0f, first_rt = 0.0f;
        HIPC(hipEventElapsedTime(&tot, t->ev0, t->ev1));
        for (int i = 0; i < t->n_seg; ++i) {
            float gi = 0.0f;
            HIPC(hipEventElapsedTime(&gi, t->ev0, t->ev_geo[i]));
            g = std::max(g, gi);
            first_rt = (i == 0) ? gi : std::min(first_rt, gi);
        }
        r = tot - first_rt;
    } else {
        HIPC(hipEventElapsedTime(&g, t->ev0, t->ev1));
    }
    if (geodesic_ms) *geodesic_ms = g;
    if (transfer_ms) *transfer_ms = r;
    return HADROS_OK;
}

int hadros_tracer_fetch(hadros_tracer* t, hadros_image_out* o, hadros_stats* st)
{
    if (!t || !t->ran) return fail(HADROS_ERR_ARG, "tracer has not run");
    const int rs = tracer_settle(t);
    if (rs) return rs;
    const size_t np = size_t(t->npix);
    const size_t nb = size_t(std::max(t->args.rt.nb, 0));
    auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
        if (dst && src && bytes) HIPC(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return HADROS_OK;
    };
    int rc = 0;
    if (o) {
        if ((rc = cp(o->tau, t->d_tau, np * 8)) || (rc = cp(o->p_surv, t->d_p, np * 8)) ||
            (rc = cp(o->intensity, t->d_I, np * 8)) || (rc = cp(o->captured, t->d_cap, np * 4)) ||
            (rc = cp(o->npoints, t->d_np, np * 4)) || (rc = cp(o->flags, t->d_flags, np * 4)) ||
            (rc = cp(o->alpha, t->d_alpha, np * 8)) || (rc = cp(o->beta, t->d_beta, np * 8)) ||
            (rc = cp(o->bin_tau, t->d_bt, np * nb * 8)) || (rc = cp(o->bin_p, t->d_bp, np * nb * 8)) ||
            (rc = cp(o->bin_i, t->d_bi, np * nb * 8)))
            return rc;
    }
    if (st) {
        unsigned long long c[4];
        HIPC(hipMemcpy(c, t->d_cnt, sizeof(c), hipMemcpyDeviceToHost));
        float f = 0.0f;
        HIPC(hipEventElapsedTime(&f, t->ev0, t->ev1));
        st->kernel_ms = f;
        st->total_ms = f;
        st->total_points = (int64_t)c[0];
        st->total_attempts = (int64_t)c[1];
        st->n_captured = int32_t(c[2]);
        st->n_exhausted = int32_t(c[3]);
        st->n_bins = t->args.rt.nb;
        st->block_size = t->block ? t->block : (t->split ? 64 : 128);
    }
    return HADROS_OK;
}

int hadros_tracer_destroy(hadros_tracer* t)
{
    if (!t) return HADROS_OK;
    dfree(t->d_sig);
    dfree(t->d_sigf);
    dfree(t->d_tau);
    dfree(t->d_p);
    dfree(t->d_I);
    dfree(t->d_alpha);
    dfree(t->d_beta);
    dfree(t->d_bt);
    dfree(t->d_bp);
    dfree(t->d_bi);
    dfree(t->d_cap);
    dfree(t->d_np);
    dfree(t->d_flags);
    dfree(t->d_cnt);
    dfree(t->d_pool);
    dfree(t->d_next);
    dfree(t->d_nst);
    dfree(t->d_rt);
    mev_free(t->mev);
    if (t->ev0) (void)hipEventDestroy(t->ev0);
    if (t->ev1) (void)hipEventDestroy(t->ev1);
    if (t->ev_fork) (void)hipEventDestroy(t->ev_fork);
    for (int i = 0; i < hadros_tracer::NSEG; ++i) {
        if (t->ev_geo[i]) (void)hipEventDestroy(t->ev_geo[i]);
        if (t->ev_end[i]) (void)hipEventDestroy(t->ev_end[i]);
        if (t->seg_s[i]) (void)hipStreamDestroy(t->seg_s[i]);
    }
    delete t;
    return HADROS_OK;
}

int hadros_trace_rt(const hadros_camera* cam, const hadros_scene* scene, const double* se, const double* ss,
                    int n_sigma, int precision, hadros_image_out* out, hadros_stats* stats)
{
    const auto t0 = std::chrono::steady_clock::now();
    hadros_tracer* t = nullptr;
    int rc = hadros_tracer_create(cam, scene, se, ss, n_sigma, precision, &t);
    if (rc) return rc;
    rc = hadros_tracer_run(t, nullptr);
    if (!rc) rc = hadros_tracer_fetch(t, out, stats);
    hadros_tracer_destroy(t);
    if (!rc && stats)
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}

}  // extern "C"

// ------------------------------------------------- allocation-free fused trace
// Device scratch of hadros_trace_rt_noalloc, one copy per device, held until
// the process ends.  Layout (bytes): [0, 32) the four counters when the caller
// passes none; [NOALLOC_RT_OFF, +sizeof(DevRt)) the scene block the fused
// kernel re-reads per path point; [NOALLOC_TAB_OFF, ...) the sigma table,
// 3 n doubles (FP64) or 3 n floats (FP32K).  Host code never names the array:
// it uses the address hipGetSymbolAddress returned for the current device.
static __device__ __align__(256) unsigned char scratch[HADROS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr size_t NOALLOC_RT_OFF = 256;
constexpr size_t NOALLOC_TAB_OFF = NOALLOC_RT_OFF + (sizeof(DevRt) + 255) / 256 * 256;
static_assert(sizeof(DevRt) % 8 == 0, "the scene block is staged in 8-byte words");
static_assert(NOALLOC_TAB_OFF + 3 * size_t(HADROS_MAX_SIGMA) * sizeof(double) <= HADROS_NOALLOC_SCRATCH_BYTES,
              "the scratch holds the largest sigma table");

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[MAX_DEVICES] = {};

// this device's copy of the scratch array (looked up once per device, with it current)
int noalloc_scratch(unsigned char** out)
{
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(HADROS_ERR_DEVICE, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(HADROS_ERR_DEVICE, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return HADROS_OK;
}

// Host data reaches the scratch as kernel arguments (2 KiB per launch), so the
// call needs no staging buffer and no host-to-device copy.
constexpr int STAGE_WORDS = 256;
struct StageBlock {
    unsigned long long w[STAGE_WORDS];
    unsigned long long* dst;
    int n;
};

__global__ void k_stage(StageBlock b)
{
    for (int t = threadIdx.x; t < b.n; t += blockDim.x) b.dst[t] = b.w[t];
}

// fill(word_index) -> 64-bit word; writes n_words words at dst on stream s
template <class Fill>
int stage_words(unsigned long long* dst, long n_words, Fill fill, hipStream_t s)
{
    StageBlock b;
    for (long w0 = 0; w0 < n_words; w0 += STAGE_WORDS) {
        b.n = int(std::min<long>(STAGE_WORDS, n_words - w0));
        b.dst = dst + w0;
        for (int t = 0; t < b.n; ++t) b.w[t] = fill(w0 + t);
        hipLaunchKernelGGL(k_stage, dim3(1), dim3(256), 0, s, b);
        HIPC(hipGetLastError());
    }
    return HADROS_OK;
}

}  // namespace

extern "C" {

int hadros_trace_rt_noalloc(const hadros_camera* cam, const hadros_scene* scene, const double* se, const double* ss,
                            int n_sigma, int precision, const hadros_image_out* out, int64_t* counters, void* stream)
{
    if (!cam || !scene || !out) return fail(HADROS_ERR_ARG, "NULL argument");
    if (precision != HADROS_PREC_FP64 && precision != HADROS_PREC_FP32K) return fail(HADROS_ERR_ARG, "unknown precision");
    int rc = ensure_device();
    if (rc) return rc;
    TraceArgs A{};
    if ((rc = make_cam(*cam, A.cam)) || (rc = make_rt(*scene, se, ss, n_sigma, cam->a_spin, A.rt))) return rc;
    const bool fp64 = precision == HADROS_PREC_FP64;
    const size_t tab_bytes = 3 * size_t(n_sigma) * (fp64 ? sizeof(double) : sizeof(float));
    if (NOALLOC_TAB_OFF + tab_bytes > size_t(HADROS_NOALLOC_SCRATCH_BYTES))
        return fail(HADROS_ERR_ARG, "sigma table larger than the allocation-free scratch (HADROS_NOALLOC_SCRATCH_BYTES)");
    unsigned char* base = nullptr;
    if ((rc = noalloc_scratch(&base))) return rc;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);

    // the counters must start at zero: a fill kernel, not a memset node
    unsigned long long* ctr =
        counters ? reinterpret_cast<unsigned long long*>(counters) : reinterpret_cast<unsigned long long*>(base);
    HIPC(launch_zero_u64(ctr, 4, s));
    const DevRt rt = A.rt;
    auto rt_word = [&rt](long w) {
        unsigned long long v;
        std::memcpy(&v, reinterpret_cast<const unsigned char*>(&rt) + 8 * w, 8);
        return v;
    };
    if ((rc = stage_words(reinterpret_cast<unsigned long long*>(base + NOALLOC_RT_OFF), long(sizeof(DevRt) / 8), rt_word,
                          s)))
        return rc;
    unsigned long long* tab = reinterpret_cast<unsigned long long*>(base + NOALLOC_TAB_OFF);
    if (fp64) {
        auto word = [&](long w) {
            const double d = sigma_table_entry(se, ss, n_sigma, int(w));
            unsigned long long v;
            std::memcpy(&v, &d, 8);
            return v;
        };
        if ((rc = stage_words(tab, 3L * n_sigma, word, s))) return rc;
    } else {
        const long nf = 3L * n_sigma;   // two floats per word, the last one padded with 0
        auto word = [&](long w) {
            float f[2] = {float(sigma_table_entry(se, ss, n_sigma, int(2 * w))), 0.0f};
            if (2 * w + 1 < nf) f[1] = float(sigma_table_entry(se, ss, n_sigma, int(2 * w + 1)));
            unsigned long long v;
            std::memcpy(&v, f, 8);
            return v;
        };
        if ((rc = stage_words(tab, (nf + 1) / 2, word, s))) return rc;
    }
    A.rtg = reinterpret_cast<const DevRt*>(base + NOALLOC_RT_OFF);
    A.sig = fp64 ? reinterpret_cast<const double*>(tab) : nullptr;
    A.sigf = fp64 ? nullptr : reinterpret_cast<const float*>(tab);
    A.out = TraceOut{out->tau, out->p_surv, out->intensity, out->alpha, out->beta, out->bin_tau, out->bin_p,
                     out->bin_i, out->captured, out->npoints, out->flags, ctr};
    HIPC(launch_trace_rt(A, precision, 128, s));
    return HADROS_OK;
}

// ------------------------------------------------------- legacy optical depth
// optical_depth::tau_along_ray (src/optical_depth.cpp:56-91) is the transfer
// kernel's tau with a scene of its own: one bin at E_inf, the black-hole mass
// fixed at 3 M_sun (:83-85), no horizon cut, and no emission.
int hadros_tau_along_ray(const hadros_cache_view* c, const hadros_medium* m, const double* se, const double* ss,
                         int n_sigma, double enu_inf, double* tau, double* p_surv)
{
    if (!c || !m) return fail(HADROS_ERR_ARG, "NULL argument");
    hadros_scene sc;
    hadros_scene_defaults(&sc);
    sc.medium = *m;
    sc.enu_obs_GeV = enu_inf;
    sc.m_bh_msun = 3.0;
    sc.spectrum.kind = HADROS_SPECTRUM_MONOCHROMATIC;
    sc.source.kind = HADROS_SOURCE_INNER_RING;
    sc.source.emax_GeV = 0.0;   // emissivity_collapsar_ring returns 0: no intensity
    hadros_replay* p = nullptr;
    int rc = hadros_replay_create(c, &sc, se, ss, n_sigma, HADROS_PREC_FP64, &p);
    if (rc) return rc;
    p->args.rt.r_h_rt = -HUGE_VAL;   // every point counts, inside the horizon too
    rc = hadros_replay_run(p, nullptr);
    hadros_image_out o{};
    o.tau = tau;
    o.p_surv = p_surv;
    if (!rc) rc = hadros_replay_fetch(p, &o, nullptr);
    hadros_replay_destroy(p);
    return rc;
}

// ---------------------------------------------------------------- replay plan
int hadros_replay_create(const hadros_cache_view* c, const hadros_scene* scene, const double* se, const double* ss,
                         int n_sigma, int precision, hadros_replay** out)
{
    if (!c || !scene || !out) return fail(HADROS_ERR_ARG, "NULL argument");
    if (c->n_rays < 0 || (c->n_rays > 0 && !c->offset)) return fail(HADROS_ERR_ARG, "bad cache view");
    if (precision != HADROS_PREC_FP64 && precision != HADROS_PREC_FP32K) return fail(HADROS_ERR_ARG, "unknown precision");
    for (int k = 0; k < c->n_rays; ++k)
        if (c->offset[k + 1] < c->offset[k]) return fail(HADROS_ERR_CACHE_TRUNCATED, "Invalid ray with negative number of points");
    int rc = ensure_device();
    if (rc) return rc;
    hadros_replay* p = new hadros_replay();
    rc = make_rt(*scene, se, ss, n_sigma, c->a_bh, p->args.rt);
    if (rc) {
        delete p;
        return rc;
    }
    p->precision = precision;
    p->n_rays = c->n_rays;
    p->n_points = c->n_rays ? long(c->offset[c->n_rays] - c->offset[0]) : 0;
    std::vector<double> sd;
    std::vector<float> sf;
    sigma_tables(se, ss, n_sigma, sd, sf);
    const size_t nr = size_t(p->n_rays), npt = size_t(p->n_points), nb = size_t(std::max(p->args.rt.nb, 0));
    if ((rc = dalloc(&p->d_sig, sd.size())) || (rc = dalloc(&p->d_sigf, sf.size())) ||
        (rc = dalloc(&p->d_off, nr + 1)) || (rc = dalloc(&p->d_r, npt)) || (rc = dalloc(&p->d_th, npt)) ||
        (rc = dalloc(&p->d_dl, npt)) || (rc = dalloc(&p->d_z, npt)) || (rc = dalloc(&p->d_tau, nr)) ||
        (rc = dalloc(&p->d_p, nr)) || (rc = dalloc(&p->d_I, nr)) || (rc = dalloc(&p->d_bt, nr * nb)) ||
        (rc = dalloc(&p->d_bp, nr * nb)) || (rc = dalloc(&p->d_bi, nr * nb)) || (rc = dalloc(&p->d_cnt, 1))) {
        hadros_replay_destroy(p);
        return rc;
    }
    HIPC(hipMemcpy(p->d_sig, sd.data(), sd.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(p->d_sigf, sf.data(), sf.size() * sizeof(float), hipMemcpyHostToDevice));
    if (nr) {
        // offsets rebased to 0
        std::vector<long> off(nr + 1);
        for (size_t k = 0; k <= nr; ++k) off[k] = long(c->offset[k] - c->offset[0]);
        HIPC(hipMemcpy(p->d_off, off.data(), (nr + 1) * sizeof(long), hipMemcpyHostToDevice));
    }
    if (npt) {
        const size_t b = size_t(c->offset[0]);
        HIPC(hipMemcpy(p->d_r, c->r + b, npt * 8, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(p->d_th, c->theta + b, npt * 8, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(p->d_dl, c->dl + b, npt * 8, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(p->d_z, c->redshift + b, npt * 8, hipMemcpyHostToDevice));
    }
    HIPC(hipEventCreate(&p->ev0));
    HIPC(hipEventCreate(&p->ev1));
    ReplayArgs& A = p->args;
    A.sig = p->d_sig;
    A.sigf = p->d_sigf;
    A.n_rays = int(nr);
    A.offset = p->d_off;
    A.r = p->d_r;
    A.th = p->d_th;
    A.dl = p->d_dl;
    A.z = p->d_z;
    A.tau = p->d_tau;
    A.p = p->d_p;
    A.I = p->d_I;
    A.btau = p->d_bt;
    A.bp = p->d_bp;
    A.bi = p->d_bi;
    A.counters = p->d_cnt;
    *out = p;
    return HADROS_OK;
}

int hadros_replay_run(hadros_replay* p, void* stream)
{
    if (!p) return fail(HADROS_ERR_ARG, "NULL replay");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    HIPC(launch_zero_u64(p->d_cnt, 1, s));
    mev_apply(p->mev, p->args);
    p->ran_mev = p->mev.on;
    HIPC(hipEventRecord(p->ev0, s));
    if (p->n_rays > 0) HIPC(launch_replay(p->args, p->precision, s));
    HIPC(hipEventRecord(p->ev1, s));
    p->ran = true;
    return HADROS_OK;
}

int hadros_replay_last_ms(hadros_replay* p, double* ms)
{
    if (!p || !ms || !p->ran) return fail(HADROS_ERR_ARG, "replay has not run");
    HIPC(hipEventSynchronize(p->ev1));
    float f = 0.0f;
    HIPC(hipEventElapsedTime(&f, p->ev0, p->ev1));
    *ms = f;
    return HADROS_OK;
}

int hadros_replay_fetch(hadros_replay* p, hadros_image_out* o, hadros_stats* st)
{
    if (!p || !p->ran) return fail(HADROS_ERR_ARG, "replay has not run");
    HIPC(hipEventSynchronize(p->ev1));
    const size_t nr = size_t(p->n_rays), nb = size_t(std::max(p->args.rt.nb, 0));
    auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
        if (dst && src && bytes) HIPC(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return HADROS_OK;
    };
    int rc = 0;
    if (o) {
        if ((rc = cp(o->tau, p->d_tau, nr * 8)) || (rc = cp(o->p_surv, p->d_p, nr * 8)) ||
            (rc = cp(o->intensity, p->d_I, nr * 8)) || (rc = cp(o->bin_tau, p->d_bt, nr * nb * 8)) ||
            (rc = cp(o->bin_p, p->d_bp, nr * nb * 8)) || (rc = cp(o->bin_i, p->d_bi, nr * nb * 8)))
            return rc;
    }
    if (st) {
        unsigned long long c = 0;
        HIPC(hipMemcpy(&c, p->d_cnt, sizeof(c), hipMemcpyDeviceToHost));
        float f = 0.0f;
        HIPC(hipEventElapsedTime(&f, p->ev0, p->ev1));
        std::memset(st, 0, sizeof(*st));
        st->kernel_ms = f;
        st->total_ms = f;
        st->total_points = (int64_t)c;
        st->n_bins = p->args.rt.nb;
        st->block_size = 256;
    }
    return HADROS_OK;
}

int hadros_replay_set_mev(hadros_replay* p, const hadros_mev* mev)
{
    if (!p) return fail(HADROS_ERR_ARG, "NULL replay");
    if (p->ran) HIPC(hipEventSynchronize(p->ev1));   // the parameters may still be in use
    return mev_set(p->mev, mev, p->args.rt.med, p->n_rays);
}

int hadros_replay_fetch_mev(hadros_replay* p, hadros_mev_out* o)
{
    if (!p || !p->ran || !o) return fail(HADROS_ERR_ARG, "replay has not run");
    HIPC(hipEventSynchronize(p->ev1));
    return mev_fetch(p->mev, p->ran_mev, p->n_rays, o);
}

int hadros_replay_destroy(hadros_replay* p)
{
    if (!p) return HADROS_OK;
    mev_free(p->mev);
    dfree(p->d_sig);
    dfree(p->d_sigf);
    dfree(p->d_off);
    dfree(p->d_r);
    dfree(p->d_th);
    dfree(p->d_dl);
    dfree(p->d_z);
    dfree(p->d_tau);
    dfree(p->d_p);
    dfree(p->d_I);
    dfree(p->d_bt);
    dfree(p->d_bp);
    dfree(p->d_bi);
    dfree(p->d_cnt);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    delete p;
    return HADROS_OK;
}

int hadros_rt_from_cache(const hadros_cache_view* c, const hadros_scene* scene, const double* se, const double* ss,
                         int n_sigma, int precision, hadros_image_out* out, hadros_stats* stats)
{
    const auto t0 = std::chrono::steady_clock::now();
    hadros_replay* p = nullptr;
    int rc = hadros_replay_create(c, scene, se, ss, n_sigma, precision, &p);
    if (rc) return rc;
    rc = hadros_replay_run(p, nullptr);
    if (!rc) rc = hadros_replay_fetch(p, out, stats);
    hadros_replay_destroy(p);
    if (!rc && stats)
        stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}


// ---------------------------------------------------------------- batched tracer
int hadros_batch_create_sub(const hadros_camera* cams, const hadros_scene* scenes, int n_cams, const double* se,
                            const double* ss, int n_sigma, int precision, int stride, int offset,
                            hadros_batch** out)
{
    if (!cams || !scenes || !out || n_cams <= 0) return fail(HADROS_ERR_ARG, "bad batch arguments");
    if (stride < 1 || offset < 0) return fail(HADROS_ERR_ARG, "sub-lattice stride must be >= 1 and offset >= 0");
    if (precision != HADROS_PREC_FP64 && precision != HADROS_PREC_FP32K) return fail(HADROS_ERR_ARG, "unknown precision");
    for (int c = 1; c < n_cams; ++c)
        if (cams[c].mode != cams[0].mode) return fail(HADROS_ERR_ARG, "all cameras of a batch must share one integrator mode");
    int rc = ensure_device();
    if (rc) return rc;
    hadros_batch* b = new hadros_batch();
    std::vector<DevCam> dc(static_cast<size_t>(n_cams));
    std::vector<DevRt> dr(static_cast<size_t>(n_cams));
    b->pix0.assign(size_t(n_cams) + 1, 0);
    b->nbc.assign(size_t(n_cams), 0);
    for (int c = 0; c < n_cams; ++c) {
        if ((rc = make_cam(cams[c], dc[c])) || (rc = make_rt(scenes[c], se, ss, n_sigma, cams[c].a_spin, dr[c]))) {
            delete b;
            return rc;
        }
        dc[c].pad_rt = c;
        dc[c].sub_stride = stride;
        dc[c].sub_offset = offset;
        dc[c].sub_nx = cams[c].nx > offset ? (cams[c].nx - offset + stride - 1) / stride : 0;
        dc[c].sub_ny = cams[c].ny > offset ? (cams[c].ny - offset + stride - 1) / stride : 0;
        b->nbc[c] = dr[c].nb;
        b->nb_max = std::max(b->nb_max, dr[c].nb);
        b->pix0[c + 1] = b->pix0[c] + long(dc[c].sub_nx) * dc[c].sub_ny;
    }
    b->mode = cams[0].mode;
    b->precision = precision;
    b->n_rays = b->pix0[n_cams];
    std::vector<double> sd;
    std::vector<float> sf;
    sigma_tables(se, ss, n_sigma, sd, sf);
    const size_t nr = size_t(b->n_rays), nb = size_t(b->nb_max);
    if ((rc = dalloc(&b->d_cams, size_t(n_cams))) || (rc = dalloc(&b->d_pix0, size_t(n_cams) + 1)) ||
        (rc = dalloc(&b->d_rts, size_t(n_cams))) || (rc = dalloc(&b->d_sig, sd.size())) ||
        (rc = dalloc(&b->d_sigf, sf.size())) || (rc = dalloc(&b->d_queue, 1)) || (rc = dalloc(&b->d_cnt, 4)) ||
        (rc = dalloc(&b->d_tau, nr)) || (rc = dalloc(&b->d_p, nr)) || (rc = dalloc(&b->d_I, nr)) ||
        (rc = dalloc(&b->d_bt, nr * nb)) || (rc = dalloc(&b->d_bp, nr * nb)) || (rc = dalloc(&b->d_bi, nr * nb)) ||
        (rc = dalloc(&b->d_cap, nr)) || (rc = dalloc(&b->d_np, nr)) || (rc = dalloc(&b->d_flags, nr))) {
        hadros_batch_destroy(b);
        return rc;
    }
    HIPC(hipMemcpy(b->d_cams, dc.data(), dc.size() * sizeof(DevCam), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(b->d_pix0, b->pix0.data(), b->pix0.size() * sizeof(long), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(b->d_rts, dr.data(), dr.size() * sizeof(DevRt), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(b->d_sig, sd.data(), sd.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(b->d_sigf, sf.data(), sf.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPC(hipEventCreate(&b->ev0));
    HIPC(hipEventCreate(&b->ev1));
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    HIPC(hipGetDeviceProperties(&prop, dev));
    b->cu = prop.multiProcessorCount;
    BatchArgs& A = b->args;
    A.cams = b->d_cams;
    A.pix0 = b->d_pix0;
    A.rts = b->d_rts;
    A.sig = b->d_sig;
    A.sigf = b->d_sigf;
    A.n_sigma = n_sigma;
    A.n_cams = n_cams;
    A.nb_out = b->nb_max;
    A.steps_per_chunk = b->spc;
    A.n_rays = b->n_rays;
    A.queue = b->d_queue;
    A.out = TraceOut{b->d_tau, b->d_p, b->d_I, nullptr, nullptr, b->d_bt, b->d_bp, b->d_bi,
                     b->d_cap, b->d_np, b->d_flags, b->d_cnt};
    *out = b;
    return HADROS_OK;
}

int hadros_batch_create(const hadros_camera* cams, const hadros_scene* scenes, int n_cams, const double* se,
                        const double* ss, int n_sigma, int precision, hadros_batch** out)
{
    return hadros_batch_create_sub(cams, scenes, n_cams, se, ss, n_sigma, precision, 1, 0, out);
}

int hadros_batch_fetch_all(hadros_batch* b, hadros_image_out* o)
{
    if (!b || !b->ran || !o) return fail(HADROS_ERR_ARG, "batch has not run");
    {
        const int rs = batch_settle(b);
        if (rs) return rs;
    }
    const size_t n = size_t(b->n_rays), nbo = size_t(b->nb_max);
    auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
        if (dst && src && bytes) HIPC(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return HADROS_OK;
    };
    int rc = 0;
    if ((rc = cp(o->tau, b->d_tau, n * 8)) || (rc = cp(o->p_surv, b->d_p, n * 8)) ||
        (rc = cp(o->intensity, b->d_I, n * 8)) || (rc = cp(o->captured, b->d_cap, n * 4)) ||
        (rc = cp(o->npoints, b->d_np, n * 4)) || (rc = cp(o->flags, b->d_flags, n * 4)) ||
        (rc = cp(o->bin_tau, b->d_bt, n * nbo * 8)) || (rc = cp(o->bin_p, b->d_bp, n * nbo * 8)) ||
        (rc = cp(o->bin_i, b->d_bi, n * nbo * 8)))
        return rc;
    return HADROS_OK;
}

int hadros_batch_fetch_all_device(hadros_batch* b, const hadros_image_out* o, void* stream)
{
    if (!b || !b->ran || !o) return fail(HADROS_ERR_ARG, "batch has not run");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (b->split) {   // a split run may have to grow its pool and run again
        const int rs = batch_settle(b);
        if (rs) return rs;
    }
    HIPC(hipStreamWaitEvent(s, b->ev1, 0));
    const size_t n = size_t(b->n_rays), nbo = size_t(b->nb_max);
    auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
        if (dst && src && bytes) HIPC(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
        return HADROS_OK;
    };
    int rc = 0;
    if ((rc = cp(o->tau, b->d_tau, n * 8)) || (rc = cp(o->p_surv, b->d_p, n * 8)) ||
        (rc = cp(o->intensity, b->d_I, n * 8)) || (rc = cp(o->captured, b->d_cap, n * 4)) ||
        (rc = cp(o->npoints, b->d_np, n * 4)) || (rc = cp(o->flags, b->d_flags, n * 4)) ||
        (rc = cp(o->bin_tau, b->d_bt, n * nbo * 8)) || (rc = cp(o->bin_p, b->d_bp, n * nbo * 8)) ||
        (rc = cp(o->bin_i, b->d_bi, n * nbo * 8)))
        return rc;
    return HADROS_OK;
}

int hadros_batch_set_tuning(hadros_batch* b, int block, int steps_per_chunk)
{
    if (!b) return fail(HADROS_ERR_ARG, "NULL batch");
    if (block < 64 || block > 256 || block % 64) return fail(HADROS_ERR_ARG, "block size must be 64, 128, 192 or 256");
    if (steps_per_chunk < 1) return fail(HADROS_ERR_ARG, "steps_per_chunk must be >= 1");
    b->block = block;
    b->spc = steps_per_chunk;
    b->args.steps_per_chunk = steps_per_chunk;
    return HADROS_OK;
}

namespace {

constexpr double CHUNK_BYTES = double(CHUNK_DOUBLES) * sizeof(double);
// default bound of the split batch's chunk pool
// (64 GB of an MI355X's 288: the 256-camera batch, 3.6e9 points, runs in
// fewer sub-batches than at 24 GB and 2 % faster; a 160 GB pool gains 1 % more)
constexpr double BATCH_POOL_MAX_BYTES = 64.0e9;
// chunks per ray the pool is first sized for: the ray's own and two more
constexpr long BATCH_CHUNKS_PER_RAY = 3;
// sub-batch rays are int indices in the kernels
constexpr long BATCH_SUB_MAX = 1L << 30;

// the split form only on request: with the queue refilling idle lanes the
// fused kernel keeps its lanes as busy as the replay does, and on the
// 256-camera workload it is the faster of the two (508 vs 516 ms)
// HADROS_PATH_AUTO: split in CFG mode (the sub-batches bound the pool), fused
// for REF-mode rays, whose ~1e4 points each would leave few rays per sub-batch
bool batch_want_split(const hadros_batch* b)
{
    return b->path == HADROS_PATH_SPLIT || (b->path == HADROS_PATH_AUTO && b->mode == HADROS_MODE_CFG);
}

double batch_pool_bound(const hadros_batch* b) { return b->pool_max > 0.0 ? b->pool_max : BATCH_POOL_MAX_BYTES; }

// (re)allocate the pool for sub-batches of `sub` rays and `cap` chunks, in
// `halves` halves of that size (and the second stream and events they use)
int batch_pool(hadros_batch* b, long sub, long cap, int halves)
{
    dfree(b->d_pool);
    dfree(b->d_next);
    dfree(b->d_nst);
    b->capacity = 0;
    b->sub = 0;
    int rc = 0;
    const size_t h = size_t(halves);
    if ((rc = dalloc(&b->d_pool, h * size_t(cap) * CHUNK_DOUBLES)) || (rc = dalloc(&b->d_next, h * size_t(cap))) ||
        (rc = dalloc(&b->d_nst, h * size_t(sub))))
        return rc;
    if (halves == 2 && !b->s2) {
        HIPC(hipStreamCreateWithFlags(&b->s2, hipStreamNonBlocking));   // default priority
        for (int k = 0; k < 2; ++k) {
            HIPC(hipEventCreateWithFlags(&b->evg[k], hipEventDisableTiming));
            HIPC(hipEventCreateWithFlags(&b->evr[k], hipEventDisableTiming));
        }
    }
    b->capacity = cap;
    b->sub = sub;
    b->halves = halves;
    return HADROS_OK;
}

// first sizing: as many rays per sub-batch as BATCH_CHUNKS_PER_RAY chunks each
// fit in the bound -- or in half of it, when one sub-batch cannot hold them all
int batch_pool_init(hadros_batch* b)
{
    const double bound = batch_pool_bound(b);
    long fit = long(bound / (CHUNK_BYTES * double(BATCH_CHUNKS_PER_RAY)));
    int halves = 1;
    if (b->n_rays > fit) {
        halves = 2;
        fit = long(0.5 * bound / (CHUNK_BYTES * double(BATCH_CHUNKS_PER_RAY)));
    }
    const long sub = std::max(1L, std::min(std::min(b->n_rays, fit), BATCH_SUB_MAX));
    return batch_pool(b, sub, sub * BATCH_CHUNKS_PER_RAY, halves);
}

// the first ray_limit rays as sub-batches: k_batch_points into the pool, then
// k_replay of the same rays, all queued on stream s
int batch_launch_split(hadros_batch* b, hipStream_t s, long ray_limit)
{
    int rc = 0;
    if (!b->d_pool && (rc = batch_pool_init(b))) return rc;
    const long n_sub = std::max(1L, (ray_limit + b->sub - 1) / b->sub);
    if (n_sub > b->n_ctr) {
        dfree(b->d_ctr);
        b->n_ctr = 0;
        if ((rc = dalloc(&b->d_ctr, size_t(n_sub)))) return rc;
        b->n_ctr = int(n_sub);
    }
    int per_cu = 0;
    HIPC(batch_points_resident_blocks(b->mode, b->block, b->args.n_cams, &per_cu));
    const int grid = std::max(1, per_cu) * b->cu;
    HIPC(launch_zero_u64(b->d_cnt, 4, s));
    HIPC(launch_zero_u64(b->d_ctr, int(n_sub), s));
    HIPC(hipEventRecord(b->ev0, s));
    const bool two = b->halves == 2;
    hipStream_t sr = two ? b->s2 : s;   // the replays' stream
    int last_half = 0;
    for (long i = 0; i < n_sub; ++i) {
        const long r0 = i * b->sub, n = std::min(b->sub, ray_limit - r0);
        if (n <= 0) break;
        const int hf = two ? int(i & 1) : 0;
        double* pool = b->d_pool + size_t(hf) * size_t(b->capacity) * CHUNK_DOUBLES;
        int* next = b->d_next + size_t(hf) * size_t(b->capacity);
        int* nst = b->d_nst + size_t(hf) * size_t(b->sub);
        // the half is free once the replay of the sub-batch before last is done
        if (two && i >= 2) HIPC(hipStreamWaitEvent(s, b->evr[hf], 0));
        BatchPointsArgs Q{};
        Q.cams = b->d_cams;
        Q.pix0 = b->d_pix0;
        Q.n_cams = b->args.n_cams;
        Q.steps_per_chunk = b->spc;
        Q.ray0 = r0;
        Q.n_rays = n;
        Q.queue = b->d_queue;
        Q.cp = ChunkPool{pool, next, reinterpret_cast<unsigned*>(b->d_ctr + i), n, b->capacity};
        Q.n_st = nst;
        Q.out = b->args.out;
        HIPC(launch_zero_u64(b->d_queue, 1, s));
        HIPC(launch_batch_points(Q, b->mode, grid, b->block, s));
        if (two) {
            HIPC(hipEventRecord(b->evg[hf], s));
            HIPC(hipStreamWaitEvent(sr, b->evg[hf], 0));
        }
        ReplayArgs R{};
        R.rt.n_sigma = b->args.n_sigma;
        R.sig = b->d_sig;
        R.sigf = b->d_sigf;
        R.n_rays = int(n);
        R.pool = pool;
        R.next = next;
        R.n_st = nst;
        R.rts = b->d_rts;
        R.cams = b->d_cams;
        R.pix0 = b->d_pix0;
        R.n_cams = b->args.n_cams;
        R.nb_out = b->nb_max;
        R.ray0 = r0;
        R.tau = b->d_tau;
        R.p = b->d_p;
        R.I = b->d_I;
        R.btau = b->d_bt;
        R.bp = b->d_bp;
        R.bi = b->d_bi;
        HIPC(launch_replay(R, b->precision, sr));
        if (two) HIPC(hipEventRecord(b->evr[hf], sr));
        last_half = hf;
    }
    if (two) HIPC(hipStreamWaitEvent(s, b->evr[last_half], 0));   // the run ends when its last replay does
    HIPC(hipEventRecord(b->ev1, s));
    b->last_limit = ray_limit;
    return HADROS_OK;
}

// after a split run: a sub-batch whose rays ran out of chunks leaves its
// counter at the chunks it needed; grow the pool to that (past the bound,
// with fewer rays per sub-batch) and run again
int batch_settle(hadros_batch* b)
{
    HIPC(hipEventSynchronize(b->ev1));
    if (!b->split) return HADROS_OK;
    for (int attempt = 0; attempt < 4; ++attempt) {
        const long n_sub = std::max(1L, (b->last_limit + b->sub - 1) / b->sub);
        std::vector<unsigned long long> ctr(static_cast<size_t>(n_sub));
        HIPC(hipMemcpy(ctr.data(), b->d_ctr, ctr.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        long need = 0;   // chunks of the most demanding sub-batch
        for (long i = 0; i < n_sub; ++i) {
            const long n = std::min(b->sub, b->last_limit - i * b->sub);
            need = std::max(need, n + long(unsigned(ctr[size_t(i)])));
        }
        if (need <= b->capacity) return HADROS_OK;
        if (attempt == 3) break;
        long cap = long(double(need) * 1.25) + 1, sub = b->sub;
        const double bound = batch_pool_bound(b) / double(b->halves);   // per half
        if (double(cap) * CHUNK_BYTES > bound) {
            sub = std::max(1L, long(double(sub) * bound / (double(cap) * CHUNK_BYTES)));
            cap = std::max(sub + 1, long(bound / CHUNK_BYTES));
        }
        int rc = 0;
        if ((rc = batch_pool(b, sub, cap, b->halves)) || (rc = batch_launch_split(b, nullptr, b->last_limit))) return rc;
        HIPC(hipEventSynchronize(b->ev1));
    }
    return fail(HADROS_ERR_MAX_POINTS, "split batch: chunk pool too small");
}

}  // namespace

static int batch_launch(hadros_batch* b, hipStream_t s, long ray_limit)
{
    b->split = batch_want_split(b);
    if (b->split) {
        const int rc = batch_launch_split(b, s, ray_limit);
        if (!rc) b->ran = true;
        return rc;
    }
    int per_cu = 0;
    HIPC(batch_resident_blocks(b->mode, b->precision, b->nb_max, b->block, b->args.n_sigma, b->args.n_cams, &per_cu));
    const int grid = std::max(1, per_cu) * b->cu;
    BatchArgs A = b->args;
    A.n_rays = ray_limit;
    HIPC(launch_zero_u64(b->d_queue, 1, s));
    HIPC(launch_zero_u64(b->d_cnt, 4, s));
    HIPC(hipEventRecord(b->ev0, s));
    HIPC(launch_trace_batch(A, b->mode, b->precision, b->nb_max, grid, b->block, s));
    HIPC(hipEventRecord(b->ev1, s));
    b->ran = true;
    return HADROS_OK;
}

int hadros_batch_run(hadros_batch* b, void* stream)
{
    if (!b) return fail(HADROS_ERR_ARG, "NULL batch");
    return batch_launch(b, reinterpret_cast<hipStream_t>(stream), b->n_rays);
}

int hadros_batch_set_path(hadros_batch* b, int path)
{
    if (!b) return fail(HADROS_ERR_ARG, "NULL batch");
    if (path != HADROS_PATH_AUTO && path != HADROS_PATH_FUSED && path != HADROS_PATH_SPLIT)
        return fail(HADROS_ERR_ARG, "path must be HADROS_PATH_AUTO, _FUSED or _SPLIT");
    b->path = path;
    return HADROS_OK;
}

int hadros_batch_set_pool(hadros_batch* b, double max_bytes)
{
    if (!b) return fail(HADROS_ERR_ARG, "NULL batch");
    if (!(max_bytes >= 0.0)) return fail(HADROS_ERR_ARG, "pool bound must be >= 0 bytes");
    if (max_bytes > 0.0 && max_bytes < 2.0 * CHUNK_BYTES) return fail(HADROS_ERR_ARG, "pool bound below two chunks");
    if (b->ran) HIPC(hipEventSynchronize(b->ev1));   // the pool may still be in use
    b->pool_max = max_bytes;
    dfree(b->d_pool);   // sized again at the next run
    dfree(b->d_next);
    dfree(b->d_nst);
    b->capacity = 0;
    b->sub = 0;
    return HADROS_OK;
}

int hadros_batch_last_ms(hadros_batch* b, double* ms)
{
    if (!b || !ms || !b->ran) return fail(HADROS_ERR_ARG, "batch has not run");
    {
        const int rs = batch_settle(b);
        if (rs) return rs;
    }
    float f = 0.0f;
    HIPC(hipEventElapsedTime(&f, b->ev0, b->ev1));
    *ms = f;
    return HADROS_OK;
}

// Times every (block, steps_per_chunk) pair on the first eighth of the rays
// (at least 2^18), after one untimed run, then times the three fastest twice
// more and keeps the pair with the least total.
int hadros_batch_autotune(hadros_batch* b, int* block_out, int* spc_out)
{
    if (!b) return fail(HADROS_ERR_ARG, "NULL batch");
    const long limit = std::min(b->n_rays, std::max(b->n_rays / 8, 1L << 18));
    const int blocks[] = {64, 128, 256};
    const int spcs[] = {1, 4, 16, 64};
    struct Cand {
        int blk, spc;
        double ms;
    };
    std::vector<Cand> cands;
    auto time_pair = [&](int blk, int spc, int runs, bool warm, double* total) {
        b->block = blk;
        b->spc = spc;   // the split form's points kernel
        b->args.steps_per_chunk = spc;   // the fused kernel
        double ms = 0.0;
        int rc = 0;
        if (warm) {
            rc = batch_launch(b, nullptr, limit);
            if (!rc) rc = hadros_batch_last_ms(b, &ms);
        }
        *total = 0.0;
        for (int k = 0; k < runs && !rc; ++k) {
            rc = batch_launch(b, nullptr, limit);
            if (!rc) rc = hadros_batch_last_ms(b, &ms);
            *total += ms;
        }
        return rc;
    };
    for (int blk : blocks)
        for (int spc : spcs) {
            double ms = 0.0;
            if (const int rc = time_pair(blk, spc, 1, true, &ms)) return rc;
            cands.push_back({blk, spc, ms});
        }
    std::sort(cands.begin(), cands.end(), [](const Cand& x, const Cand& y) { return x.ms < y.ms; });
    cands.resize(3);
    for (Cand& c : cands) {
        double ms = 0.0;
        if (const int rc = time_pair(c.blk, c.spc, 2, false, &ms)) return rc;
        c.ms += ms;
    }
    const Cand best = *std::min_element(cands.begin(), cands.end(),
                                        [](const Cand& x, const Cand& y) { return x.ms < y.ms; });
    b->block = best.blk;
    b->spc = best.spc;
    b->args.steps_per_chunk = best.spc;
    b->ran = false;
    if (block_out) *block_out = best.blk;
    if (spc_out) *spc_out = best.spc;
    return HADROS_OK;
}

int hadros_batch_fetch(hadros_batch* b, int cam, hadros_image_out* o)
{
    if (!b || !b->ran) return fail(HADROS_ERR_ARG, "batch has not run");
    if (cam < 0 || cam >= b->args.n_cams) return fail(HADROS_ERR_ARG, "camera index out of range");
    {
        const int rs = batch_settle(b);
        if (rs) return rs;
    }
    const size_t k0 = size_t(b->pix0[cam]), n = size_t(b->pix0[cam + 1] - b->pix0[cam]);
    const size_t nbo = size_t(b->nb_max), nbc = size_t(std::max(b->nbc[cam], 0));
    auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
        if (dst && src && bytes) HIPC(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        return HADROS_OK;
    };
    int rc = 0;
    if ((rc = cp(o->tau, b->d_tau + k0, n * 8)) || (rc = cp(o->p_surv, b->d_p + k0, n * 8)) ||
        (rc = cp(o->intensity, b->d_I + k0, n * 8)) || (rc = cp(o->captured, b->d_cap + k0, n * 4)) ||
        (rc = cp(o->npoints, b->d_np + k0, n * 4)) || (rc = cp(o->flags, b->d_flags + k0, n * 4)))
        return rc;
    if (nbc && (o->bin_tau || o->bin_p || o->bin_i)) {
        std::vector<double> t(n * nbo);
        double* dsts[3] = {o->bin_tau, o->bin_p, o->bin_i};
        const double* srcs[3] = {b->d_bt, b->d_bp, b->d_bi};
        for (int q = 0; q < 3; ++q) {
            if (!dsts[q]) continue;
            HIPC(hipMemcpy(t.data(), srcs[q] + k0 * nbo, n * nbo * 8, hipMemcpyDeviceToHost));
            for (size_t k = 0; k < n; ++k)
                for (size_t c = 0; c < nbc; ++c) dsts[q][k * nbc + c] = t[k * nbo + c];
        }
    }
    if (o->alpha || o->beta) {
        // tan-plane coordinates of the camera's pixels (src/kerr_camera.cpp:196-200)
        DevCam dc;
        HIPC(hipMemcpy(&dc, b->d_cams + cam, sizeof(DevCam), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; ++k) {
            const int i = dc.sub_offset + dc.sub_stride * int(k / size_t(dc.sub_ny));
            const int j = dc.sub_offset + dc.sub_stride * int(k % size_t(dc.sub_ny));
            if (o->alpha) o->alpha[k] = (2.0 * (i + 0.5) / dc.nx - 1.0) * dc.tanhalf;
            if (o->beta) o->beta[k] = (2.0 * (j + 0.5) / dc.ny - 1.0) * dc.tanhalf;
        }
    }
    return HADROS_OK;
}

int hadros_batch_stats(hadros_batch* b, hadros_stats* st)
{
    if (!b || !st || !b->ran) return fail(HADROS_ERR_ARG, "batch has not run");
    {
        const int rs = batch_settle(b);
        if (rs) return rs;
    }
    unsigned long long c[4];
    HIPC(hipMemcpy(c, b->d_cnt, sizeof(c), hipMemcpyDeviceToHost));
    float f = 0.0f;
    HIPC(hipEventElapsedTime(&f, b->ev0, b->ev1));
    st->kernel_ms = f;
    st->total_ms = f;
    st->total_points = (int64_t)c[0];
    st->total_attempts = (int64_t)c[1];
    st->n_captured = int32_t(c[2]);
    st->n_exhausted = int32_t(c[3]);
    st->n_bins = b->nb_max;
    st->block_size = b->block;
    return HADROS_OK;
}

int hadros_batch_info(hadros_batch* b, int64_t* n_rays, int* block, int* steps_per_chunk, int* grid)
{
    if (!b) return fail(HADROS_ERR_ARG, "NULL batch");
    int per_cu = 0;
    if (batch_want_split(b)) HIPC(batch_points_resident_blocks(b->mode, b->block, b->args.n_cams, &per_cu));
    else HIPC(batch_resident_blocks(b->mode, b->precision, b->nb_max, b->block, b->args.n_sigma, b->args.n_cams, &per_cu));
    if (n_rays) *n_rays = b->n_rays;
    if (block) *block = b->block;
    if (steps_per_chunk) *steps_per_chunk = b->spc;
    if (grid) *grid = std::max(1, per_cu) * b->cu;
    return HADROS_OK;
}

int hadros_batch_destroy(hadros_batch* b)
{
    if (!b) return HADROS_OK;
    dfree(b->d_cams);
    dfree(b->d_pix0);
    dfree(b->d_rts);
    dfree(b->d_sig);
    dfree(b->d_sigf);
    dfree(b->d_queue);
    dfree(b->d_cnt);
    dfree(b->d_tau);
    dfree(b->d_p);
    dfree(b->d_I);
    dfree(b->d_bt);
    dfree(b->d_bp);
    dfree(b->d_bi);
    dfree(b->d_cap);
    dfree(b->d_np);
    dfree(b->d_flags);
    dfree(b->d_pool);
    dfree(b->d_next);
    dfree(b->d_nst);
    dfree(b->d_ctr);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    for (int k = 0; k < 2; ++k) {
        if (b->evg[k]) (void)hipEventDestroy(b->evg[k]);
        if (b->evr[k]) (void)hipEventDestroy(b->evr[k]);
    }
    if (b->s2) (void)hipStreamDestroy(b->s2);
    delete b;
    return HADROS_OK;
}

// ---------------------------------------------------------------- geodesic cache
int hadros_trace_cache_range(const hadros_camera* cam, int64_t k0, int64_t k1, int64_t* offset, int32_t* captured,
                             double* points, int64_t cap)
{
    if (!cam || !offset) return fail(HADROS_ERR_ARG, "NULL argument");
    int rc = ensure_device();
    if (rc) return rc;
    CacheArgs A{};
    rc = make_cam(*cam, A.cam);
    if (rc) return rc;
    const long npix = long(cam->nx) * long(cam->ny);
    if (k0 < 0 || k1 > npix || k1 < k0) return fail(HADROS_ERR_ARG, "pixel range outside the camera");
    A.k0 = long(k0);
    A.k1 = long(k1);
    const size_t np = size_t(k1 - k0);
    if (np == 0) {
        offset[0] = 0;
        return HADROS_OK;
    }
    long* d_counts = nullptr;
    int* d_cap = nullptr;
    long* d_off = nullptr;
    double* d_pts = nullptr;
    auto cleanup = [&]() {
        dfree(d_counts);
        dfree(d_cap);
        dfree(d_off);
        dfree(d_pts);
    };
    if ((rc = dalloc(&d_counts, np)) || (rc = dalloc(&d_cap, np)) || (rc = dalloc(&d_off, np + 1))) {
        cleanup();
        return rc;
    }
    // pass 1: point counts (the trace is deterministic, so pass 2 reproduces it)
    A.counts = d_counts;
    A.captured = d_cap;
    hipError_t e = launch_trace_cache(A, false, nullptr);
    std::vector<long> counts(np);
    std::vector<int> capv(np);
    if (e == hipSuccess) e = hipMemcpy(counts.data(), d_counts, np * sizeof(long), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(capv.data(), d_cap, np * sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        cleanup();
        return fail(HADROS_ERR_DEVICE, std::string("trace_cache count pass: ") + hipGetErrorString(e));
    }
    std::vector<long> off(np + 1, 0);
    for (size_t k = 0; k < np; ++k) off[k + 1] = off[k] + counts[k];
    for (size_t k = 0; k <= np; ++k) offset[k] = off[k];
    if (captured)
        for (size_t k = 0; k < np; ++k) captured[k] = capv[k];
    if (points) {
        if (cap < off[np]) {
            cleanup();
            return fail(HADROS_ERR_MAX_POINTS, "point buffer smaller than the traced point count");
        }
        if ((rc = dalloc(&d_pts, size_t(off[np]) * 7))) {
            cleanup();
            return rc;
        }
        e = hipMemcpy(d_off, off.data(), (np + 1) * sizeof(long), hipMemcpyHostToDevice);
        A.offset = d_off;
        A.points = d_pts;
        if (e == hipSuccess) e = launch_trace_cache(A, true, nullptr);
        if (e == hipSuccess && off[np] > 0)
            e = hipMemcpy(points, d_pts, size_t(off[np]) * 7 * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    cleanup();
    if (e != hipSuccess) return fail(HADROS_ERR_DEVICE, std::string("trace_cache: ") + hipGetErrorString(e));
    return HADROS_OK;
}

int hadros_trace_cache(const hadros_camera* cam, int64_t* offset, int32_t* captured, double* points, int64_t cap)
{
    if (!cam) return fail(HADROS_ERR_ARG, "NULL camera");
    return hadros_trace_cache_range(cam, 0, int64_t(cam->nx) * cam->ny, offset, captured, points, cap);
}

}  // extern "C"

// ---------------------------------------------------------------- opacity surfaces
extern "C" int hadros_opacity_surface(const hadros_medium* m, const double* se, const double* ss, int n_sigma,
                                      double enu_GeV, double m_bh_msun, double tau_surface, int nr, int ntheta,
                                      double* theta, double* r_tau, double* tau_total, double* classification,
                                      int* crossing)
{
    if (!m || !theta || !r_tau || !tau_total || !classification || !crossing)
        return fail(HADROS_ERR_ARG, "NULL argument");
    if (ntheta < 2 || nr < 4 || m->r_max <= m->r_min)
        return fail(HADROS_ERR_ARG, "Invalid opacity-surface grid parameters.");
    if (m->kind < 0 || m->kind > 7) return fail(HADROS_ERR_KIND, "unknown density profile kind");
    int rc = check_sigma(se, ss, n_sigma);
    if (rc) return rc;
    if (enu_GeV < se[0] || enu_GeV > se[n_sigma - 1]) return fail(HADROS_ERR_RANGE, "Energy outside sigma table range.");
    if ((rc = ensure_device())) return rc;
    std::vector<double> sd;
    std::vector<float> sf;
    sigma_tables(se, ss, n_sigma, sd, sf);
    OpacityArgs A{};
    A.m = make_medium(*m);
    A.n_sigma = n_sigma;
    A.nr = nr;
    A.ntheta = ntheta;
    sigma_guess(se, n_sigma, A.sig_lo, A.sig_inv_d);
    A.ln_enu = std::log(enu_GeV);
    A.rg_cm = G_CGS * m_bh_msun * MSUN_G / (C_CGS * C_CGS);   // include/constants.hpp:17-20
    A.r_min = m->r_min;
    A.dr = (m->r_max - m->r_min) / static_cast<double>(nr - 1);
    A.tau_s = tau_surface;
    double* d_sig = nullptr;
    double* d_out = nullptr;   // theta, r_tau, tau_total, cls
    int* d_cross = nullptr;
    auto cleanup = [&]() {
        dfree(d_sig);
        dfree(d_out);
        dfree(d_cross);
    };
    const size_t nt = size_t(ntheta);
    if ((rc = dalloc(&d_sig, sd.size())) || (rc = dalloc(&d_out, 4 * nt)) || (rc = dalloc(&d_cross, nt))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpy(d_sig, sd.data(), sd.size() * 8, hipMemcpyHostToDevice);
    A.sig = d_sig;
    A.theta = d_out;
    A.r_tau = d_out + nt;
    A.tau_total = d_out + 2 * nt;
    A.cls = d_out + 3 * nt;
    A.crossing = d_cross;
    if (e == hipSuccess) e = launch_opacity_surface(A, nullptr);
    std::vector<double> h(4 * nt);
    if (e == hipSuccess) e = hipMemcpy(h.data(), d_out, 4 * nt * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(crossing, d_cross, nt * sizeof(int), hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return fail(HADROS_ERR_DEVICE, hipGetErrorString(e));
    std::memcpy(theta, h.data(), nt * 8);
    std::memcpy(r_tau, h.data() + nt, nt * 8);
    std::memcpy(tau_total, h.data() + 2 * nt, nt * 8);
    std::memcpy(classification, h.data() + 3 * nt, nt * 8);
    return HADROS_OK;
}

// ---------------------------------------------------------------- point-wise hooks
namespace {
template <class F>
int run_eval(const std::vector<std::pair<const double*, size_t>>& ins, double* out, size_t out_n, F&& launch)
{
    int rc = ensure_device();
    if (rc) return rc;
    std::vector<double*> d(ins.size(), nullptr);
    double* d_out = nullptr;
    auto cleanup = [&]() {
        for (auto& p : d) dfree(p);
        dfree(d_out);
    };
    for (size_t k = 0; k < ins.size(); ++k) {
        if ((rc = dalloc(&d[k], ins[k].second))) {
            cleanup();
            return rc;
        }
        if (ins[k].second) {
            hipError_t e = hipMemcpy(d[k], ins[k].first, ins[k].second * 8, hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                cleanup();
                return fail(HADROS_ERR_DEVICE, hipGetErrorString(e));
            }
        }
    }
    if ((rc = dalloc(&d_out, out_n))) {
        cleanup();
        return rc;
    }
    hipError_t e = launch(d, d_out);
    if (e == hipSuccess && out_n) e = hipMemcpy(out, d_out, out_n * 8, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return fail(HADROS_ERR_DEVICE, hipGetErrorString(e));
    return HADROS_OK;
}
}  // namespace

extern "C" {

int hadros_eval_rho(const hadros_medium* m, const double* r, const double* th, double* out, int64_t n)
{
    if (!m || n < 0) return fail(HADROS_ERR_ARG, "bad argument");
    if (m->kind < 0 || m->kind > 7) return fail(HADROS_ERR_KIND, "unknown density profile kind");
    const DevMedium dm = make_medium(*m);
    return run_eval({{r, size_t(n)}, {th, size_t(n)}}, out, size_t(n), [&](std::vector<double*>& d, double* o) {
        return launch_eval_rho(dm, d[0], d[1], o, long(n));
    });
}

int hadros_eval_sigma(const double* se, const double* ss, int n_sigma, const double* e, double* out, int64_t n)
{
    int rc = check_sigma(se, ss, n_sigma);
    if (rc) return rc;
    std::vector<double> sd;
    std::vector<float> sf;
    sigma_tables(se, ss, n_sigma, sd, sf);
    const double lo = se[0], hi = se[n_sigma - 1];
    double lo0 = 0.0, inv_d = 0.0;
    sigma_guess(se, n_sigma, lo0, inv_d);
    return run_eval({{sd.data(), sd.size()}, {e, size_t(n)}}, out, size_t(n),
                    [&](std::vector<double*>& d, double* o) {
                        return launch_eval_sigma(d[0], n_sigma, lo, hi, lo0, inv_d, d[1], o, long(n));
                    });
}

int hadros_eval_emissivity(const hadros_scene* scene, const double* r, const double* th, const double* e, double* out,
                           int64_t n)
{
    if (!scene) return fail(HADROS_ERR_ARG, "NULL scene");
    // a two-point dummy table: emissivity does not use sigma
    const double te[2] = {1.0, 2.0}, ts[2] = {1.0, 1.0};
    DevRt P;
    int rc = make_rt(*scene, te, ts, 2, 0.0, P);
    if (rc) return rc;
    return run_eval({{r, size_t(n)}, {th, size_t(n)}, {e, size_t(n)}}, out, size_t(n),
                    [&](std::vector<double*>& d, double* o) { return launch_eval_emis(P, d[0], d[1], d[2], o, long(n)); });
}

int hadros_eval_mev(const hadros_scene* scene, const hadros_mev* mev, const double* r, const double* th,
                    const double* g, double* out, int64_t n)
{
    if (!scene || !mev) return fail(HADROS_ERR_ARG, "NULL argument");
    int rc = check_kinds(*scene);
    if (rc || (rc = check_mev(*mev))) return rc;
    const DevMedium med = make_medium(scene->medium);
    const DevMev M = make_mev(*mev, med);
    return run_eval({{r, size_t(n)}, {th, size_t(n)}, {g, size_t(n)}}, out, size_t(n) * 4,
                    [&](std::vector<double*>& d, double* o) { return launch_eval_mev(med, M, d[0], d[1], d[2], o, long(n)); });
}

int hadros_eval_step(double a, int kind, double h, const double* y, double* out, int64_t n)
{
    if (kind < EVAL_STEP_RK4 || kind > EVAL_STEP_RKF45_CFG_SHIFT) return fail(HADROS_ERR_ARG, "unknown step kind");
    return run_eval({{y, size_t(n) * 8}}, out, size_t(n) * 11,
                    [&](std::vector<double*>& d, double* o) { return launch_eval_step(a, kind, h, d[0], o, long(n)); });
}

int hadros_eval_transfer_point(const hadros_scene* scene, const double* se, const double* ss, int n_sigma,
                               double a_bh, int precision, const double* r, const double* th, const double* dl,
                               const double* z, double* out, int64_t n)
{
    if (!scene || n < 0) return fail(HADROS_ERR_ARG, "bad argument");
    if (precision != HADROS_PREC_FP64 && precision != HADROS_PREC_FP32K) return fail(HADROS_ERR_ARG, "unknown precision");
    DevRt P;
    int rc = make_rt(*scene, se, ss, n_sigma, a_bh, P);
    if (rc) return rc;
    std::vector<double> sd;
    std::vector<float> sf;
    sigma_tables(se, ss, n_sigma, sd, sf);
    // the float table travels in double-sized words
    std::vector<double> sf_words((sf.size() + 1) / 2, 0.0);
    std::memcpy(sf_words.data(), sf.data(), sf.size() * sizeof(float));
    return run_eval({{sd.data(), sd.size()}, {sf_words.data(), sf_words.size()}, {r, size_t(n)}, {th, size_t(n)},
                     {dl, size_t(n)}, {z, size_t(n)}},
                    out, size_t(n) * EVAL_TRANSFER_COLS, [&](std::vector<double*>& d, double* o) {
                        return launch_eval_transfer_point(P, precision, d[0], reinterpret_cast<const float*>(d[1]),
                                                          d[2], d[3], d[4], d[5], o, long(n));
                    });
}

int hadros_stream_probe(int kind, double* buf, int64_t n, double* partial, double value, void* stream)
{
    if (kind != 0 && kind != 1) return fail(HADROS_ERR_ARG, "unknown probe kind");
    if (n < 0 || (!buf && n > 0) || (kind == 0 && !partial)) return fail(HADROS_ERR_ARG, "bad argument");
    if (reinterpret_cast<uintptr_t>(buf) % 16 != 0) return fail(HADROS_ERR_ARG, "buffer not 16-byte aligned");
    int rc = ensure_device();
    if (rc) return rc;
    const hipError_t e = launch_stream_probe(kind, buf, long(n), partial, value, HADROS_STREAM_GRID,
                                             static_cast<hipStream_t>(stream));
    return e == hipSuccess ? HADROS_OK : fail(HADROS_ERR_DEVICE, hipGetErrorString(e));
}

int hadros_eval_rhs(double a, int mode, const double* y, double* dy, int64_t n)
{
    if (mode != HADROS_MODE_REF && mode != HADROS_MODE_CFG) return fail(HADROS_ERR_ARG, "unknown mode");
    return run_eval({{y, size_t(n) * 8}}, dy, size_t(n) * 8,
                    [&](std::vector<double*>& d, double* o) { return launch_eval_rhs(a, mode, d[0], o, long(n)); });
}

int hadros_eval_metric(double a, const double* r, const double* th, double* out, int64_t n)
{
    return run_eval({{r, size_t(n)}, {th, size_t(n)}}, out, size_t(n) * 10,
                    [&](std::vector<double*>& d, double* o) { return launch_eval_metric(a, d[0], d[1], o, long(n)); });
}

}  // extern "C"
