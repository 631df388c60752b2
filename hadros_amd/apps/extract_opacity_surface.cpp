// extract_opacity_surface -- axisymmetric opacity surfaces on the GPU.
//
// Drop-in for apps/extract_opacity_surface.cpp: same positional argv
//   ENU_GEV M_BH_MSUN SIGMA_PATH PROFILE RHO0 R0 SIGMA_R H_OVER_R RADIAL_POWER
//   FUNNEL_DEPLETION FUNNEL_THETA_DEG ENVELOPE_RHO0 ENVELOPE_ALPHA R_MIN R_MAX
//   RHO_FLOOR TAU_SURFACE NTHETA NR OUTPUT_PATH
// (defaults :17-37, parsing :72-93), same header lines and columns
// (:131-154, :192-198) and the same "Saved:" line.  The tau(r, theta) integral
// runs in hadros_opacity_surface (one GPU lane per polar angle).  Errors exit
// 1 with the reference's messages; where the reference lets an exception
// escape (bad sigma table, unknown profile) this prints it and exits 1.
#include "cli_common.h"

#include <iomanip>

int main(int argc, char** argv)
{
    using namespace cli;
    const double enu = argf(argc, argv, 1, 1.0e5);
    const double m_bh = argf(argc, argv, 2, 3.0);
    const std::string sigma_path = args(argc, argv, 3, "data/sigma/sigma_nuN_CC_GBW.dat");
    const std::string profile = args(argc, argv, 4, "gaussian");
    hadros_medium m{};
    m.rho0 = argf(argc, argv, 5, 1.0e-2);
    m.r0 = argf(argc, argv, 6, 10.0);
    m.sigma_r = argf(argc, argv, 7, 5.0);
    m.h_over_r = argf(argc, argv, 8, 0.25);
    m.radial_power = argf(argc, argv, 9, 2.0);
    m.funnel_depletion = argf(argc, argv, 10, 0.0);
    const double funnel_theta_deg = argf(argc, argv, 11, 15.0);
    m.funnel_theta_rad = funnel_theta_deg * M_PI / 180.0;
    m.envelope_rho0 = argf(argc, argv, 12, 0.0);
    m.envelope_alpha = argf(argc, argv, 13, 2.5);
    m.r_min = argf(argc, argv, 14, 4.0);
    m.r_max = argf(argc, argv, 15, 60.0);
    m.rho_floor = argf(argc, argv, 16, 1.0e-99);
    const double tau_surface = argf(argc, argv, 17, 1.0);
    const int ntheta = argi(argc, argv, 18, 181);
    const int nr = argi(argc, argv, 19, 1200);
    const std::string out_path = args(argc, argv, 20, "output/opacity_surfaces/tau_surface_tau1.dat");

    if (ntheta < 2 || nr < 4 || m.r_max <= m.r_min) {
        std::cerr << "Invalid opacity-surface grid parameters.\n";
        return 1;
    }
    const Sigma sig = load_sigma(sigma_path);
    if (enu < sig.e.front() || enu > sig.e.back()) {
        std::cerr << "Energy outside sigma table range.\n";
        return 1;
    }
    check(hadros_parse_profile(profile.c_str(), &m.kind), "profile");

    std::ofstream out(out_path);
    if (!out) {
        std::cerr << "Could not open " << out_path << "\n";
        return 1;
    }
    double sigma_cm2 = 0.0;
    check(hadros_eval_sigma(sig.e.data(), sig.s.data(), int(sig.e.size()), &enu, &sigma_cm2, 1), "sigma");
    std::vector<double> theta(static_cast<size_t>(ntheta)), r_tau(theta.size()), tau_total(theta.size()),
        cls(theta.size());
    std::vector<int> crossing(theta.size());
    check(hadros_opacity_surface(&m, sig.e.data(), sig.s.data(), int(sig.e.size()), enu, m_bh, tau_surface, nr,
                                 ntheta, theta.data(), r_tau.data(), tau_total.data(), cls.data(), crossing.data()),
          "opacity surface");

    out << "# opacity_surface axisymmetric_radial\n"
        << "# Enu_GeV " << enu << "\n"
        << "# tau_surface " << tau_surface << "\n"
        << "# profile_type " << profile << "\n"
        << "# rho0 " << m.rho0 << "\n"
        << "# r0 " << m.r0 << "\n"
        << "# sigma_r " << m.sigma_r << "\n"
        << "# H_over_R " << m.h_over_r << "\n"
        << "# radial_power " << m.radial_power << "\n"
        << "# funnel_depletion " << m.funnel_depletion << "\n"
        << "# funnel_theta_deg " << funnel_theta_deg << "\n"
        << "# envelope_rho0 " << m.envelope_rho0 << "\n"
        << "# envelope_alpha " << m.envelope_alpha << "\n"
        << "# R_min " << m.r_min << "\n"
        << "# R_max " << m.r_max << "\n"
        << "# rho_floor " << m.rho_floor << "\n"
        << "# sigma_path " << sigma_path << "\n"
        << "# sigma_cm2 " << sigma_cm2 << "\n"
        << "# M_bh_msun " << m_bh << "\n"
        << "# angular_sampling_uniform_theta " << ntheta << "\n"
        << "# radial_sampling_uniform_rg " << nr << "\n"
        << "# interpolation linear_in_tau_between_radial_samples\n"
        << "# no_crossing_r_tau_rg -1\n"
        << "# future_interface r_tau(theta,phi)\n"
        << "# columns theta_rad theta_deg r_tau_rg tau_total classification crossing_found\n";
    for (int it = 0; it < ntheta; ++it) {
        const size_t k = size_t(it);
        out << std::scientific << std::setprecision(10) << theta[k] << " " << theta[k] * 180.0 / M_PI << " "
            << r_tau[k] << " " << tau_total[k] << " " << cls[k] << " " << crossing[k] << "\n";
    }
    std::cout << "Saved: " << out_path << "\n";
    return 0;
}
