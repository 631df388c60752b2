// compute_kerr_image -- one-pass trace + tau + transfer on the GPU.
//
// Takes the argv of the reference's GPU program compute_kerr_image_cuda
// (usage apps/compute_kerr_image_cuda.cu:2222-2235, parsing :2303-2358):
//   ENU ASPIN MBH_MSUN TORUS_RHO0 TORUS_R0 TORUS_SIGMA TORUS_H_OVER_R
//   SOURCE_R SOURCE_SIGMA SOURCE_THETA_DEG SOURCE_POWERLAW SOURCE_EMAX SOURCE_NORM
//   MEV_ENU MEV_NORM CAM_THETA_DEG CAM_R_OBS CAM_FOV CAM_NX CAM_NY CAM_R_MAX CAM_STEP
//   GPU_TILE GPU_STEPS_PER_LAUNCH GPU_MAX_STEPS GPU_TOL
//   [PROFILE RADIAL_POWER FUNNEL_DEPLETION FUNNEL_THETA_DEG ENVELOPE_RHO0 ENVELOPE_ALPHA
//    R_MIN R_MAX RHO_FLOOR SOURCE_MODEL SOURCE_FUNNEL_THETA_DEG SOURCE_DENSITY_Q
//    SOURCE_RADIAL_S SOURCE_GRADIENT_DR SOURCE_GRADIENT_DTHETA_DEG SOURCE_RHO_REF
//    SOURCE_CUTOFF_MIN SOURCE_CUTOFF_MAX]                                (argv 27-44)
// and, beyond the reference (which hard-codes the GBW table and is
// monochromatic only):
//   [SIGMA_PATH SPECTRAL_MODEL GAMMA ECUT E_MIN E_MAX N_BINS MODE PRECISION]  (45-53)
// MODE ref|cfg (default ref: the CPU integrator's semantics at GPU_TOL),
// PRECISION fp64|fp32k.  GPU_TILE and GPU_STEPS_PER_LAUNCH are accepted and
// ignored: the whole image is one launch of the fused kernel.
//
// Writes output/images/kerr_image_cuda_<tags>.bin (KIMG, :279-332) and .dat
// (:2517-2586) in the reference's j-major row order and prints Saved: lines
// and a timing line.  The mode flags forward to the two stage programs:
//   --geodesic-cache ...   -> the KGEO writer of compute_kerr_geodesics
//                             (argv of :1624-1660; writes output/rays/kerr_geodesics_cuda.bin)
//   --image-from-cache ... -> compute_kerr_image_from_cache on that cache
#include "cli_common.h"

#include <chrono>
#include <sys/wait.h>
#include <unistd.h>

namespace {

int run_child(const std::string& prog, const std::vector<std::string>& args)
{
    std::vector<char*> av;
    av.push_back(const_cast<char*>(prog.c_str()));
    for (const auto& s : args) av.push_back(const_cast<char*>(s.c_str()));
    av.push_back(nullptr);
    const pid_t pid = fork();
    if (pid == 0) {
        execv(prog.c_str(), av.data());
        std::perror("execv");
        _exit(127);
    }
    int status = 0;
    waitpid(pid, &status, 0);
    return WIFEXITED(status) ? WEXITSTATUS(status) : 1;
}

std::string sibling(const char* argv0, const char* name)
{
    std::string p(argv0);
    const size_t s = p.rfind('/');
    return (s == std::string::npos ? std::string("./") : p.substr(0, s + 1)) + name;
}

}  // namespace

int main(int argc, char** argv)
{
    using namespace cli;
    if (argc > 1 && std::string(argv[1]) == "--geodesic-cache") {
        // --geodesic-cache ASPIN CAM_R_OBS CAM_THETA_DEG CAM_FOV_DEG CAM_NX CAM_NY CAM_R_MAX CAM_STEP
        //                  GPU_CACHE_RAYS GPU_CACHE_MAX_POINTS GPU_TOL
        std::vector<std::string> a;
        const char* def[] = {"0.9", "60", "80", "25", "100", "100", "120", "0.001"};
        for (int i = 0; i < 8; ++i) a.push_back(argc > i + 2 ? argv[i + 2] : def[i]);
        a.push_back("output/rays/kerr_geodesics_cuda.bin");
        a.push_back("ref");
        a.push_back(argc > 12 ? argv[12] : "1e-8");
        a.push_back("200000");
        return run_child(sibling(argv[0], "compute_kerr_geodesics"), a);
    }
    if (argc > 1 && std::string(argv[1]) == "--image-from-cache") {
        // --image-from-cache ENU ASPIN MBH RHO0 R0 SIGMA HOR SRC_R SRC_SIG SRC_TH SRC_PL SRC_EMAX SRC_NORM
        //                    MEV_ENU MEV_NORM CAM_THETA GPU_CACHE_RAYS [PROFILE ... SOURCE_GRADIENT_DTHETA_DEG]
        auto A = [&](int i, const char* d) { return std::string(argc > i ? argv[i] : d); };
        std::vector<std::string> a;
        const char* d16[] = {"1e9", "0.9", "3.0", "1e-2", "10.0", "5.0", "0.25", "3.5", "1.0", "15.0",
                             "2.0", "1e12", "1.0", "10.0", "1.0", "80.0"};
        for (int i = 0; i < 16; ++i) a.push_back(A(i + 2, d16[i]));
        a.push_back(A(37, "data/sigma/sigma_nuN_CC_GBW.dat"));
        const char* dprof[] = {"gaussian", "2.0", "0.0", "15.0", "0.0", "2.5", "4.0", "60.0", "1e-99"};
        for (int i = 0; i < 9; ++i) a.push_back(A(19 + i, dprof[i]));
        a.push_back("60.0");
        a.push_back("1");
        a.push_back("output/rays/kerr_geodesics_cuda.bin");
        const char* dsrc[] = {"inner_ring", "20.0", "1.0", "2.0", "0.1", "1.0", "-1.0", "0.0", "100.0"};
        for (int i = 0; i < 9; ++i) a.push_back(A(28 + i, dsrc[i]));
        return run_child(sibling(argv[0], "compute_kerr_image_from_cache"), a);
    }

    hadros_scene sc;
    hadros_scene_defaults(&sc);
    hadros_medium& m = sc.medium;
    hadros_source& so = sc.source;
    hadros_spectrum& sp = sc.spectrum;
    // defaults of apps/compute_kerr_image_cuda.cu:2254-2296
    sc.enu_obs_GeV = argf(argc, argv, 1, 1.0e9);
    const double a_spin = argf(argc, argv, 2, 0.9);
    sc.m_bh_msun = argf(argc, argv, 3, 3.0);
    m.rho0 = argf(argc, argv, 4, 1.0e-2);
    m.r0 = argf(argc, argv, 5, 10.0);
    m.sigma_r = argf(argc, argv, 6, 5.0);
    m.h_over_r = argf(argc, argv, 7, 0.25);
    so.r_center = argf(argc, argv, 8, 3.5);
    so.sigma_r = argf(argc, argv, 9, 1.0);
    so.theta_width_rad = argf(argc, argv, 10, 15.0) * M_PI / 180.0;
    so.powerlaw = argf(argc, argv, 11, 2.0);
    so.emax_GeV = argf(argc, argv, 12, 1.0e12);
    so.norm = argf(argc, argv, 13, 1.0);
    const double mev_enu = argf(argc, argv, 14, 10.0);
    const double mev_norm = argf(argc, argv, 15, 1.0);
    hadros_camera cam;
    hadros_camera_defaults(&cam);
    cam.a_spin = a_spin;
    cam.theta_obs_rad = argf(argc, argv, 16, 80.0) * M_PI / 180.0;
    cam.r_obs = argf(argc, argv, 17, 60.0);
    cam.fov_deg = argc > 18 ? std::atof(argv[18]) : 25.0;
    cam.nx = argi(argc, argv, 19, 100);
    cam.ny = argi(argc, argv, 20, 100);
    cam.r_max = argf(argc, argv, 21, 120.0);
    cam.h = argf(argc, argv, 22, 0.001);
    const int gpu_tile = argi(argc, argv, 23, 32);
    const int steps_per_launch = argi(argc, argv, 24, 128);
    cam.max_steps = argi(argc, argv, 25, 200000);
    cam.tol = argf(argc, argv, 26, 1.0e-8);
    std::string profile = args(argc, argv, 27, "gaussian");
    m.radial_power = argf(argc, argv, 28, 2.0);
    m.funnel_depletion = argf(argc, argv, 29, 0.0);
    const double funnel_theta_deg = argf(argc, argv, 30, 15.0);
    m.funnel_theta_rad = funnel_theta_deg * M_PI / 180.0;
    m.envelope_rho0 = argf(argc, argv, 31, 0.0);
    m.envelope_alpha = argf(argc, argv, 32, 2.5);
    m.r_min = argf(argc, argv, 33, 4.0);
    m.r_max = argf(argc, argv, 34, 60.0);
    m.rho_floor = argf(argc, argv, 35, 1.0e-99);
    std::string source_name = args(argc, argv, 36, "inner_ring");
    so.funnel_theta_rad = argf(argc, argv, 37, 20.0) * M_PI / 180.0;
    so.density_q = argf(argc, argv, 38, 1.0);
    so.radial_s = argf(argc, argv, 39, 2.0);
    so.gradient_dr = argf(argc, argv, 40, 0.1);
    so.gradient_dtheta_rad = argf(argc, argv, 41, 1.0) * M_PI / 180.0;
    so.rho_ref = argf(argc, argv, 42, -1.0);
    so.cutoff_min = argf(argc, argv, 43, 0.0);
    so.cutoff_max = argf(argc, argv, 44, 1.0e2);
    const std::string sigma_path = args(argc, argv, 45, "data/sigma/sigma_nuN_CC_GBW.dat");
    std::string spectral_name = args(argc, argv, 46, "monochromatic");
    sp.gamma = argf(argc, argv, 47, 2.0);
    sp.ecut_GeV = argf(argc, argv, 48, 1.0e12);
    sp.e_min_GeV = argf(argc, argv, 49, 1.0e5);
    sp.e_max_GeV = argf(argc, argv, 50, 1.0e12);
    sp.n_bins = argi(argc, argv, 51, 8);
    const std::string mode = args(argc, argv, 52, "ref");
    const std::string prec = args(argc, argv, 53, "fp64");
    if (cam.nx <= 0 || cam.ny <= 0 || gpu_tile <= 0 || steps_per_launch <= 0 || cam.max_steps <= 0) {
        std::cerr << "Usage: " << argv[0]
                  << " ENU ASPIN MBH_MSUN TORUS_RHO0 TORUS_R0 TORUS_SIGMA TORUS_H_OVER_R SOURCE_R SOURCE_SIGMA"
                     " SOURCE_THETA_DEG SOURCE_POWERLAW SOURCE_EMAX SOURCE_NORM MEV_ENU MEV_NORM CAM_THETA_DEG"
                     " CAM_R_OBS CAM_FOV CAM_NX CAM_NY CAM_R_MAX CAM_STEP GPU_TILE GPU_STEPS_PER_LAUNCH"
                     " GPU_MAX_STEPS GPU_TOL [PROFILE ... SOURCE_CUTOFF_MAX] [SIGMA_PATH SPECTRAL_MODEL GAMMA"
                     " ECUT E_MIN E_MAX N_BINS MODE PRECISION]\n";
        return 1;
    }
    if (mode != "ref" && mode != "cfg") {
        std::cerr << "MODE must be ref or cfg\n";
        return 1;
    }
    if (prec != "fp64" && prec != "fp32k") {
        std::cerr << "PRECISION must be fp64 or fp32k\n";
        return 1;
    }
    cam.mode = mode == "cfg" ? HADROS_MODE_CFG : HADROS_MODE_REF;
    if (hadros_parse_profile(profile.c_str(), &m.kind) || hadros_parse_source(source_name.c_str(), &so.kind) ||
        hadros_parse_spectrum(spectral_name.c_str(), &sp.kind)) {
        std::cerr << hadros_last_error() << "\n";
        return 1;
    }
    const Sigma sig = load_sigma(sigma_path);
    char dev[256];
    int cus = 0, arch = 0;
    check(hadros_device_info(dev, sizeof(dev), &cus, &arch), "GPU");
    std::cout << "GPU device: " << dev << ", " << cus << " CUs\n";
    std::cout << "Grid: " << cam.nx << " x " << cam.ny << "\n";

    const auto t0 = std::chrono::steady_clock::now();
    const size_t npix = size_t(cam.nx) * size_t(cam.ny);
    std::vector<double> tau(npix), p(npix), I(npix), al(npix), be(npix);
    std::vector<int32_t> capd(npix), np(npix), fl(npix);
    hadros_image_out o;
    std::memset(&o, 0, sizeof(o));
    o.tau = tau.data();
    o.p_surv = p.data();
    o.intensity = I.data();
    o.captured = capd.data();
    o.npoints = np.data();
    o.flags = fl.data();
    o.alpha = al.data();
    o.beta = be.data();
    hadros_stats st;
    check(hadros_trace_rt(&cam, &sc, sig.e.data(), sig.s.data(), int(sig.e.size()),
                          prec == "fp32k" ? HADROS_PREC_FP32K : HADROS_PREC_FP64, &o, &st),
          "trace");
    // rows in j-major order (image[j*nx + i], apps/compute_kerr_image_cuda.cu:1620)
    std::vector<Row> rows;
    rows.reserve(npix);
    for (int j = 0; j < cam.ny; ++j)
        for (int i = 0; i < cam.nx; ++i) {
            const size_t k = size_t(i) * cam.ny + j;
            rows.push_back(Row{i, j, capd[k], al[k], be[k], tau[k], p[k], I[k]});
        }
    const double cam_theta_deg = cam.theta_obs_rad * 180.0 / M_PI;
    const std::string stem = std::string("output/images/kerr_image_cuda_rho0_torus_") +
                             make_compact_scientific_tag(m.rho0) +
                             (profile == "gaussian" ? std::string("") : "_Profile_" + profile) +
                             (so.kind == HADROS_SOURCE_INNER_RING ? std::string("") : "_Source_" + source_name) +
                             "_Enu_" + make_tag("%.0e", sc.enu_obs_GeV) + "_MeVEnu_" + make_tag("%.0e", mev_enu) +
                             "_MeVNorm_" + make_tag("%.0e", mev_norm) + "_CamTheta_" + make_tag("%.1f", cam_theta_deg);
    const std::string bin_name = stem + ".bin", dat_name = stem + ".dat";
    if (!write_kimg(bin_name, rows, cam.nx, cam.ny, sc.enu_obs_GeV, mev_enu, mev_norm, cam_theta_deg, 3)) {
        std::cerr << "Could not open binary output file: " << bin_name << "\n";
        return 1;
    }
    std::ofstream out(dat_name);
    if (!out) {
        std::cerr << "Could not open output file: " << dat_name << "\n";
        return 1;
    }
    out << "# profile_type " << profile << "\n# rho0 " << m.rho0 << "\n# rho0_gcm3 " << m.rho0 << "\n"
        << "# r0 " << m.r0 << "\n# r0_rg " << m.r0 << "\n# sigma_r " << m.sigma_r << "\n# sigma_r_rg " << m.sigma_r
        << "\n# H_over_R " << m.h_over_r << "\n# radial_power " << m.radial_power << "\n# funnel_depletion "
        << m.funnel_depletion << "\n# funnel_theta_deg " << funnel_theta_deg << "\n# rho_floor " << m.rho_floor
        << "\n# rho_floor_gcm3 " << m.rho_floor << "\n# envelope_rho0 " << m.envelope_rho0
        << "\n# envelope_rho0_gcm3 " << m.envelope_rho0 << "\n# envelope_alpha " << m.envelope_alpha
        << "\n# R_min " << m.r_min << "\n# R_min_rg " << m.r_min << "\n# R_max " << m.r_max << "\n# R_max_rg "
        << m.r_max << "\n# spin " << a_spin << "\n# observer_distance " << cam.r_obs << "\n# observer_distance_rg "
        << cam.r_obs << "\n# observer_inclination " << cam_theta_deg << "\n# observer_inclination_deg "
        << cam_theta_deg << "\n# DIS_model " << sigma_model_tag(sigma_path) << "\n# use_F3 1\n# source_model "
        << source_name << "\n# source_r_rg " << so.r_center << "\n# source_sigma_rg " << so.sigma_r
        << "\n# source_theta_deg " << so.theta_width_rad * 180.0 / M_PI << "\n# source_powerlaw " << so.powerlaw
        << "\n# source_emax_GeV " << so.emax_GeV << "\n# source_norm " << so.norm << "\n# source_funnel_theta_deg "
        << so.funnel_theta_rad * 180.0 / M_PI << "\n# source_rho_ref " << so.rho_ref << "\n# rho_ref " << so.rho_ref
        << "\n# source_q " << so.density_q << "\n# source_s " << so.radial_s << "\n# source_density_power_q "
        << so.density_q << "\n# source_radial_power_s " << so.radial_s << "\n# source_cutoff_min " << so.cutoff_min
        << "\n# source_cutoff_max " << so.cutoff_max << "\n# source_gradient_dr_rg " << so.gradient_dr
        << "\n# source_gradient_dtheta_deg " << so.gradient_dtheta_rad * 180.0 / M_PI << "\n# spectral_model "
        << spectral_name << "\n# integrator_mode " << mode << "\n# precision " << prec << "\n";
    out << row_columns();
    for (const Row& r : rows) write_row(out, r);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "Saved: " << bin_name << "\n";
    std::cout << "Saved: " << dat_name << "\n";
    std::cout << "GPU timing: " << s << " s, " << double(npix) / std::max(s, 1e-300)
              << " rays/s, kernel " << st.kernel_ms << " ms, accepted steps=" << st.total_points
              << ", attempts=" << st.total_attempts << ", exhausted=" << st.n_exhausted << "\n";
    return 0;
}
