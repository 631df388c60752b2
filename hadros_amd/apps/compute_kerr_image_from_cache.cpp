// compute_kerr_image_from_cache -- stage 2 (cache replay) on the GPU.
//
// Drop-in for apps/compute_kerr_image_from_cache.cpp: the same 71 positional
// arguments (:57-305), the same cache lookup (argv[29], else the GPU cache,
// else the CPU cache, :353-360), the same header checks and error exits
// (:365-406, :672-700), the same output file name (:408-500), metadata header
// (:512-597) and rows in cache order (:635-654), and the same stdout lines.
// The UHE columns and the five MeV columns (argv 14-15 and 45-71, the model of
// src/radiative_transfer.cpp:513-638) both come from the GPU replay kernel
// (hadros_replay_* with hadros_replay_set_mev).
#include "cli_common.h"

#include <chrono>
#include <sstream>
#include <utility>

namespace {

struct MevArgs {   // argv 14-15 and 45-71 (:200-305)
    double enu = 10.0, norm = 1.0;
    std::string model = "physical", flavor = "anti_nu_e", thermal = "inner_hot_torus",
                ye = "neutron_rich_torus", spectral = "monochromatic";
    int urca = 1, pair = 1, brems = 1, absorption = 1, scattering = 1;
    double T0 = 6.0, T_floor = 0.1, T_power = 0.2, Ye_torus = 0.25, Ye_funnel = 0.55, Ye_envelope = 0.45,
           Ye_floor = 0.01, Ye_ceil = 0.60, E_min = 3.0, E_max = 50.0;
    int n_bins = 8, degeneracy = 0, abs_n = 1, abs_p = 1, scat_n = 1, scat_p = 1, scat_e = 1;
};

}  // namespace

int main(int argc, char** argv)
{
    using namespace cli;
    hadros_scene sc;
    hadros_scene_defaults(&sc);
    MevArgs mev;
    hadros_medium& m = sc.medium;
    hadros_source& so = sc.source;
    hadros_spectrum& sp = sc.spectrum;
    // argv 1-29 (:57-145)
    sc.enu_obs_GeV = argf(argc, argv, 1, 1.0e9);
    const double a_spin = argf(argc, argv, 2, 0.9);
    sc.m_bh_msun = argf(argc, argv, 3, 3.0);
    m.rho0 = argf(argc, argv, 4, 1.0e10);
    m.r0 = argf(argc, argv, 5, 10.0);
    m.sigma_r = argf(argc, argv, 6, 5.0);
    m.h_over_r = argf(argc, argv, 7, 0.01);
    so.r_center = argf(argc, argv, 8, so.r_center);
    so.sigma_r = argf(argc, argv, 9, so.sigma_r);
    if (argc > 10) so.theta_width_rad = std::atof(argv[10]) * M_PI / 180.0;
    so.powerlaw = argf(argc, argv, 11, so.powerlaw);
    so.emax_GeV = argf(argc, argv, 12, so.emax_GeV);
    so.norm = argf(argc, argv, 13, so.norm);
    mev.enu = argf(argc, argv, 14, mev.enu);
    mev.norm = argf(argc, argv, 15, mev.norm);
    const double cam_theta_deg = argf(argc, argv, 16, 0.0);
    const std::string sigma_path = args(argc, argv, 17, "data/sigma/sigma_nuN_CC_GBW.dat");
    const std::string profile = args(argc, argv, 18, "gaussian");
    m.radial_power = argf(argc, argv, 19, 2.0);
    m.funnel_depletion = argf(argc, argv, 20, 0.0);
    const double funnel_theta_deg = argf(argc, argv, 21, 15.0);
    m.envelope_rho0 = argf(argc, argv, 22, 0.0);
    m.envelope_alpha = argf(argc, argv, 23, 2.5);
    m.r_min = argf(argc, argv, 24, 4.0);
    m.r_max = argf(argc, argv, 25, 60.0);
    m.rho_floor = argf(argc, argv, 26, 1.0e-99);
    const double observer_distance = argf(argc, argv, 27, 60.0);
    const int use_F3 = argi(argc, argv, 28, 1);
    const std::string requested_cache = args(argc, argv, 29, "");
    std::string source_name = "inner_ring";
    if (argc > 30) {
        source_name = argv[30];
        if (hadros_parse_source(argv[30], &so.kind)) {
            std::cerr << hadros_last_error() << "\n";
            return 1;
        }
    }
    if (argc > 31) so.funnel_theta_rad = std::atof(argv[31]) * M_PI / 180.0;
    so.density_q = argf(argc, argv, 32, so.density_q);
    so.radial_s = argf(argc, argv, 33, so.radial_s);
    so.gradient_dr = argf(argc, argv, 34, so.gradient_dr);
    if (argc > 35) so.gradient_dtheta_rad = std::atof(argv[35]) * M_PI / 180.0;
    so.rho_ref = argf(argc, argv, 36, so.rho_ref);
    so.cutoff_min = argf(argc, argv, 37, so.cutoff_min);
    so.cutoff_max = argf(argc, argv, 38, so.cutoff_max);
    std::string spectral_name = "monochromatic";
    if (argc > 39) {
        spectral_name = argv[39];
        if (hadros_parse_spectrum(argv[39], &sp.kind)) {
            std::cerr << hadros_last_error() << "\n";
            return 1;
        }
    }
    sp.gamma = argf(argc, argv, 40, sp.gamma);
    sp.ecut_GeV = argf(argc, argv, 41, sp.ecut_GeV);
    sp.e_min_GeV = argf(argc, argv, 42, sp.e_min_GeV);
    sp.e_max_GeV = argf(argc, argv, 43, sp.e_max_GeV);
    sp.n_bins = argi(argc, argv, 44, sp.n_bins);
    mev.model = args(argc, argv, 45, mev.model);
    mev.flavor = args(argc, argv, 46, mev.flavor);
    mev.urca = argi(argc, argv, 47, 1) != 0;
    mev.pair = argi(argc, argv, 48, 1) != 0;
    mev.brems = argi(argc, argv, 49, 1) != 0;
    mev.absorption = argi(argc, argv, 50, 1) != 0;
    mev.scattering = argi(argc, argv, 51, 1) != 0;
    mev.thermal = args(argc, argv, 52, mev.thermal);
    mev.ye = args(argc, argv, 53, mev.ye);
    mev.T0 = argf(argc, argv, 54, mev.T0);
    mev.T_floor = argf(argc, argv, 55, mev.T_floor);
    mev.T_power = argf(argc, argv, 56, mev.T_power);
    mev.Ye_torus = argf(argc, argv, 57, mev.Ye_torus);
    mev.Ye_funnel = argf(argc, argv, 58, mev.Ye_funnel);
    mev.Ye_envelope = argf(argc, argv, 59, mev.Ye_envelope);
    mev.Ye_floor = argf(argc, argv, 60, mev.Ye_floor);
    mev.Ye_ceil = argf(argc, argv, 61, mev.Ye_ceil);
    mev.spectral = args(argc, argv, 62, mev.spectral);
    mev.E_min = argf(argc, argv, 63, mev.E_min);
    mev.E_max = argf(argc, argv, 64, mev.E_max);
    mev.n_bins = argi(argc, argv, 65, mev.n_bins);
    mev.degeneracy = argi(argc, argv, 66, 0) != 0;
    mev.abs_n = argi(argc, argv, 67, 1) != 0;
    mev.abs_p = argi(argc, argv, 68, 1) != 0;
    mev.scat_n = argi(argc, argv, 69, 1) != 0;
    mev.scat_p = argi(argc, argv, 70, 1) != 0;
    mev.scat_e = argi(argc, argv, 71, 1) != 0;

    // MeVThermalParams of argv 14-15 and 45-71; the reference's parsers throw on
    // an unknown name, here the message is printed and the program exits 1
    hadros_mev hm;
    hadros_mev_defaults(&hm);
    hm.enu_obs_MeV = mev.enu;
    hm.norm = mev.norm;
    const std::pair<int (*)(const char*, int32_t*), std::pair<const std::string*, int32_t*>> names[] = {
        {hadros_parse_mev_model, {&mev.model, &hm.model}},
        {hadros_parse_mev_flavor, {&mev.flavor, &hm.flavor}},
        {hadros_parse_mev_thermal, {&mev.thermal, &hm.thermal_profile}},
        {hadros_parse_mev_ye, {&mev.ye, &hm.ye_profile}},
        {hadros_parse_mev_spectral, {&mev.spectral, &hm.spectral_mode}}};
    for (const auto& n : names)
        if (n.first(n.second.first->c_str(), n.second.second)) {
            std::cerr << hadros_last_error() << "\n";
            return 1;
        }
    const int bits[] = {mev.urca, mev.pair, mev.brems, mev.absorption, mev.scattering, mev.degeneracy,
                        mev.abs_n, mev.abs_p, mev.scat_n, mev.scat_p, mev.scat_e};
    hm.processes = 0;
    for (int b = 0; b < 11; ++b)
        if (bits[b]) hm.processes |= 1u << b;
    hm.T0_MeV = mev.T0;
    hm.T_floor_MeV = mev.T_floor;
    hm.T_power = mev.T_power;
    hm.Ye_torus = mev.Ye_torus;
    hm.Ye_funnel = mev.Ye_funnel;
    hm.Ye_envelope = mev.Ye_envelope;
    hm.Ye_floor = mev.Ye_floor;
    hm.Ye_ceil = mev.Ye_ceil;
    hm.E_min_MeV = mev.E_min;
    hm.E_max_MeV = mev.E_max;
    hm.n_bins = mev.n_bins;

    const std::string sigma_model = sigma_model_tag(sigma_path);

    const Sigma sig = load_sigma(sigma_path);
    if (hadros_parse_profile(profile.c_str(), &m.kind)) {
        std::cerr << hadros_last_error() << "\n";
        return 1;
    }
    m.funnel_theta_rad = funnel_theta_deg * M_PI / 180.0;

    const std::string gpu_cache = "output/rays/kerr_geodesics_cuda.bin";
    const std::string cpu_cache = "output/rays/kerr_geodesics.bin";
    const std::string cache_path =
        !requested_cache.empty() ? requested_cache
                                 : (std::ifstream(gpu_cache, std::ios::binary) ? gpu_cache : cpu_cache);
    KgeoReader rd;
    if (!rd.open(cache_path)) {
        std::cerr << "Could not open geodesic cache (tried " << gpu_cache << " and " << cpu_cache << ")\n";
        return 1;
    }
    std::cout << "Using geodesic cache: " << cache_path << "\n";
    std::cout << "Using sigma model: " << sigma_model << " (" << sigma_path << ")\n";
    if (rd.magic != KGEO_MAGIC) {
        std::cerr << "Invalid binary geodesic cache: bad magic number\n";
        return 1;
    }
    if (rd.version != 1) {
        std::cerr << "Unsupported binary geodesic cache version: " << rd.version << "\n";
        return 1;
    }
    if (std::abs(rd.a - a_spin) > 1.0e-12)
        std::cerr << "Warning: cache spin a=" << rd.a << " differs from requested ASPIN=" << a_spin << "\n";

    const std::string mev_band_tag =
        hm.spectral_mode == HADROS_MEV_SPECTRAL_MONOCHROMATIC   // the parsed mode: aliases such as "mono" too
            ? std::string("")
            : std::string("_MeVSpectrum_") + mev.spectral + "_E" + make_tag("%.0e", mev.E_min) + "_" +
                  make_tag("%.0e", mev.E_max);
    const std::string image_filename =
        std::string("output/images/kerr_image_cuda_cache_") + sigma_model + "_rho0_torus_" +
        make_compact_scientific_tag(m.rho0) + (profile == "gaussian" ? std::string("") : "_Profile_" + profile) +
        (so.kind == HADROS_SOURCE_INNER_RING ? std::string("") : "_Source_" + source_name) +
        (sp.kind == HADROS_SPECTRUM_MONOCHROMATIC ? std::string("") : "_Spectrum_" + spectral_name) + "_Enu_" +
        make_tag("%.0e", sc.enu_obs_GeV) + "_MeVEnu_" + make_tag("%.0e", mev.enu) + "_MeVNorm_" +
        make_tag("%.0e", mev.norm) + mev_band_tag + "_CamTheta_" + make_tag("%.1f", cam_theta_deg) + ".dat";
    std::ofstream out(image_filename);
    if (!out) {
        std::cerr << "Could not open " << image_filename << "\n";
        return 1;
    }
    // metadata header, same keys and order as :512-597
    out << "# profile_type " << profile << "\n"
        << "# rho0 " << m.rho0 << "\n# rho0_gcm3 " << m.rho0 << "\n"
        << "# r0 " << m.r0 << "\n# r0_rg " << m.r0 << "\n"
        << "# sigma_r " << m.sigma_r << "\n# sigma_r_rg " << m.sigma_r << "\n"
        << "# H_over_R " << m.h_over_r << "\n"
        << "# radial_power " << m.radial_power << "\n"
        << "# funnel_depletion " << m.funnel_depletion << "\n"
        << "# funnel_theta_deg " << funnel_theta_deg << "\n"
        << "# rho_floor " << m.rho_floor << "\n# rho_floor_gcm3 " << m.rho_floor << "\n"
        << "# envelope_rho0 " << m.envelope_rho0 << "\n# envelope_rho0_gcm3 " << m.envelope_rho0 << "\n"
        << "# envelope_alpha " << m.envelope_alpha << "\n"
        << "# R_min " << m.r_min << "\n# R_min_rg " << m.r_min << "\n"
        << "# R_max " << m.r_max << "\n# R_max_rg " << m.r_max << "\n"
        << "# spin " << a_spin << "\n"
        << "# observer_distance " << observer_distance << "\n# observer_distance_rg " << observer_distance << "\n"
        << "# observer_inclination " << cam_theta_deg << "\n# observer_inclination_deg " << cam_theta_deg << "\n"
        << "# DIS_model " << sigma_model << "\n"
        << "# use_F3 " << use_F3 << "\n"
        << "# source_model " << source_name << "\n"
        << "# source_r_rg " << so.r_center << "\n"
        << "# source_sigma_rg " << so.sigma_r << "\n"
        << "# source_theta_deg " << so.theta_width_rad * 180.0 / M_PI << "\n"
        << "# source_powerlaw " << so.powerlaw << "\n"
        << "# source_emax_GeV " << so.emax_GeV << "\n"
        << "# source_norm " << so.norm << "\n"
        << "# source_funnel_theta_deg " << so.funnel_theta_rad * 180.0 / M_PI << "\n"
        << "# source_rho_ref " << so.rho_ref << "\n# rho_ref " << so.rho_ref << "\n"
        << "# source_q " << so.density_q << "\n# source_s " << so.radial_s << "\n"
        << "# source_density_power_q " << so.density_q << "\n# source_radial_power_s " << so.radial_s << "\n"
        << "# source_cutoff_min " << so.cutoff_min << "\n# source_cutoff_max " << so.cutoff_max << "\n"
        << "# source_gradient_dr_rg " << so.gradient_dr << "\n"
        << "# source_gradient_dtheta_deg " << so.gradient_dtheta_rad * 180.0 / M_PI << "\n"
        << "# spectral_model " << spectral_name << "\n"
        << "# spectral_gamma " << sp.gamma << "\n"
        << "# spectral_ecut_GeV " << sp.ecut_GeV << "\n"
        << "# spectral_E_min_GeV " << sp.e_min_GeV << "\n"
        << "# spectral_E_max_GeV " << sp.e_max_GeV << "\n"
        << "# spectral_N_bins " << sp.n_bins << "\n"
        << "# MEV_MODEL " << mev.model << "\n# MEV_CPU_MODEL " << mev.model << "\n"
        << "# MEV_CUDA_STATUS computed_gpu_" << (hm.model == HADROS_MEV_MODEL_TOY ? "toy" : "physical") << "\n"
        << "# MEV_FLAVOR " << mev.flavor << "\n"
        << "# MEV_ENERGY_MEV " << mev.enu << "\n"
        << "# MEV_INCLUDE_URCA " << mev.urca << "\n# MEV_INCLUDE_PAIR " << mev.pair << "\n"
        << "# MEV_INCLUDE_BREMS " << mev.brems << "\n# MEV_INCLUDE_ABSORPTION " << mev.absorption << "\n"
        << "# MEV_INCLUDE_SCATTERING " << mev.scattering << "\n"
        << "# MEV_THERMAL_PROFILE " << mev.thermal << "\n# MEV_YE_PROFILE " << mev.ye << "\n"
        << "# MEV_T0_MEV " << mev.T0 << "\n# MEV_T_FLOOR_MEV " << mev.T_floor << "\n"
        << "# MEV_T_POWER " << mev.T_power << "\n"
        << "# MEV_YE_TORUS " << mev.Ye_torus << "\n# MEV_YE_FUNNEL " << mev.Ye_funnel << "\n"
        << "# MEV_YE_ENVELOPE " << mev.Ye_envelope << "\n# MEV_YE_FLOOR " << mev.Ye_floor << "\n"
        << "# MEV_YE_CEIL " << mev.Ye_ceil << "\n"
        << "# MEV_SPECTRAL_MODE " << mev.spectral << "\n"
        << "# MEV_E_MIN_MEV " << mev.E_min << "\n# MEV_E_MAX_MEV " << mev.E_max << "\n"
        << "# MEV_N_BINS " << mev.n_bins << "\n"
        << "# MEV_USE_DEGENERACY_CORRECTION " << mev.degeneracy << "\n"
        << "# MEV_INCLUDE_ABS_N " << mev.abs_n << "\n# MEV_INCLUDE_ABS_P " << mev.abs_p << "\n"
        << "# MEV_INCLUDE_SCAT_N " << mev.scat_n << "\n# MEV_INCLUDE_SCAT_P " << mev.scat_p << "\n"
        << "# MEV_INCLUDE_SCAT_E " << mev.scat_e << "\n"
        << "# mev_emissivity_units local_spectral_emissivity_proxy_per_MeV\n"
        << "# mev_opacity_units cm^-1\n";
    out << row_columns();

    // ---- stream the cache in chunks, replay each chunk on the GPU
    const char* env = std::getenv("HADROS_REPLAY_CHUNK_POINTS");
    const long chunk_points = env ? std::atol(env) : (64L << 20);
    std::vector<RayHeader> hdrs;
    std::vector<int64_t> off{0};
    std::vector<double> r, th, dl, z, pts;
    size_t loaded = 0, processed = 0;
    double gpu_ms = 0.0;
    bool device_shown = false;
    auto flush = [&]() {
        if (hdrs.empty()) return;
        if (!device_shown) {   // the GPU is first needed here; file errors above exit before it
            char dev[256];
            int cus = 0, arch = 0;
            check(hadros_device_info(dev, sizeof(dev), &cus, &arch), "GPU");
            std::cout << "GPU device = " << dev << "\n";
            device_shown = true;
        }
        hadros_cache_view v;
        std::memset(&v, 0, sizeof(v));
        v.n_rays = int32_t(hdrs.size());
        v.a_bh = rd.a;
        v.offset = off.data();
        v.r = r.data();
        v.theta = th.data();
        v.dl = dl.data();
        v.redshift = z.data();
        const size_t n = hdrs.size();
        std::vector<double> tau(n), p(n), I(n), mcol[5];
        for (auto& c : mcol) c.resize(n);
        hadros_image_out o;
        std::memset(&o, 0, sizeof(o));
        o.tau = tau.data();
        o.p_surv = p.data();
        o.intensity = I.data();
        hadros_mev_out mo{mcol[0].data(), mcol[1].data(), mcol[2].data(), mcol[3].data(), mcol[4].data()};
        hadros_stats st;
        hadros_replay* plan = nullptr;
        check(hadros_replay_create(&v, &sc, sig.e.data(), sig.s.data(), int(sig.e.size()), HADROS_PREC_FP64, &plan),
              "replay");
        check(hadros_replay_set_mev(plan, &hm), "MeV channel");
        check(hadros_replay_run(plan, nullptr), "replay");
        check(hadros_replay_fetch(plan, &o, &st), "replay");
        check(hadros_replay_fetch_mev(plan, &mo), "MeV channel");
        hadros_replay_destroy(plan);
        gpu_ms += st.kernel_ms;
        for (size_t k = 0; k < n; ++k) {
            Row row{hdrs[k].pixel_i, hdrs[k].pixel_j, hdrs[k].captured != 0 ? 1 : 0, hdrs[k].alpha_rg,
                    hdrs[k].beta_rg, tau[k], p[k], I[k]};
            for (int c = 0; c < 5; ++c) row.mev[c] = mcol[c][k];
            write_row(out, row);
        }
        processed += hdrs.size();
        hdrs.clear();
        off.assign(1, 0);
        r.clear();
        th.clear();
        dl.clear();
        z.clear();
    };
    while (true) {
        RayHeader h;
        const int st = rd.next(h, pts);
        if (st == 1) break;
        if (st == -1) {
            std::cerr << "Invalid ray with negative number of points\n";
            return 1;
        }
        if (st == -2) {
            std::cerr << "Unexpected end of file while reading points\n";
            return 1;
        }
        hdrs.push_back(h);
        for (int q = 0; q < h.npoints; ++q) {
            const double* pp = pts.data() + 7 * q;
            r.push_back(pp[0]);
            th.push_back(pp[1]);
            dl.push_back(pp[5]);
            z.push_back(pp[6]);
        }
        off.push_back(int64_t(r.size()));
        ++loaded;
        if (long(r.size()) >= chunk_points || hdrs.size() >= (1u << 20)) flush();
    }
    flush();
    std::cout << "Cache grid: " << rd.nx << " x " << rd.ny << "\n";
    std::cout << "Loaded rays: " << loaded << "\n";
    std::cout << "Processed rays: " << processed << "\n";
    std::cout << "GPU replay kernels: " << gpu_ms << " ms\n";
    std::cout << "Saved: " << image_filename << "\n";
    return 0;
}
