// compute_kerr_geodesics -- stage 1 of the two-stage pipeline on the GPU.
//
// Drop-in for apps/compute_kerr_geodesics.cpp: same positional argv
//   ASPIN R_OBS THETA_DEG FOV_DEG NX NY R_MAX H OUT_PATH
// (defaults :26-34), same KGEO output (header :95-104, 40-byte ray headers
// and 56-byte PathPoints :106-140), same "Camera grid:" / "Saved:" lines.
// Optional trailing arguments select this library's integrator knobs:
//   MODE (ref|cfg, default ref)  TOL (default 1e-6 ref, 1e-8 cfg)
//   MAX_STEPS (default 200000 ref, 20000 cfg)
// Rays are written in ray_id order (the reference's OpenMP order is arbitrary).
// Pixels are traced on the GPU in batches whose point buffers stay below
// HADROS_CACHE_BATCH_MB (default 2048) of host memory.
#include "cli_common.h"

#include <chrono>

int main(int argc, char** argv)
{
    using namespace cli;
    hadros_camera cam;
    hadros_camera_defaults(&cam);
    const double a = argf(argc, argv, 1, 0.9);
    const double r_obs = argf(argc, argv, 2, 60.0);
    const double theta_deg = argf(argc, argv, 3, 74.0);
    const double fov = argf(argc, argv, 4, 45.0);
    const int nx = argi(argc, argv, 5, 100);
    const int ny = argi(argc, argv, 6, 100);
    const double r_max = argf(argc, argv, 7, 120.0);
    const double h = argf(argc, argv, 8, 0.001);
    const std::string out_path = args(argc, argv, 9, "output/rays/kerr_geodesics.bin");
    const std::string mode = args(argc, argv, 10, "ref");
    cam.a_spin = a;
    cam.r_obs = r_obs;
    cam.theta_obs_rad = M_PI * theta_deg / 180.0;
    cam.fov_deg = fov;
    cam.nx = nx;
    cam.ny = ny;
    cam.r_max = r_max;
    cam.h = h;
    cam.mode = (mode == "cfg") ? HADROS_MODE_CFG : HADROS_MODE_REF;
    if (mode != "cfg" && mode != "ref") {
        std::cerr << "MODE must be ref or cfg\n";
        return 1;
    }
    cam.tol = argf(argc, argv, 11, cam.mode == HADROS_MODE_CFG ? 1.0e-8 : 1.0e-6);
    cam.max_steps = argi(argc, argv, 12, cam.mode == HADROS_MODE_CFG ? 20000 : 200000);

    char dev[256];
    int cus = 0, arch = 0;
    check(hadros_device_info(dev, sizeof(dev), &cus, &arch), "GPU");
    std::cout << "GPU device = " << dev << "\n";

    KgeoWriter w;
    if (!w.open(out_path, nx, ny, a)) {
        std::cerr << "Could not open " << out_path << "\n";
        return 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const long npix = long(nx) * long(ny);
    const char* env = std::getenv("HADROS_CACHE_BATCH_MB");
    const double budget = (env ? std::atof(env) : 2048.0) * 1024.0 * 1024.0;
    long k0 = 0, total_points = 0;
    long batch = std::min<long>(npix, 4096);
    std::vector<int64_t> off;
    std::vector<int32_t> cap;
    std::vector<double> pts;
    while (k0 < npix) {
        const long k1 = std::min(npix, k0 + batch);
        off.assign(size_t(k1 - k0 + 1), 0);
        cap.assign(size_t(k1 - k0), 0);
        check(hadros_trace_cache_range(&cam, k0, k1, off.data(), cap.data(), nullptr, 0), "trace");
        const int64_t n = off.back();
        if (double(n) * 56.0 > budget && k1 - k0 > 1) {   // shrink the batch and retry
            batch = std::max<long>(1, batch / 4);
            continue;
        }
        pts.resize(size_t(std::max<int64_t>(n, 1)) * 7);
        check(hadros_trace_cache_range(&cam, k0, k1, off.data(), cap.data(), pts.data(), n), "trace");
        for (long k = k0; k < k1; ++k) {
            const long q = k - k0;
            RayHeader rh;
            std::memset(&rh, 0, sizeof(rh));
            rh.ray_id = int32_t(k);
            rh.pixel_i = int32_t(k / ny);
            rh.pixel_j = int32_t(k % ny);
            rh.captured = cap[q];
            rh.npoints = int32_t(off[q + 1] - off[q]);
            const double th = std::tan(0.5 * (fov * M_PI / 180.0));
            rh.alpha_rg = (2.0 * (rh.pixel_i + 0.5) / nx - 1.0) * th;
            rh.beta_rg = (2.0 * (rh.pixel_j + 0.5) / ny - 1.0) * th;
            w.ray(rh, pts.data() + 7 * off[q]);
        }
        total_points += n;
        k0 = k1;
        if (double(n) * 56.0 * 4.0 < budget && batch < (1L << 22)) batch *= 2;
    }
    if (!w.out) {
        std::cerr << "Could not write " << out_path << "\n";
        return 1;
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "Camera grid: " << nx << " x " << ny << "\n";
    std::cout << "GPU geodesic cache: " << s << " s, " << double(npix) / std::max(s, 1e-300) << " rays/s, "
              << total_points << " points\n";
    std::cout << "Saved: " << out_path << "\n";
    return 0;
}
