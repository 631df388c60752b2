// cli_common.h -- file formats and argv helpers shared by the drop-in CLIs.
//
// Formats kept byte-compatible with the reference:
//   KGEO geodesic cache  apps/compute_kerr_geodesics.cpp:14-22,95-140
//   image .dat           apps/compute_kerr_image_from_cache.cpp:408-654
//   KIMG image .bin      apps/compute_kerr_image_cuda.cu:279-332
#pragma once

#include "../../include/hadros.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

namespace cli {

constexpr int32_t KGEO_MAGIC = 0x4B47454F;   // "KGEO"
constexpr int32_t KIMG_MAGIC = 0x4B494D47;   // "KIMG"

// BinaryRayHeader: 5 x int32, 4 bytes of padding, 2 x f64 = 40 bytes.
struct RayHeader {
    int32_t ray_id;
    int32_t pixel_i;
    int32_t pixel_j;
    int32_t captured;
    int32_t npoints;
    double alpha_rg;
    double beta_rg;
};
static_assert(sizeof(RayHeader) == 40, "KGEO ray header must be 40 bytes");

inline void check(int rc, const char* what)
{
    if (rc != HADROS_OK) {
        std::cerr << what << ": " << hadros_last_error() << "\n";
        std::exit(1);
    }
}

inline double argf(int argc, char** argv, int i, double def) { return argc > i ? std::atof(argv[i]) : def; }
inline int argi(int argc, char** argv, int i, int def) { return argc > i ? std::atoi(argv[i]) : def; }
inline std::string args(int argc, char** argv, int i, const std::string& def) { return argc > i ? argv[i] : def; }

// make_tag of apps/compute_kerr_image_from_cache.cpp:408-425
inline std::string make_tag(const char* fmt, double value)
{
    char buf[64];
    std::snprintf(buf, sizeof(buf), fmt, value);
    std::string t(buf);
    std::replace(t.begin(), t.end(), '+', '_');
    std::replace(t.begin(), t.end(), '-', 'm');
    std::replace(t.begin(), t.end(), '.', 'p');
    return t;
}

// make_compact_scientific_tag (:427-453): "%.0e" with the exponent as an integer
inline std::string make_compact_scientific_tag(double value)
{
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.0e", value);
    std::string t(buf);
    const size_t epos = t.find('e');
    if (epos == std::string::npos) return t;
    return t.substr(0, epos) + "e" + std::to_string(std::stoi(t.substr(epos + 1)));
}

// DIS model tag from the table path (:307-322)
inline std::string sigma_model_tag(const std::string& path)
{
    if (path.find("CTW_reference") != std::string::npos) return "CTW_reference";
    if (path.find("CSMS_reference") != std::string::npos) return "CSMS_reference";
    if (path.find("literature_powerlaw_scale") != std::string::npos ||
        path.find("gandhi_like_fit") != std::string::npos)
        return "literature_powerlaw_scale";
    if (path.find("IIM") != std::string::npos) return "IIM";
    return "GBW";
}

struct Sigma {
    std::vector<double> e, g, s;
};

inline Sigma load_sigma(const std::string& path)
{
    Sigma t;
    int n = 0;
    check(hadros_sigma_load(path.c_str(), nullptr, nullptr, nullptr, 0, &n), "sigma table");
    t.e.resize(n);
    t.g.resize(n);
    t.s.resize(n);
    check(hadros_sigma_load(path.c_str(), t.e.data(), t.g.data(), t.s.data(), n, &n), "sigma table");
    return t;
}

// ---- KGEO writer ---------------------------------------------------------
struct KgeoWriter {
    std::ofstream out;
    bool open(const std::string& path, int nx, int ny, double a)
    {
        out.open(path, std::ios::binary);
        if (!out) return false;
        const int32_t version = 1, nx32 = nx, ny32 = ny;
        out.write(reinterpret_cast<const char*>(&KGEO_MAGIC), 4);
        out.write(reinterpret_cast<const char*>(&version), 4);
        out.write(reinterpret_cast<const char*>(&nx32), 4);
        out.write(reinterpret_cast<const char*>(&ny32), 4);
        out.write(reinterpret_cast<const char*>(&a), 8);
        return bool(out);
    }
    void ray(const RayHeader& h, const double* pts)
    {
        out.write(reinterpret_cast<const char*>(&h), sizeof(h));
        if (h.npoints > 0) out.write(reinterpret_cast<const char*>(pts), std::streamsize(h.npoints) * 56);
    }
};

// ---- KGEO reader (streaming) --------------------------------------------
struct KgeoReader {
    std::ifstream in;
    int32_t magic = 0, version = 0, nx = 0, ny = 0;
    double a = 0.0;
    bool open(const std::string& path)
    {
        in.open(path, std::ios::binary);
        if (!in) return false;
        in.read(reinterpret_cast<char*>(&magic), 4);
        in.read(reinterpret_cast<char*>(&version), 4);
        in.read(reinterpret_cast<char*>(&nx), 4);
        in.read(reinterpret_cast<char*>(&ny), 4);
        in.read(reinterpret_cast<char*>(&a), 8);
        return true;
    }
    // 0 = ray read, 1 = clean end of file, -1 = negative npoints, -2 = truncated points
    int next(RayHeader& h, std::vector<double>& pts)
    {
        in.read(reinterpret_cast<char*>(&h), sizeof(h));
        if (!in) return 1;
        if (h.npoints < 0) return -1;
        pts.resize(size_t(h.npoints) * 7);
        if (h.npoints > 0) {
            in.read(reinterpret_cast<char*>(pts.data()), std::streamsize(h.npoints) * 56);
            if (!in) return -2;
        }
        return 0;
    }
};

// ---- KIMG writer (apps/compute_kerr_image_cuda.cu:279-332) -----------------
// MeV columns tau_mev P_surv_mev I_obs_mev r_neutrinosphere_rg leakage_mev; a row
// without them carries the RTResult defaults (include/radiative_transfer.hpp:11-21).
struct Row {
    int i, j, captured;
    double alpha, beta, tau, p, I;
    double mev[5] = {0.0, 1.0, 0.0, -1.0, 1.0};
};

inline bool write_kimg(const std::string& path, const std::vector<Row>& rows, int nx, int ny, double enu,
                       double mev_enu, double mev_norm, double cam_theta_deg, int32_t backend)
{
    std::ofstream out(path, std::ios::binary);
    if (!out) return false;
    const int32_t hdr[7] = {KIMG_MAGIC, 1, backend, nx, ny, 13, 0};
    out.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
    const double f[4] = {enu, mev_enu, mev_norm, cam_theta_deg};
    out.write(reinterpret_cast<const char*>(f), sizeof(f));
    for (const Row& r : rows) {
        const double v[13] = {double(r.i), double(r.j), r.alpha, r.beta, r.tau, r.p, r.I, double(r.captured),
                              r.mev[0], r.mev[1], r.mev[2], r.mev[3], r.mev[4]};
        out.write(reinterpret_cast<const char*>(v), sizeof(v));
    }
    return bool(out);
}

inline void write_row(std::ostream& out, const Row& r)
{
    out << std::scientific << std::setprecision(8) << r.i << " " << r.j << " " << r.alpha << " " << r.beta << " "
        << r.tau << " " << r.p << " " << r.I << " " << r.captured << " " << r.mev[0] << " " << r.mev[1] << " "
        << r.mev[2] << " " << r.mev[3] << " " << r.mev[4] << "\n";
}

inline const char* row_columns()
{
    return "# i j alpha beta tau_uhe P_surv_uhe I_obs_uhe captured "
           "tau_mev P_surv_mev I_obs_mev r_neutrinosphere_rg leakage_mev\n";
}

}  // namespace cli
