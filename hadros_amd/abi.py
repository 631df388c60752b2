"""ctypes mirror of include/hadros.h (structs, enums, status codes).

Plain data only: this module does not load the library (see hadros_amd.ffi).
The structs replace the reference's C++ parameter objects:
KerrCamera ctor args (include/kerr_camera.hpp:10-19), TorusProfile ctor args
(include/torus_profile.hpp:19-33), UHESourceParams / UHESpectralParams
(include/radiative_transfer.hpp:29-64).
"""
import ctypes as C
import math

ABI_VERSION = 1

OK = 0
ERR_ARG = 1
ERR_TABLE = 2
ERR_RANGE = 3
ERR_KIND = 4
ERR_CACHE_MAGIC = 5
ERR_CACHE_VERSION = 6
ERR_CACHE_TRUNCATED = 7
ERR_DEVICE = 8
ERR_IO = 9
ERR_MAX_POINTS = 10

MODE_REF = 0
MODE_CFG = 1
PREC_FP64 = 0
PREC_FP32K = 1
PATH_AUTO, PATH_FUSED, PATH_SPLIT = 0, 1, 2
TRACER_MAX_SEGMENTS = 4                  # hadros_tracer_set_segments

MAX_BINS = 16
EVAL_TRANSFER_COLS = 6 + 2 * MAX_BINS   # hadros_eval_transfer_point columns
STREAM_GRID = 2048                       # hadros_stream_probe partial sums
MAX_SIGMA = 1024
NOALLOC_SCRATCH_BYTES = 32768
FLAG_EXHAUSTED = 1

# names accepted by the reference parsers (src/torus_profile.cpp:44-91,
# src/radiative_transfer.cpp:80-170); value = enum of include/hadros.h
PROFILE_KINDS = {
    "gaussian": 0, "torus": 0, "gaussian_torus": 0,
    "powerlaw": 1, "powerlaw_disk": 1,
    "funnel": 2, "gaussian_funnel": 2,
    "powerlaw_funnel": 3,
    "gaussian_envelope": 4,
    "powerlaw_envelope": 5,
    "powerlaw_funnel_envelope": 6,
    "collapsar_ndaf_like": 7, "collapsar_ndaf": 7,
}
PROFILE_NAMES = ["gaussian", "powerlaw", "gaussian_funnel", "powerlaw_funnel",
                 "gaussian_envelope", "powerlaw_envelope", "powerlaw_funnel_envelope",
                 "collapsar_ndaf_like"]
SOURCE_KINDS = {
    "inner_ring": 0, "INNER_RING": 0, "ring": 0,
    "funnel_wall": 1, "FUNNEL_WALL": 1,
    "jet_base": 2, "JET_BASE": 2,
    "shock_layer": 3, "SHOCK_LAYER": 3,
    "density_weighted": 4, "DENSITY_WEIGHTED": 4,
}
SOURCE_NAMES = ["inner_ring", "funnel_wall", "jet_base", "shock_layer", "density_weighted"]
SPECTRUM_KINDS = {
    "monochromatic": 0, "MONOCHROMATIC": 0, "mono": 0,
    "powerlaw": 1, "POWERLAW": 1, "power_law": 1,
    "powerlaw_cutoff": 2, "POWERLAW_CUTOFF": 2, "power_law_cutoff": 2,
}
SPECTRUM_NAMES = ["monochromatic", "powerlaw", "powerlaw_cutoff"]

DEG = math.pi / 180.0

# MeV channel (include/hadros.h hadros_mev; src/mev_neutrino_physics.cpp:36-233 names)
MEV_MODELS = {"toy": 0, "TOY": 0, "physical": 1, "PHYSICAL": 1, "phys": 1}
MEV_FLAVORS = {"nu_e": 0, "nue": 0, "NU_E": 0, "anti_nu_e": 1, "antinue": 1, "anti-nu-e": 1, "ANTI_NU_E": 1,
               "nu_x": 2, "nux": 2, "NU_X": 2}
MEV_THERMAL = {"constant": 0, "CONSTANT": 0, "inner_hot_torus": 1, "INNER_HOT_TORUS": 1, "radial_powerlaw": 2,
               "RADIAL_POWERLAW": 2, "torus_plus_cool_envelope": 3, "TORUS_PLUS_COOL_ENVELOPE": 3,
               "collapsar_inner_hot": 4, "COLLAPSAR_INNER_HOT": 4}
MEV_YE = {"constant": 0, "CONSTANT": 0, "neutron_rich_torus": 1, "NEUTRON_RICH_TORUS": 1, "funnel_proton_rich": 2,
          "FUNNEL_PROTON_RICH": 2, "torus_envelope_contrast": 3, "TORUS_ENVELOPE_CONTRAST": 3,
          "collapsar_neutron_rich": 4, "COLLAPSAR_NEUTRON_RICH": 4}
MEV_SPECTRAL = {"monochromatic": 0, "MONOCHROMATIC": 0, "mono": 0, "fermi_dirac_band": 1, "FERMI_DIRAC_BAND": 1,
                "fd_band": 1}
# process bits, in the order of argv 47-51 and 66-71 of apps/compute_kerr_image_from_cache.cpp
MEV_URCA, MEV_PAIR, MEV_BREMS, MEV_ABSORPTION, MEV_SCATTERING, MEV_DEGENERACY = 1, 2, 4, 8, 16, 32
MEV_ABS_N, MEV_ABS_P, MEV_SCAT_N, MEV_SCAT_P, MEV_SCAT_E = 64, 128, 256, 512, 1024
MEV_DEFAULT_PROCESSES = (MEV_URCA | MEV_PAIR | MEV_BREMS | MEV_ABSORPTION | MEV_SCATTERING | MEV_ABS_N | MEV_ABS_P
                         | MEV_SCAT_N | MEV_SCAT_P | MEV_SCAT_E)


class Camera(C.Structure):
    _fields_ = [("a_spin", C.c_double), ("r_obs", C.c_double), ("theta_obs_rad", C.c_double),
                ("fov_deg", C.c_double), ("nx", C.c_int32), ("ny", C.c_int32),
                ("r_max", C.c_double), ("h", C.c_double), ("tol", C.c_double),
                ("mode", C.c_int32), ("max_steps", C.c_int32)]


class Medium(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("rho0", C.c_double), ("r0", C.c_double),
                ("sigma_r", C.c_double), ("h_over_r", C.c_double), ("radial_power", C.c_double),
                ("funnel_depletion", C.c_double), ("funnel_theta_rad", C.c_double),
                ("envelope_rho0", C.c_double), ("envelope_alpha", C.c_double), ("r_min", C.c_double),
                ("r_max", C.c_double), ("rho_floor", C.c_double)]


class Source(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("r_center", C.c_double),
                ("sigma_r", C.c_double), ("theta_width_rad", C.c_double), ("powerlaw", C.c_double),
                ("emax_GeV", C.c_double), ("norm", C.c_double), ("funnel_theta_rad", C.c_double),
                ("density_q", C.c_double), ("radial_s", C.c_double), ("rho_ref", C.c_double),
                ("cutoff_min", C.c_double), ("cutoff_max", C.c_double), ("gradient_dr", C.c_double),
                ("gradient_dtheta_rad", C.c_double)]


class Spectrum(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n_bins", C.c_int32), ("gamma", C.c_double),
                ("ecut_GeV", C.c_double), ("e_min_GeV", C.c_double), ("e_max_GeV", C.c_double)]


class Scene(C.Structure):
    _fields_ = [("medium", Medium), ("source", Source), ("spectrum", Spectrum),
                ("enu_obs_GeV", C.c_double), ("m_bh_msun", C.c_double)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Mev(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("model", C.c_int32), ("flavor", C.c_int32),
                ("thermal_profile", C.c_int32), ("ye_profile", C.c_int32), ("spectral_mode", C.c_int32),
                ("n_bins", C.c_int32), ("processes", C.c_uint32),
                ("enu_obs_MeV", C.c_double), ("norm", C.c_double), ("sigma_abs0_cm2", C.c_double),
                ("sigma_scat0_cm2", C.c_double), ("sigma_scat_e0_cm2", C.c_double),
                ("neutrinosphere_tau", C.c_double), ("T0_MeV", C.c_double), ("T_floor_MeV", C.c_double),
                ("T_power", C.c_double), ("Ye_torus", C.c_double), ("Ye_funnel", C.c_double),
                ("Ye_envelope", C.c_double), ("Ye_floor", C.c_double), ("Ye_ceil", C.c_double),
                ("E_min_MeV", C.c_double), ("E_max_MeV", C.c_double)]


class MevOut(C.Structure):
    _fields_ = [("tau", C.c_void_p), ("p_surv", C.c_void_p), ("intensity", C.c_void_p),
                ("r_neutrinosphere", C.c_void_p), ("leakage", C.c_void_p)]


def mev(model="physical", flavor="anti_nu_e", thermal="inner_hot_torus", ye="neutron_rich_torus",
        spectral="monochromatic", processes=None, enu=10.0, norm=1.0, sigma_abs0=9.6e-44, sigma_scat0=1.7e-44,
        sigma_scat_e0=3.4e-45, neutrinosphere_tau=2.0 / 3.0, T0=6.0, T_floor=0.1, T_power=0.2, Ye_torus=0.25,
        Ye_funnel=0.55, Ye_envelope=0.45, Ye_floor=0.01, Ye_ceil=0.60, E_min=3.0, E_max=50.0, n_bins=8):
    """hadros_mev from MeVThermalParams fields (include/radiative_transfer.hpp:66-106 defaults)."""
    m = Mev()
    m.enabled = 1
    m.model, m.flavor = MEV_MODELS[model], MEV_FLAVORS[flavor]
    m.thermal_profile, m.ye_profile, m.spectral_mode = MEV_THERMAL[thermal], MEV_YE[ye], MEV_SPECTRAL[spectral]
    m.n_bins = n_bins
    m.processes = MEV_DEFAULT_PROCESSES if processes is None else processes
    m.enu_obs_MeV, m.norm = enu, norm
    m.sigma_abs0_cm2, m.sigma_scat0_cm2, m.sigma_scat_e0_cm2 = sigma_abs0, sigma_scat0, sigma_scat_e0
    m.neutrinosphere_tau = neutrinosphere_tau
    m.T0_MeV, m.T_floor_MeV, m.T_power = T0, T_floor, T_power
    m.Ye_torus, m.Ye_funnel, m.Ye_envelope, m.Ye_floor, m.Ye_ceil = Ye_torus, Ye_funnel, Ye_envelope, Ye_floor, Ye_ceil
    m.E_min_MeV, m.E_max_MeV = E_min, E_max
    return m


class ImageOut(C.Structure):
    _fields_ = [("tau", _dp), ("p_surv", _dp), ("intensity", _dp), ("captured", _ip),
                ("npoints", _ip), ("flags", _ip), ("alpha", _dp), ("beta", _dp),
                ("bin_tau", _dp), ("bin_p", _dp), ("bin_i", _dp)]


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("total_ms", C.c_double), ("total_points", C.c_int64),
                ("total_attempts", C.c_int64), ("n_captured", C.c_int32), ("n_exhausted", C.c_int32),
                ("n_bins", C.c_int32), ("block_size", C.c_int32)]


class CacheView(C.Structure):
    _fields_ = [("n_rays", C.c_int32), ("reserved", C.c_int32), ("a_bh", C.c_double),
                ("offset", C.POINTER(C.c_int64)), ("r", _dp), ("theta", _dp), ("dl", _dp),
                ("redshift", _dp)]


def camera(a_spin=0.9, r_obs=60.0, theta_obs_deg=74.0, fov_deg=45.0, nx=100, ny=100, r_max=120.0,
           h=0.001, tol=None, mode=MODE_REF, max_steps=None):
    """Camera with the argv defaults of apps/compute_kerr_geodesics.cpp:26-34.

    theta is converted as the reference CLI does (M_PI * deg / 180, :74);
    REF mode uses tolerance 1e-6 (include/kerr_geodesic.hpp:12) and 200000
    steps (src/kerr_camera.cpp:222); CFG mode defaults to rtol 1e-8 and 20000.
    """
    if tol is None:
        tol = 1.0e-6 if mode == MODE_REF else 1.0e-8
    if max_steps is None:
        max_steps = 200000 if mode == MODE_REF else 20000
    return Camera(a_spin, r_obs, math.pi * theta_obs_deg / 180.0, fov_deg, nx, ny, r_max, h, tol,
                  mode, max_steps)


def scene(**kw):
    """Scene with the argv defaults of apps/compute_kerr_image_from_cache.cpp:33-55
    and the struct defaults of include/radiative_transfer.hpp:29-64.

    Keyword names follow the reference argv / Makefile variables, lower case;
    angles in degrees are converted with x * M_PI / 180 like the CLI.
    """
    d = dict(enu=1.0e9, m_bh=3.0, rho0=1.0e10, r0=10.0, sigma_r=5.0, h_over_r=0.01,
             profile="gaussian", radial_power=2.0, funnel_depletion=0.0, funnel_theta_deg=15.0,
             envelope_rho0=0.0, envelope_alpha=2.5, r_min=4.0, r_max=60.0, rho_floor=1.0e-99,
             source="inner_ring", source_r=3.5, source_sigma=1.0, source_theta_deg=8.0,
             source_powerlaw=2.0, source_emax=1.0e12, source_norm=1.0, source_funnel_theta_deg=20.0,
             source_q=1.0, source_s=2.0, source_rho_ref=-1.0, source_cutoff_min=0.0,
             source_cutoff_max=1.0e2, source_gradient_dr=0.1, source_gradient_dtheta_deg=1.0,
             spectrum="monochromatic", gamma=2.0, ecut=1.0e12, e_min=1.0e5, e_max=1.0e12, n_bins=8)
    for k in kw:
        if k not in d:
            raise KeyError(k)
    d.update(kw)
    s = Scene()
    m = s.medium
    m.kind = PROFILE_KINDS[d["profile"]] if isinstance(d["profile"], str) else int(d["profile"])
    m.rho0, m.r0, m.sigma_r, m.h_over_r = d["rho0"], d["r0"], d["sigma_r"], d["h_over_r"]
    m.radial_power, m.funnel_depletion = d["radial_power"], d["funnel_depletion"]
    m.funnel_theta_rad = d["funnel_theta_deg"] * math.pi / 180.0
    m.envelope_rho0, m.envelope_alpha = d["envelope_rho0"], d["envelope_alpha"]
    m.r_min, m.r_max, m.rho_floor = d["r_min"], d["r_max"], d["rho_floor"]
    so = s.source
    so.kind = SOURCE_KINDS[d["source"]] if isinstance(d["source"], str) else int(d["source"])
    so.r_center, so.sigma_r = d["source_r"], d["source_sigma"]
    so.theta_width_rad = d["source_theta_deg"] * math.pi / 180.0
    so.powerlaw, so.emax_GeV, so.norm = d["source_powerlaw"], d["source_emax"], d["source_norm"]
    so.funnel_theta_rad = d["source_funnel_theta_deg"] * math.pi / 180.0
    so.density_q, so.radial_s, so.rho_ref = d["source_q"], d["source_s"], d["source_rho_ref"]
    so.cutoff_min, so.cutoff_max = d["source_cutoff_min"], d["source_cutoff_max"]
    so.gradient_dr = d["source_gradient_dr"]
    so.gradient_dtheta_rad = d["source_gradient_dtheta_deg"] * math.pi / 180.0
    sp = s.spectrum
    sp.kind = SPECTRUM_KINDS[d["spectrum"]] if isinstance(d["spectrum"], str) else int(d["spectrum"])
    sp.n_bins = int(d["n_bins"])
    sp.gamma, sp.ecut_GeV, sp.e_min_GeV, sp.e_max_GeV = d["gamma"], d["ecut"], d["e_min"], d["e_max"]
    s.enu_obs_GeV, s.m_bh_msun = d["enu"], d["m_bh"]
    return s


def medium_params(m):
    """TorusProfile constructor argument list (include/torus_profile.hpp:19-33) of a Medium."""
    return [m.rho0, m.r0, m.sigma_r, m.h_over_r, m.radial_power, m.funnel_depletion, m.funnel_theta_rad,
            m.envelope_rho0, m.envelope_alpha, m.r_min, m.r_max, m.rho_floor]


def source_params(so):
    """UHESourceParams field list (include/radiative_transfer.hpp:47-64) of a Source."""
    return [so.r_center, so.sigma_r, so.theta_width_rad, so.powerlaw, so.emax_GeV, so.norm,
            so.funnel_theta_rad, so.density_q, so.radial_s, so.rho_ref, so.cutoff_min, so.cutoff_max,
            so.gradient_dr, so.gradient_dtheta_rad]


def spectrum_params(sp):
    return [sp.gamma, sp.ecut_GeV, sp.e_min_GeV, sp.e_max_GeV, float(sp.n_bins)]


def opacity_args(argv):
    """argv 1-19 of apps/extract_opacity_surface.cpp:72-93 (strings, numbers or
    defaults for missing entries) -> (Medium, enu_GeV, m_bh_msun, sigma_path,
    tau_surface, ntheta, nr)."""
    dflt = ["1e5", "3", "data/sigma/sigma_nuN_CC_GBW.dat", "gaussian", "1e-2", "10", "5", "0.25", "2", "0",
            "15", "0", "2.5", "4", "60", "1e-99", "1", "181", "1200"]
    a = [str(x) for x in argv] + dflt[len(argv):]
    m = Medium()
    m.kind = PROFILE_KINDS[a[3]]
    m.rho0, m.r0, m.sigma_r, m.h_over_r = float(a[4]), float(a[5]), float(a[6]), float(a[7])
    m.radial_power, m.funnel_depletion = float(a[8]), float(a[9])
    m.funnel_theta_rad = float(a[10]) * math.pi / 180.0
    m.envelope_rho0, m.envelope_alpha = float(a[11]), float(a[12])
    m.r_min, m.r_max, m.rho_floor = float(a[13]), float(a[14]), float(a[15])
    return m, float(a[0]), float(a[1]), a[2], float(a[16]), int(a[17]), int(a[18])
