"""Concrete inputs of the benchmark workloads (cameras, scenes, sigma table).

The parameter mappings from the workload descriptions onto the reference's
parameters (tan-plane camera, m_u, funnel depletion amplitude, envelope start)
are the ones recorded in DESIGN.md; each function states them.
"""
import math

import numpy as np

from . import abi

# sigma_GeV^-2 = sigma_cm2 / (hbar c)^2 with (hbar c)^2 = 0.389379338e-27 cm^2 GeV^2
HBARC2_CM2_GEV2 = 0.389379338e-27


def synthetic_sigma(n=64, e_lo=1.0e4, e_hi=1.0e12, A=7.84e-36, B=0.363):
    """Power-law DIS table sigma = A E^B on n log-spaced points.

    Returns (E_GeV, sigma_GeV^-2, sigma_cm2), the three columns of the
    reference's table files (src/sigma_table.cpp:19-38).
    """
    k = np.arange(n)
    lo, hi = math.log10(e_lo), math.log10(e_hi)
    e = 10.0 ** (lo + (hi - lo) * k / (n - 1))
    s = A * e ** B
    return e, s / HBARC2_CM2_GEV2, s


def write_sigma_file(path, e, g, s, model="synthetic_powerlaw"):
    with open(path, "w") as f:
        f.write(f"# model {model}\n# Enu_GeV sigma_GeV_minus2 sigma_cm2\n")
        for a, b, c in zip(e, g, s):
            f.write(f"{a:.17g} {b:.17g} {c:.17g}\n")


def fov_for_half_width(half_width_rg, r_obs):
    """Full field of view [deg] whose tan-plane half-width at r_obs is half_width_rg
    (the camera's u, v are tan-plane coordinates, src/kerr_camera.cpp:112-116)."""
    return 2.0 * math.degrees(math.atan(half_width_rg / r_obs))


def c1_camera(nx=64, ny=64, mode=abi.MODE_CFG):
    """Spin 0.9, r_obs 1000, inclination 60 deg, alpha, beta in [-20, 20] M,
    stop at r < 1.01 r_+ or r > 1100, rtol 1e-8, <= 20000 steps."""
    return abi.camera(a_spin=0.9, r_obs=1000.0, theta_obs_deg=60.0, fov_deg=fov_for_half_width(20.0, 1000.0),
                      nx=nx, ny=ny, r_max=1100.0, h=1.0, mode=mode)


def c1_scene():
    """Gaussian torus r0 10, sigma_r 3, H/R 0.3, rho0 1e10, floor 1e-3;
    monochromatic E = 1e9 GeV; inner_ring source at 6 M, width 1 M, theta width 0.1 rad."""
    return abi.scene(enu=1.0e9, m_bh=3.0, rho0=1.0e10, r0=10.0, sigma_r=3.0, h_over_r=0.3, profile="gaussian",
                     rho_floor=1.0e-3, source="inner_ring", source_r=6.0, source_sigma=1.0,
                     source_theta_deg=math.degrees(0.1), source_powerlaw=2.0, source_emax=1.0e12, source_norm=1.0,
                     spectrum="monochromatic")


def c2_camera(nx=512, ny=512, mode=abi.MODE_CFG):
    """Spin 0.99, inclination 80 deg, camera otherwise as c1_camera."""
    return abi.camera(a_spin=0.99, r_obs=1000.0, theta_obs_deg=80.0, fov_deg=fov_for_half_width(20.0, 1000.0),
                      nx=nx, ny=ny, r_max=1100.0, h=1.0, mode=mode)


def c2_scene():
    """powerlaw_funnel_envelope: slope 1.5 from r 3 to 50, funnel half-angle 20 deg with
    depletion amplitude 1 - 1e-4, envelope 1e-2 rho0 from r0 = 30, floor 1e-3, rho0 1e10;
    powerlaw_cutoff gamma 2, Ecut 1e12, 8 bins on 1e5-1e12; funnel_wall source (reference defaults)."""
    return abi.scene(enu=1.0e9, m_bh=3.0, rho0=1.0e10, r0=30.0, sigma_r=3.0, h_over_r=0.3,
                     profile="powerlaw_funnel_envelope", radial_power=1.5, funnel_depletion=1.0 - 1.0e-4,
                     funnel_theta_deg=20.0, envelope_rho0=1.0e8, envelope_alpha=2.5, r_min=3.0, r_max=50.0,
                     rho_floor=1.0e-3, source="funnel_wall", source_r=3.5, source_sigma=1.0, source_theta_deg=15.0,
                     source_funnel_theta_deg=20.0, spectrum="powerlaw_cutoff", gamma=2.0, ecut=1.0e12,
                     e_min=1.0e5, e_max=1.0e12, n_bins=8)


def c3_manifest(seed=0, n_cameras=256):
    """The batched workload: n cameras drawn from numpy default_rng(seed).

    resolution uniform over {64, 128, 256, 512}^2, spin U[0, 0.998],
    inclination U[5, 85] deg, profile uniform over the 8 morphologies,
    log10 rho0 U[8, 12], source uniform over the 5 models, spectral model
    uniform over the 3, 8 bins on 1e5-1e12 GeV.  Draw order per camera:
    res, spin, inclination, profile, log rho0, source, spectrum.
    """
    rng = np.random.default_rng(seed)
    out = []
    for c in range(n_cameras):
        res = int(rng.choice([64, 128, 256, 512]))
        a = float(rng.uniform(0.0, 0.998))
        inc = float(rng.uniform(5.0, 85.0))
        prof = abi.PROFILE_NAMES[int(rng.integers(0, 8))]
        lrho = float(rng.uniform(8.0, 12.0))
        src = abi.SOURCE_NAMES[int(rng.integers(0, 5))]
        spec = abi.SPECTRUM_NAMES[int(rng.integers(0, 3))]
        out.append({"camera": c, "res": res, "spin": a, "inclination_deg": inc, "profile": prof,
                    "log10_rho0": lrho, "source": src, "spectrum": spec})
    return out


def manifest_camera(m, mode=abi.MODE_CFG, res=None):
    r = res or m["res"]
    return abi.camera(a_spin=m["spin"], r_obs=1000.0, theta_obs_deg=m["inclination_deg"],
                      fov_deg=fov_for_half_width(20.0, 1000.0), nx=r, ny=r, r_max=1100.0, h=1.0, mode=mode)


def manifest_scene(m, slope=2.0):
    """Scene of one manifest entry: the c2_scene medium conventions (funnel
    depletion 1 - 1e-4 at 20 deg, envelope 1e-2 rho0, r in [3, 50], floor 1e-3),
    the c1 inner_ring / c2 funnel_wall source widths, reference defaults for
    the other sources, gamma 2 and Ecut 1e12 for the spectra."""
    rho0 = 10.0 ** m["log10_rho0"]
    src = m["source"]
    kw = dict(source_r=3.5, source_sigma=1.0, source_theta_deg=15.0)
    if src == "inner_ring":
        kw = dict(source_r=6.0, source_sigma=1.0, source_theta_deg=math.degrees(0.1))
    return abi.scene(enu=1.0e9, m_bh=3.0, rho0=rho0, r0=10.0, sigma_r=3.0, h_over_r=0.3, profile=m["profile"],
                     radial_power=slope, funnel_depletion=1.0 - 1.0e-4, funnel_theta_deg=20.0,
                     envelope_rho0=1.0e-2 * rho0, envelope_alpha=2.5, r_min=3.0, r_max=50.0, rho_floor=1.0e-3,
                     source=src, source_funnel_theta_deg=20.0, spectrum=m["spectrum"], gamma=2.0, ecut=1.0e12,
                     e_min=1.0e5, e_max=1.0e12, n_bins=8, **kw)


WORKLOADS = {
    "c1": (c1_camera, c1_scene),
    "c2": (c2_camera, c2_scene),
}
