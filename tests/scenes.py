"""Scenes and cameras shared by the parity tests.

The benchmark scenes (rho0 >= 1e8 g/cm^3) are opaque to UHE neutrinos on
almost every ray (P = 0 exactly), so the parity sets add low-density scenes
where tau ~ 1e-3 .. 30 and survival and intensity are non-trivial.
"""
import math

from hadros_amd import abi, workloads

SIG_E, SIG_G, SIG_S = workloads.synthetic_sigma()


def small_ref_camera(nx=6, ny=6, a=0.9, theta_deg=80.0, h=0.1):
    """The reference's Makefile camera (r_obs 60, fov 25 deg, r_max 120) at a coarse step."""
    return abi.camera(a_spin=a, r_obs=60.0, theta_obs_deg=theta_deg, fov_deg=25.0, nx=nx, ny=ny, r_max=120.0,
                      h=h, mode=abi.MODE_REF)


def cfg_camera(nx=24, ny=24, a=0.9, theta_deg=60.0, r_obs=1000.0, half=20.0):
    return abi.camera(a_spin=a, r_obs=r_obs, theta_obs_deg=theta_deg,
                      fov_deg=workloads.fov_for_half_width(half, r_obs), nx=nx, ny=ny, r_max=1.1 * r_obs, h=1.0,
                      mode=abi.MODE_CFG)


def parity_scenes():
    """(name, scene) pairs covering the 8 profiles, 5 sources and 3 spectra."""
    out = []
    profiles = abi.PROFILE_NAMES
    sources = abi.SOURCE_NAMES
    spectra = abi.SPECTRUM_NAMES
    for k, prof in enumerate(profiles):
        src = sources[k % len(sources)]
        spec = spectra[k % len(spectra)]
        rho0 = [1.0e-2, 1.0e1, 1.0e3][k % 3]
        sc = abi.scene(enu=1.0e9, rho0=rho0, r0=10.0, sigma_r=3.0, h_over_r=0.3, profile=prof, radial_power=2.0,
                       funnel_depletion=0.9, funnel_theta_deg=20.0, envelope_rho0=rho0 * 1e-2, envelope_alpha=2.5,
                       r_min=3.0, r_max=60.0, rho_floor=1.0e-3, source=src, source_r=6.0, source_sigma=1.5,
                       source_theta_deg=15.0, spectrum=spec, n_bins=8, e_min=1.0e5, e_max=1.0e12)
        out.append((f"{prof}-{src}-{spec}", sc))
    return out


def c1_lowdensity():
    sc = workloads.c1_scene()
    sc.medium.rho0 = 10.0
    return sc


DEG = math.pi / 180.0
