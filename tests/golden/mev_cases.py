"""MeV channel cases shared by make_golden.py (which writes mev.npz from the
compiled reference) and the tests that read it: parameter sets covering every
model, flavor, thermal and Ye profile and both spectral modes, and the media
they are evaluated in."""
import numpy as np

from hadros_amd import abi


# MeV channel cases: (model, flavor, thermal, ye, spectral) names and argv-style
# overrides of MeVThermalParams (include/radiative_transfer.hpp:66-106)
MEV_CASES = {
    "default": (("physical", "anti_nu_e", "inner_hot_torus", "neutron_rich_torus", "monochromatic"), {}),
    "nue_degenerate": (("physical", "nu_e", "radial_powerlaw", "funnel_proton_rich", "monochromatic"),
                       dict(enu=20.0, degeneracy=1, T0=8.0, T_power=0.35)),
    "nux_band": (("physical", "nu_x", "torus_plus_cool_envelope", "torus_envelope_contrast", "fermi_dirac_band"),
                 dict(E_min=3.0, E_max=50.0, n_bins=8)),
    "toy_collapsar": (("toy", "anti_nu_e", "collapsar_inner_hot", "collapsar_neutron_rich", "monochromatic"),
                      dict(enu=5.0, norm=2.5)),
    "constant_band_subset": (("physical", "anti_nu_e", "constant", "constant", "fermi_dirac_band"),
                             dict(E_min=1.0, E_max=80.0, n_bins=16, brems=0, scat_e=0, T0=4.0, Ye_torus=0.4)),
    "nue_absorption_only": (("physical", "nu_e", "inner_hot_torus", "neutron_rich_torus", "mono"),
                            dict(enu=30.0, scattering=0, abs_p=0)),
    "toy_band": (("TOY", "nux", "RADIAL_POWERLAW", "FUNNEL_PROTON_RICH", "fd_band"),
                 dict(E_min=2.0, E_max=40.0, n_bins=5, absorption=0)),
}
MEV_BITS = ["urca", "pair", "brems", "absorption", "scattering", "degeneracy", "abs_n", "abs_p", "scat_n",
            "scat_p", "scat_e"]


def mev_params(over):
    """mp array of the reference harness (oracle/ref_harness.cpp:fill_mev) and the abi.mev kwargs."""
    d = dict(enu=10.0, norm=1.0, sigma_abs0=9.6e-44, sigma_scat0=1.7e-44, sigma_scat_e0=3.4e-45,
             neutrinosphere_tau=2.0 / 3.0, T0=6.0, T_floor=0.1, T_power=0.2, Ye_torus=0.25, Ye_funnel=0.55,
             Ye_envelope=0.45, Ye_floor=0.01, Ye_ceil=0.60, E_min=3.0, E_max=50.0, n_bins=8)
    flags = dict(urca=1, pair=1, brems=1, absorption=1, scattering=1, degeneracy=0, abs_n=1, abs_p=1, scat_n=1,
                 scat_p=1, scat_e=1)
    for k, v in over.items():
        (flags if k in flags else d)[k] = v
    bits = sum(1 << i for i, b in enumerate(MEV_BITS) if flags[b])
    mp = [d["enu"], d["norm"], d["sigma_abs0"], d["sigma_scat0"], d["sigma_scat_e0"], d["neutrinosphere_tau"],
          d["T0"], d["T_floor"], d["T_power"], d["Ye_torus"], d["Ye_funnel"], d["Ye_envelope"], d["Ye_floor"],
          d["Ye_ceil"], d["E_min"], d["E_max"], float(d["n_bins"]), float(bits)]
    return np.array(mp), d, flags, bits


# media of the MeV fixtures: (profile, rho0) with the image fixture's other parameters
MEV_MEDIA = [("gaussian", 1e11), ("collapsar_ndaf_like", 1e12), ("powerlaw_funnel_envelope", 1e10)]


def mev_scene(profile, rho0, source="inner_ring", spectrum="monochromatic"):
    return abi.scene(enu=1e9, m_bh=3.0, rho0=rho0, r0=10.0, sigma_r=5.0, h_over_r=0.25, profile=profile,
                     radial_power=2.0, funnel_depletion=0.9, funnel_theta_deg=20.0, envelope_rho0=1e-2 * rho0,
                     envelope_alpha=2.5, r_min=4.0, r_max=60.0, rho_floor=1e-3, source=source, source_r=6.0,
                     source_sigma=1.5, source_theta_deg=15.0, source_powerlaw=2.0, source_emax=1e12,
                     source_norm=1.0, source_funnel_theta_deg=20.0, source_q=1.0, source_s=2.0,
                     source_gradient_dr=0.1, source_gradient_dtheta_deg=1.0, source_rho_ref=-1.0,
                     source_cutoff_min=0.0, source_cutoff_max=100.0, spectrum=spectrum, gamma=2.0, ecut=1e12,
                     e_min=1e5, e_max=1e12, n_bins=8)


def mev_struct(names, over):
    """abi.Mev of a case."""
    mp, d, flags, bits = mev_params(over)
    return abi.mev(model=names[0], flavor=names[1], thermal=names[2], ye=names[3], spectral=names[4],
                   processes=bits, enu=d["enu"], norm=d["norm"], sigma_abs0=d["sigma_abs0"],
                   sigma_scat0=d["sigma_scat0"], sigma_scat_e0=d["sigma_scat_e0"],
                   neutrinosphere_tau=d["neutrinosphere_tau"], T0=d["T0"], T_floor=d["T_floor"],
                   T_power=d["T_power"], Ye_torus=d["Ye_torus"], Ye_funnel=d["Ye_funnel"],
                   Ye_envelope=d["Ye_envelope"], Ye_floor=d["Ye_floor"], Ye_ceil=d["Ye_ceil"], E_min=d["E_min"],
                   E_max=d["E_max"], n_bins=d["n_bins"])
