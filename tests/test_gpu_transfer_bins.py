"""GPU parity of the split tracer's transfer over every bin layout and sigma
lookup, against the FP64 CPU oracle.

The transfer kernel lays a chunk's (point, bin) pairs out bin-major: with NBP
= n_bins rounded up to a power of two, each bin owns 64 / NBP lanes, and
sixteen bins take each chunk in two halves.  Every lane count (64, 32, 16, 8,
4) is traced here, with each of the four sigma lookups the host can select:

  powerlaw  a table on one power law (the benchmark's A E^0.363): n_N dl z^s
            per point times sigma(E_mid) per bin
  exact     log-uniform nodes, not a power law: interval by one ceil
  guess     nearly log-uniform nodes (the reference's printed GBW table):
            index guess and one comparison each way
  search    uneven nodes: binary search

Both precision forms.  Bars (written here): the image bars of the parity
tests (FP64: rel tau, survival and sum I <= 1e-6 on rays whose point count
equals the oracle's; FP32-Kahan: 2e-5, 5e-5 and 1e-5), plus rel(sum over
pixels of I_b) <= 1e-6 (FP32-Kahan 1e-5) for every bin with a non-zero
total.
The power-law lookup and the exact lookup of the same table, nudged off its
power law by 1e-11, agree to 1e-9 relative in every output.
"""
import math
import os

import numpy as np
import pytest

from hadros_amd import abi, ffi
from oracle import py_oracle as po

from image_metrics import compare_images, rel
from scenes import SIG_E, SIG_S, cfg_camera

pytestmark = pytest.mark.gpu

_GBW = np.load(os.path.join(os.path.dirname(__file__), "golden", "sigma_GBW.npz"))


def _tables():
    k = np.arange(SIG_E.size)
    rng = np.random.default_rng(7)
    uneven = np.sort(10 ** rng.uniform(4, 12, 40))
    return {
        "powerlaw": (SIG_E, SIG_S),
        "exact": (SIG_E, SIG_S * (1.0 + 0.2 * np.sin(0.7 * k))),
        "guess": (np.ascontiguousarray(_GBW["e"]), np.ascontiguousarray(_GBW["s"])),
        "search": (uneven, 5.53e-36 * uneven ** 0.363 * (1 + 0.3 * rng.uniform(size=40))),
    }


# densities [g/cm^3] at which each table leaves survival spread over (0, 1)
RHO0 = {"powerlaw": 100.0, "exact": 100.0, "guess": 3.0, "search": 100.0}


def _scene(n_bins, rho0=100.0):
    # a translucent powerlaw_funnel_envelope medium
    return abi.scene(enu=1.0e9, rho0=rho0, r0=10.0, sigma_r=3.0, h_over_r=0.3, profile="powerlaw_funnel_envelope",
                     radial_power=1.5, funnel_depletion=0.9999, funnel_theta_deg=20.0, envelope_rho0=1.0e-2 * rho0,
                     envelope_alpha=2.5, r_min=3.0, r_max=50.0, rho_floor=1.0e-3, source="funnel_wall",
                     source_r=6.0, source_sigma=1.5, source_theta_deg=15.0, spectrum="powerlaw_cutoff",
                     n_bins=n_bins, e_min=1.0e5, e_max=1.0e12)


def _bin_totals_close(img, ref, label, tol=1e-6):
    nb = ref.bin_i.shape[1]
    for b in range(nb):
        s, sr = img.bin_i[:, b].sum(), ref.bin_i[:, b].sum()
        if sr != 0.0:
            assert abs(s - sr) <= tol * abs(sr), f"{label}: bin {b} sum I {s} vs {sr}"


@pytest.mark.parametrize("n_bins", [1, 2, 3, 5, 8, 12, 16])
@pytest.mark.parametrize("table", ["powerlaw", "exact", "guess", "search"])
@pytest.mark.parametrize("prec", ["fp64", "fp32k"])
def test_transfer_bins_and_lookups(prec, table, n_bins):
    se, ss = _tables()[table]
    cam = cfg_camera(nx=20, ny=20, a=0.95, theta_deg=70.0)
    sc = _scene(n_bins, RHO0[table])
    p = abi.PREC_FP64 if prec == "fp64" else abi.PREC_FP32K
    img, st = ffi.trace_rt(cam, sc, se, ss, p, path=abi.PATH_SPLIT)
    ref, rst = po.trace_image(cam, sc, se, ss)
    assert st.n_bins == rst.n_bins == n_bins
    assert 0.01 < ref.p.mean() < 0.99, "the medium should be translucent"
    label = f"{prec} {table} nb={n_bins}"
    if prec == "fp64":
        rep = compare_images(img, ref, label=label)
        _bin_totals_close(img, ref, label)
    else:   # the FP32-Kahan bars of the parity tests
        rep = compare_images(img, ref, tol_tau=2e-5, tol_p=5e-5, tol_i=1e-5, label=label)
        _bin_totals_close(img, ref, label, tol=1e-5)
    print(rep)


@pytest.mark.parametrize("n_bins", [1, 8, 16])
def test_powerlaw_lookup_matches_table_lookup(n_bins):
    """The same camera with the power-law table and with that table nudged
    by 1e-11 at one interior node (off the power law, so the kernel takes
    the exact-uniform lookup): every ray's outputs agree to 1e-9."""
    cam = cfg_camera(nx=16, ny=16, a=0.9, theta_deg=60.0)
    sc = _scene(n_bins)
    nudged = SIG_S.copy()
    nudged[SIG_S.size // 2] *= 1.0 + 1e-11
    a, sa = ffi.trace_rt(cam, sc, SIG_E, SIG_S, abi.PREC_FP64, path=abi.PATH_SPLIT)
    b, sb = ffi.trace_rt(cam, sc, SIG_E, nudged, abi.PREC_FP64, path=abi.PATH_SPLIT)
    assert sa.total_points == sb.total_points
    assert np.array_equal(a.npoints, b.npoints)
    for f in ("tau", "p", "I", "bin_tau", "bin_p", "bin_i"):
        x, y = getattr(a, f), getattr(b, f)
        assert rel(x, y, floor=1e-300).max() <= 1e-9 or np.abs(x - y).max() <= 1e-300, f
    assert math.isfinite(a.I.sum()) and a.I.sum() > 0.0


@pytest.mark.parametrize("prec", ["fp64", "fp32k"])
def test_spectrum_outside_table_integrates_nothing(prec):
    """A spectrum whose range misses the sigma table integrates no bin
    (src/radiative_transfer.cpp:698-705): every ray keeps the defaults, tau 0,
    P 1, I 0, as in the oracle."""
    cam = cfg_camera(nx=12, ny=12, a=0.9, theta_deg=60.0)
    sc = _scene(8)
    sc.spectrum.e_min_GeV, sc.spectrum.e_max_GeV = 1.0e13, 1.0e14
    p = abi.PREC_FP64 if prec == "fp64" else abi.PREC_FP32K
    img, st = ffi.trace_rt(cam, sc, SIG_E, SIG_S, p, path=abi.PATH_SPLIT)
    ref, rst = po.trace_image(cam, sc, SIG_E, SIG_S)
    assert st.n_bins in (0, -1) and rst.n_bins in (0, -1)
    assert np.array_equal(img.captured, ref.captured)
    for f in ("tau", "p", "I"):
        assert np.array_equal(getattr(img, f), getattr(ref, f)), f
    assert np.all(img.tau == 0.0) and np.all(img.p == 1.0) and np.all(img.I == 0.0)
