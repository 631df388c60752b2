"""CPU: pin the oracle's MeV channel (src/mev_neutrino_physics.cpp and the MeV
block of src/radiative_transfer.cpp:500-640) against the reference's own code
compiled in this container (tests/golden/mev.npz, written by
tests/golden/make_golden.py --only mev), bit for bit, and against what the
formulas fix on their own."""
import math
import os
import sys

import numpy as np
import pytest

from hadros_amd import abi, ffi, workloads
from oracle import py_oracle as po

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
import mev_cases as mc  # noqa: E402

MEV = np.load(os.path.join(GOLD, "mev.npz"))


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    same = (a.view(np.int64) == b.view(np.int64)) | (np.isnan(a) & np.isnan(b)) | ((a == 0) & (b == 0))
    return bool(same.all())


def ref_paths():
    """The 16 rays of the reference's small KGEO cache, retraced by the oracle
    in REF mode (bit-exact with the reference's cache: test_oracle_golden)."""
    d = np.load(os.path.join(GOLD, "kgeo_small.npz"))
    a, r_obs, th_deg, fov, nx, ny, r_max, h = [float(x) for x in d["argv"]]
    cam = abi.camera(a_spin=a, r_obs=r_obs, theta_obs_deg=th_deg, fov_deg=fov, nx=int(nx), ny=int(ny), r_max=r_max,
                     h=h, mode=abi.MODE_REF)
    return a, [po.trace_pixel(cam, int(i), int(j))[0] for (_, i, j, _, _) in d["rays"]]


def view_of(a_bh, paths):
    """hadros_cache_view of 7-double paths (the returned Cache keeps the arrays alive)."""
    off = np.zeros(len(paths) + 1, np.int64)
    off[1:] = np.cumsum([p.shape[0] for p in paths])
    pts = np.concatenate(paths)
    c = ffi.Cache(a_bh, off, pts[:, 0], pts[:, 1], pts[:, 5], pts[:, 6])
    return c.view(), c


@pytest.mark.parametrize("name", list(mc.MEV_CASES))
def test_mev_point_bitexact(name):
    """T and Ye of every profile; kappa and j at E = E_obs / g for the
    monochromatic cases (the band is pinned through whole rays below)."""
    names, over = mc.MEV_CASES[name]
    m = mc.mev_struct(names, over)
    mono = abi.MEV_SPECTRAL[names[4]] == 0
    for mi, (prof, rho0) in enumerate(mc.MEV_MEDIA):
        sc = mc.mev_scene(prof, rho0)
        got = po.mev_point(sc, m, MEV["r"], MEV["th"], MEV["g"])
        ref = MEV[f"{name}_point{mi}"]
        cols = [0, 1, 2, 3] if mono else [0, 1]
        assert bits_equal(got[:, cols], ref[:, cols]), (name, prof)


@pytest.mark.parametrize("name", list(mc.MEV_CASES))
def test_mev_rays_bitexact(name):
    """Whole rays through integrate_kerr_ray_spectral: a dense medium whose
    rays cross the neutrinosphere and a thin one with a binned UHE spectrum."""
    names, over = mc.MEV_CASES[name]
    m = mc.mev_struct(names, over)
    a, paths = ref_paths()
    v, keep = view_of(a, paths)
    e, _, s = workloads.synthetic_sigma()
    for si, (prof, rho0, spec) in enumerate((("gaussian", 1e12, "monochromatic"),
                                             ("powerlaw_funnel_envelope", 1e3, "powerlaw_cutoff"))):
        sc = mc.mev_scene(prof, rho0, "funnel_wall", spec)
        got = po.rt_from_cache_mev(v, sc, e, s, m)
        assert bits_equal(got, MEV[f"{name}_rt{si}"][:, 3:8]), (name, si)
    if name == "default":
        assert (MEV["default_rt0"][:, 6] > 0).sum() >= 4   # the dense medium puts neutrinospheres on the rays


def test_mev_columns_of_reference_cli_rows():
    """The MeV columns of the reference CLI's rows (%.8e) with argv 45-71 set."""
    a, paths = ref_paths()
    v, keep = view_of(a, paths)
    e, _, s = workloads.synthetic_sigma()
    sc = mc.mev_scene("gaussian", 1e12, "funnel_wall", "monochromatic")
    d = np.load(os.path.join(GOLD, "kgeo_small.npz"))
    for name, (names, over) in mc.MEV_CASES.items():
        got = po.rt_from_cache_mev(v, sc, e, s, mc.mev_struct(names, over))
        rows = {(int(r[0]), int(r[1])): r for r in MEV[name + "_rows"]}
        for k, (_, i, j, _, _) in enumerate(d["rays"]):
            rw = rows[(int(i), int(j))]
            for c in range(5):
                assert float("%.8e" % got[k, c]) == rw[8 + c], (name, k, c)


def test_mev_closed_forms():
    """What the formulas fix: with the constant thermal and Ye profiles T = T0
    and Ye = Ye_torus everywhere; a zero norm gives j = 0; the toy absorption
    (scattering off) is n_b sigma_abs0 E^2 ((1 - Ye) + Ye / 4); the optical
    depth of a ray is linear in sigma (absorption only, so tau doubles exactly
    when sigma_abs0 does), and P = exp(-tau)."""
    sc = mc.mev_scene("gaussian", 1e11)
    r = np.array([8.0, 10.0, 14.0]), np.array([1.5, 1.57, 1.62])
    m = abi.mev(thermal="constant", ye="constant", T0=4.5, Ye_torus=0.33, norm=0.0)
    out = po.mev_point(sc, m, r[0], r[1], np.ones(3))
    assert np.all(out[:, 0] == 4.5) and np.all(out[:, 1] == 0.33) and np.all(out[:, 3] == 0.0)
    m = abi.mev(model="toy", thermal="constant", ye="constant", Ye_torus=0.2, enu=12.0,
                processes=abi.MEV_ABSORPTION)
    out = po.mev_point(sc, m, r[0], r[1], np.ones(3))
    rho = np.array([po.lib().or_rho(sc.medium, r[0][k], r[1][k]) for k in range(3)])
    want = rho / 1.66053906660e-24 * 9.6e-44 * 144.0 * (0.8 + 0.25 * 0.2)
    assert np.allclose(out[:, 2], want, rtol=1e-14)
    a, paths = ref_paths()
    v, keep = view_of(a, paths)
    e, _, s = workloads.synthetic_sigma()
    base = dict(model="toy", thermal="constant", ye="constant", processes=abi.MEV_ABSORPTION)
    t1 = po.rt_from_cache_mev(v, sc, e, s, abi.mev(sigma_abs0=1e-44, **base))
    t2 = po.rt_from_cache_mev(v, sc, e, s, abi.mev(sigma_abs0=2e-44, **base))
    assert np.allclose(t2[:, 0], 2.0 * t1[:, 0], rtol=1e-13)
    assert np.allclose(t1[:, 1], np.exp(-t1[:, 0]), rtol=4e-16, atol=0)
    assert np.all(t1[:, 4] == 1.0)   # leakage: two identical sums
