"""GPU parity: the HIP path through the C-ABI against the FP64 CPU oracle.

Tolerances (written here, applied per test):
  point-wise metric / rho / sigma / emissivity     rel 1e-12 (libm vs device ulps)
  rhs, REF (finite differences amplify ulps 1e5x)  rel 1e-8
  rhs, CFG (analytic)                              rel 1e-12
  traced images (every ray bounded, _compare_images):
    FP64, rays whose point count equals the oracle's   rel(tau) <= 1e-6 where tau_ref > 0;
                                                       rel(P) <= 1e-6 max(1, tau_ref) where
                                                       P_ref > 1e-300 (dP/P = -dtau), else |dP| <= 1e-6;
                                                       the same per energy bin
    FP64, "chaotic" rays (the adaptive step sequence   rel(tau) <= 1e-3, |dP| <= 1e-4, and at most
      forked on a rounding difference, so the two      1% of the rays (10% in REF mode)
      sides integrate different points)
    FP64, "gate" rays (a point within rounding of a   rel(tau) <= 1e-4, survival 1e-4, and at most
      hard density gate, image_metrics)                0.1% of the rays
    FP64 image                                         rel(sum I) <= 1e-6, captured flags equal
    FP32-Kahan                                         rel(tau) <= 2e-5, rel(P) <= 5e-5 max(1, tau_ref)
                                                       (|dP| <= 5e-5 where P_ref <= 1e-300),
                                                       rel(sum I) <= 1e-5
  cache replay, FP64 (same points on both sides)   rel 1e-12
"""
import numpy as np
import pytest

from hadros_amd import abi, ffi, workloads
from oracle import py_oracle as po

from image_metrics import compare_images as _compare_images, rel
from scenes import SIG_E, SIG_S, cfg_camera, parity_scenes, small_ref_camera

pytestmark = pytest.mark.gpu


def test_device_is_gfx950():
    name, cus, arch = ffi.device_info()
    print(name, cus)
    assert arch == 950


def test_eval_metric_and_rhs():
    L = po.lib()
    rng = np.random.default_rng(0)
    for a in (0.0, 0.5, 0.9, 0.998):
        r = rng.uniform(1.05 * (1 + np.sqrt(1 - a * a)) + 0.1, 200.0, 500)
        th = rng.uniform(0.05, np.pi - 0.05, 500)
        g = ffi.eval_metric(a, r, th)
        ref = np.zeros((r.size, 10))
        for k in range(r.size):
            L.or_metric(a, r[k], th[k], po.dptr(ref[k, :5]))
            L.or_inverse_metric(a, r[k], th[k], po.dptr(ref[k, 5:]))
        m = np.abs(ref) > 1e-300
        assert rel(g[m], ref[m]).max() < 1e-12
        y = np.zeros((r.size, 8))
        y[:, 1], y[:, 2], y[:, 3] = r, th, rng.uniform(0, 6, r.size)
        y[:, 4] = -1.0
        y[:, 5], y[:, 6], y[:, 7] = rng.normal(size=r.size), 4 * rng.normal(size=r.size), 3 * rng.normal(size=r.size)
        for mode, tol in ((abi.MODE_REF, 1e-8), (abi.MODE_CFG, 1e-12)):
            d = ffi.eval_rhs(a, mode, y)
            dref = np.zeros_like(y)
            for k in range(r.size):
                L.or_rhs(a, mode, po.dptr(y[k]), po.dptr(dref[k]))
            for c in (1, 2, 3, 5, 6):
                scale = np.abs(dref[:, c]) + 1e-3 * np.abs(dref[:, c]).max()
                assert (np.abs(d[:, c] - dref[:, c]) / scale).max() < tol, (mode, c)


@pytest.mark.parametrize("name,scene", parity_scenes())
def test_eval_rho_emissivity(name, scene):
    L = po.lib()
    rng = np.random.default_rng(1)
    r = rng.uniform(1.2, 80.0, 2000)
    th = rng.uniform(0.0, np.pi, 2000)
    e = 10 ** rng.uniform(3, 13, 2000)
    g = ffi.eval_rho(scene.medium, r, th)
    ref = np.array([L.or_rho(scene.medium, r[k], th[k]) for k in range(r.size)])
    assert rel(g, ref).max() < 1e-12
    j = ffi.eval_emissivity(scene, r, th, e)
    jref = np.array([L.or_emissivity(scene, r[k], th[k], e[k]) for k in range(r.size)])
    assert rel(j, jref, 1e-300).max() < 1e-11


def test_eval_sigma():
    rng = np.random.default_rng(2)
    e = 10 ** rng.uniform(4, 12, 5000)
    e[:3] = [SIG_E[0], SIG_E[-1], SIG_E[17]]
    g = ffi.eval_sigma(SIG_E, SIG_S, e)
    ref = np.array([po.lib().or_sigma(po.dptr(SIG_E), po.dptr(SIG_S), SIG_E.size, x) for x in e])
    assert rel(g, ref).max() < 1e-13
    # exact power law: log-log interpolation reproduces 7.84e-36 E^0.363
    assert rel(g, 7.84e-36 * e ** 0.363).max() < 1e-12
    # outside the table the reference throws; the device hook returns NaN
    out = ffi.eval_sigma(SIG_E, SIG_S, np.array([1.0, 1e13]))
    assert np.isnan(out).all()


def test_eval_sigma_lookup_paths():
    """The index-guess lookup (tables uniform in ln E) and the binary search
    (other tables) both bracket exactly like the oracle's lower_bound: every
    node, every midpoint, and values one ulp either side of a node."""
    rng = np.random.default_rng(3)
    uneven = np.sort(10 ** rng.uniform(4, 12, 40))
    uneven_s = 5.53e-36 * uneven ** 0.363 * (1 + 0.3 * rng.uniform(size=40))
    for se, ss in ((SIG_E, SIG_S), (uneven, uneven_s)):
        mids = np.sqrt(se[1:] * se[:-1])
        e = np.concatenate([se, mids, np.nextafter(se[1:], 0), np.nextafter(se[:-1], np.inf)])
        g = ffi.eval_sigma(se, ss, e)
        ref = np.array([po.lib().or_sigma(po.dptr(se), po.dptr(ss), se.size, x) for x in e])
        assert rel(g, ref).max() < 1e-13


PATHS = [("fused", abi.PATH_FUSED), ("split", abi.PATH_SPLIT)]


@pytest.mark.parametrize("path", [p for _, p in PATHS], ids=[n for n, _ in PATHS])
@pytest.mark.parametrize("name,scene", parity_scenes())
def test_trace_cfg_fp64(name, scene, path):
    cam = cfg_camera(nx=20, ny=20, a=0.95, theta_deg=70.0)
    img, st = ffi.trace_rt(cam, scene, SIG_E, SIG_S, abi.PREC_FP64, path=path)
    ref, rst = po.trace_image(cam, scene, SIG_E, SIG_S)
    assert st.n_bins == rst.n_bins
    _compare_images(img, ref, label=name)
    print(st.total_points, rst.total_points, st.kernel_ms)


def test_trace_workloads_fp64():
    e, _, s = workloads.synthetic_sigma()
    for name in ("c1", "c2"):
        camf, scf = workloads.WORKLOADS[name]
        cam = camf(nx=32, ny=32)
        sc = scf()
        img, st = ffi.trace_rt(cam, sc, e, s, abi.PREC_FP64)
        ref, rst = po.trace_image(cam, sc, e, s)
        _compare_images(img, ref, label=name)
        assert st.total_points == rst.total_points or abs(st.total_points - rst.total_points) < 0.01 * rst.total_points


@pytest.mark.parametrize("path", [p for _, p in PATHS], ids=[n for n, _ in PATHS])
@pytest.mark.parametrize("name,scene", parity_scenes()[:4])
def test_trace_ref_fp64(name, scene, path):
    cam = small_ref_camera(nx=6, ny=6)
    img, st = ffi.trace_rt(cam, scene, SIG_E, SIG_S, abi.PREC_FP64, path=path)
    ref, rst = po.trace_image(cam, scene, SIG_E, SIG_S)
    _compare_images(img, ref, max_chaotic=0.1, label=name)
    print(st.total_points, rst.total_points)


@pytest.mark.parametrize("path", [p for _, p in PATHS], ids=[n for n, _ in PATHS])
@pytest.mark.parametrize("name,scene", parity_scenes())
def test_trace_fp32k(name, scene, path):
    cam = cfg_camera(nx=20, ny=20, a=0.95, theta_deg=70.0)
    img, st = ffi.trace_rt(cam, scene, SIG_E, SIG_S, abi.PREC_FP32K, path=path)
    ref, rst = po.trace_image(cam, scene, SIG_E, SIG_S)
    _compare_images(img, ref, tol_tau=2e-5, tol_p=5e-5, tol_i=1e-5, label=f"{name} FP32K")


def test_replay_matches_oracle():
    cam = cfg_camera(nx=16, ny=16, a=0.9, theta_deg=60.0)
    rays = []
    for k in range(cam.nx * cam.ny):
        pts, cp, ex, at = po.trace_pixel(cam, k // cam.ny, k % cam.ny)
        rays.append(pts)
    off = np.zeros(len(rays) + 1, np.int64)
    off[1:] = np.cumsum([p.shape[0] for p in rays])
    allp = np.concatenate(rays)
    cache = ffi.Cache(cam.a_spin, off, allp[:, 0], allp[:, 1], allp[:, 5], allp[:, 6])
    for name, sc in parity_scenes():
        for prec, tol in ((abi.PREC_FP64, 1e-12), (abi.PREC_FP32K, 1e-4)):
            img, st = ffi.rt_from_cache(cache, sc, SIG_E, SIG_S, prec)
            ref = po.rt_from_cache(cache.view(), sc, SIG_E, SIG_S)
            m = ref.tau > 0
            assert rel(img.tau[m], ref.tau[m]).max() < tol, name
            assert np.abs(img.p - ref.p).max() < tol, name
            mi = np.abs(ref.I) > 0
            if mi.any():
                assert rel(img.I[mi], ref.I[mi]).max() < (1e-11 if prec == abi.PREC_FP64 else 1e-4), name


def test_split_pool_grows_and_matches_fused():
    """The split form starts with a pool of 3 chunks per pixel; a camera whose
    rays need more (long REF-mode rays) triggers the grow-and-rerun path.
    Split and fused forms integrate the same points: agreement to rounding."""
    cam = small_ref_camera(nx=6, ny=6)
    sc = parity_scenes()[1][1]
    a, sa = ffi.trace_rt(cam, sc, SIG_E, SIG_S, abi.PREC_FP64, path=abi.PATH_FUSED)
    b, sb = ffi.trace_rt(cam, sc, SIG_E, SIG_S, abi.PREC_FP64, path=abi.PATH_SPLIT)
    assert sb.total_points / (cam.nx * cam.ny) > 3 * 64     # needed the pool to grow
    assert np.array_equal(a.captured, b.captured) and np.array_equal(a.npoints, b.npoints)
    assert sa.total_attempts == sb.total_attempts
    m = a.tau > 0
    assert rel(b.tau[m], a.tau[m]).max() < 1e-12 and np.abs(b.p - a.p).max() < 1e-12
    assert abs(b.I.sum() - a.I.sum()) <= 1e-12 * abs(a.I.sum())
    t = ffi.Tracer(cam, sc, SIG_E, SIG_S, abi.PREC_FP64, path=abi.PATH_SPLIT)
    for _ in range(2):                                       # pool kept from the first run
        t.run()
        c, _ = t.fetch()
        assert np.array_equal(c.p, b.p) and np.array_equal(c.tau, b.tau)
    t.close()


def test_split_segments_bitwise_invariant():
    """The split form's tile-row bands are scheduling only: every segment
    count gives the same image bit for bit, on ragged cameras (a partial last
    tile row and tile column): 70 x 54 (9 tile rows: the automatic bands are
    equal ones) and 130 x 44 (17 tile rows: the automatic bands split at 3/8,
    1/2 and 5/8 of them), in both precision forms, with the thin medium so that
    every output column is non-trivial.  Counts beyond the camera's tile rows
    are cut to them; counts outside [0, 4] are rejected."""
    sc = workloads.c2_scene()
    sc.medium.rho0 *= 1e-8
    sc.medium.envelope_rho0 *= 1e-8
    cols = ("tau", "p", "I", "captured", "npoints", "flags", "alpha", "beta", "bin_tau", "bin_p", "bin_i")
    for cam, prec in ((cfg_camera(nx=70, ny=54), abi.PREC_FP64), (cfg_camera(nx=70, ny=54), abi.PREC_FP32K),
                      (cfg_camera(nx=130, ny=44), abi.PREC_FP64)):
        base = None
        for nseg in (1, 0, 2, 3, 4):
            t = ffi.Tracer(cam, sc, SIG_E, SIG_S, prec, path=abi.PATH_SPLIT, segments=nseg)
            for _ in range(2):
                t.run()
                img, st = t.fetch()
                g, r = t.kernel_ms()
                tot = t.last_ms()
                assert 0.0 < g <= tot and 0.0 < r <= tot, (nseg, g, r, tot)
                if nseg == 1:
                    assert g + r <= tot * 1.01 + 0.01, (g, r, tot)
                if base is None:
                    base = img
                    assert 0.0 < img.p.mean() < 1.0
                    continue
                for c in cols:
                    assert np.array_equal(getattr(img, c), getattr(base, c)), (cam.nx, prec, nseg, c)
            t.close()
    t = ffi.Tracer(cfg_camera(nx=70, ny=54), sc, SIG_E, SIG_S, abi.PREC_FP64, path=abi.PATH_SPLIT)
    for bad in (-1, abi.TRACER_MAX_SEGMENTS + 1):
        with pytest.raises(ffi.HadrosError):
            t.set_segments(bad)
    t.close()


def test_trace_cache_matches_oracle_points():
    cam = small_ref_camera(nx=4, ny=4)
    cache = ffi.trace_cache(cam)
    same = 0
    for k in range(cam.nx * cam.ny):
        pts, cp, ex, at = po.trace_pixel(cam, k // cam.ny, k % cam.ny)
        g = np.stack([cache.r, cache.theta, cache.x, cache.y, cache.z, cache.dl, cache.redshift], 1)[
            cache.offset[k]:cache.offset[k + 1]]
        assert cache.captured[k] == cp
        if g.shape == pts.shape:
            same += 1
            n = min(200, g.shape[0])
            assert np.allclose(g[:n], pts[:n], rtol=1e-9, atol=1e-9)
    assert same >= 0.75 * cam.nx * cam.ny
