"""GPU: the persistent multi-camera tracer.

Every image of a batch must equal the plain per-camera tracer's, run in the
same form (fused, or split into geodesic + replay kernels), bit for bit (same
device arithmetic per ray; only the scheduling differs), for every workgroup
size / steps-per-chunk the autotuner may pick."""
import numpy as np
import pytest

from hadros_amd import abi, ffi, workloads
from oracle import py_oracle as po

pytestmark = pytest.mark.gpu

E, _, S = workloads.synthetic_sigma()


def small_batch(n=14, shrink=8, lowdens=True):
    man = workloads.c3_manifest()[:n]
    cams, scenes = [], []
    for m in man:
        cams.append(workloads.manifest_camera(m, res=max(8, m["res"] // shrink)))
        sc = workloads.manifest_scene(m)
        if lowdens:   # keep survival and intensity non-trivial
            sc.medium.rho0 = 10.0 ** (m["log10_rho0"] - 10.0)
            sc.medium.envelope_rho0 = 1e-2 * sc.medium.rho0
        scenes.append(sc)
    return cams, scenes


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("path", [abi.PATH_FUSED, abi.PATH_SPLIT])
@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_FP32K])
def test_batch_equals_plain_bitwise(precision, path):
    cams, scenes = small_batch()
    b = ffi.Batch(cams, scenes, E, S, precision, path=path)
    for block, spc in ((64, 1), (128, 4), (256, 64)):
        b.set_tuning(block, spc)
        b.run()
        b.last_ms()
        st = b.stats()
        tot = 0
        for c, (cam, sc) in enumerate(zip(cams, scenes)):
            img, pst = ffi.trace_rt(cam, sc, E, S, precision, path=path)
            got = b.fetch(c, pst.n_bins)
            for f in ("tau", "p", "I", "captured", "npoints", "flags", "bin_tau", "bin_p", "bin_i"):
                assert same(getattr(got, f), getattr(img, f)), (block, spc, c, f)
            tot += pst.total_points
        assert st.total_points == tot


def test_batch_autotune_and_oracle():
    cams, scenes = small_batch(n=6, shrink=16)
    b = ffi.Batch(cams, scenes, E, S, abi.PREC_FP64)
    blk, spc = b.autotune()
    assert blk in (64, 128, 256) and spc in (1, 4, 16, 64)
    b.run()
    b.last_ms()
    for c, (cam, sc) in enumerate(zip(cams, scenes)):
        ref, rst = po.trace_image(cam, sc, E, S)
        got = b.fetch(c, max(rst.n_bins, 0))
        assert np.array_equal(got.captured, ref.captured)
        assert np.abs(got.p - ref.p).max() <= 1e-6
        si, sr = got.I.sum(), ref.I.sum()
        assert abs(si - sr) <= 1e-6 * max(abs(sr), 1e-300)


def test_batch_split_sub_batches_and_pool_growth():
    """A pool bound of ~1000 rays' worth of chunks splits the batch into many
    sub-batches, and rays longer than the first sizing overflow it: the
    results after growth must equal the one-sub-batch run bit for bit."""
    cams, scenes = small_batch()
    ref = ffi.Batch(cams, scenes, E, S, abi.PREC_FP64, path=abi.PATH_SPLIT)
    ref.run()
    want = ref.fetch_all(None)
    wst = ref.stats()
    n = ref.info()["n_rays"]
    assert n > 5000
    b = ffi.Batch(cams, scenes, E, S, abi.PREC_FP64, path=abi.PATH_SPLIT)
    b.set_pool(3 * 2048 * 1000)
    for _ in range(2):   # the second run reuses the grown pool
        b.run()
        got = b.fetch_all(None)
        st = b.stats()
        for f in ("tau", "p", "I", "captured", "npoints", "flags", "bin_tau", "bin_p", "bin_i"):
            assert same(getattr(got, f), getattr(want, f)), f
        assert (st.total_points, st.total_attempts) == (wst.total_points, wst.total_attempts)
    with pytest.raises(ffi.HadrosError):
        b.set_pool(100.0)   # below two chunks


@pytest.mark.parametrize("path", [abi.PATH_FUSED, abi.PATH_SPLIT])
def test_batch_tile_order_odd_resolutions(path):
    """Rays are queued 8x8 tile by tile, then the strips right of and below
    the whole tiles: resolutions that are not multiples of 8 (and below 8)
    exercise every part of that order; each image must still equal the plain
    tracer's bit for bit."""
    man = workloads.c3_manifest()[:5]
    cams, scenes = [], []
    for m, (nx, ny) in zip(man, ((13, 21), (5, 7), (16, 3), (24, 19), (9, 40))):
        cam = workloads.manifest_camera(m, res=8)
        cam.nx, cam.ny = nx, ny
        cams.append(cam)
        sc = workloads.manifest_scene(m)
        sc.medium.rho0 = 10.0 ** (m["log10_rho0"] - 10.0)
        sc.medium.envelope_rho0 = 1e-2 * sc.medium.rho0
        scenes.append(sc)
    b = ffi.Batch(cams, scenes, E, S, abi.PREC_FP64, path=path)
    b.set_tuning(64, 4)
    b.run()
    for c, (cam, sc) in enumerate(zip(cams, scenes)):
        img, pst = ffi.trace_rt(cam, sc, E, S, abi.PREC_FP64, path=path)
        got = b.fetch(c, pst.n_bins)
        for f in ("tau", "p", "I", "captured", "npoints", "flags", "bin_tau", "bin_p", "bin_i", "alpha", "beta"):
            assert same(getattr(got, f), getattr(img, f)), (c, f)
