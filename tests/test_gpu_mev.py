"""GPU: the MeV channel (src/mev_neutrino_physics.cpp, the MeV block of
src/radiative_transfer.cpp:500-640) against the oracle, which is pinned bit
for bit to the reference's own code (tests/test_oracle_mev.py).

- point-wise T, Ye, kappa, j (hadros_eval_mev) for every case and medium;
- whole rays through the replay (the reference's 16-ray cache and a GPU-traced
  64x64 camera through a dense medium);
- the split tracer's MeV columns against the replay of its own cache;
- compute_kerr_image_from_cache with argv 45-71 against the reference CLI's rows.

Tolerances: the device's exp / pow / cbrt / tanh differ from glibc's in the
last ulp, and the wavefront accumulates tau and the affine intensity
recurrence I -> a I + b in a different association than the reference's
sequential loop.  r_neutrinosphere is the radius of the first point whose
optical depth reaches the threshold, so a ray whose depth passes the
threshold within rounding of it can pick the neighbouring point: on the
reference's rays it must be equal, on whole GPU images equal on >= 99 % of rays."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import kgeo
from hadros_amd import abi, ffi, workloads
from oracle import py_oracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, GOLD)
import mev_cases as mc  # noqa: E402

MEV = np.load(os.path.join(GOLD, "mev.npz"))
E, _, S = workloads.synthetic_sigma()


def rel(a, b, floor=1e-300):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def close_columns(got, ref, what):
    """got, ref: [n x 5] MeV columns."""
    assert np.all(rel(got[:, 0], ref[:, 0]) <= 1e-11), (what, "tau")
    assert np.all(np.abs(got[:, 1] - ref[:, 1]) <= 1e-12 + 1e-11 * ref[:, 1]), (what, "P")
    assert np.all(rel(got[:, 2], ref[:, 2]) <= 1e-10), (what, "I")
    assert np.array_equal(got[:, 4], ref[:, 4]), (what, "leakage")
    diff = got[:, 3] != ref[:, 3]
    return diff


def ref_paths():
    d = np.load(os.path.join(GOLD, "kgeo_small.npz"))
    a, r_obs, th_deg, fov, nx, ny, r_max, h = [float(x) for x in d["argv"]]
    cam = abi.camera(a_spin=a, r_obs=r_obs, theta_obs_deg=th_deg, fov_deg=fov, nx=int(nx), ny=int(ny), r_max=r_max,
                     h=h, mode=abi.MODE_REF)
    paths = [po.trace_pixel(cam, int(i), int(j))[0] for (_, i, j, _, _) in d["rays"]]
    off = np.zeros(len(paths) + 1, np.int64)
    off[1:] = np.cumsum([p.shape[0] for p in paths])
    pts = np.concatenate(paths)
    return ffi.Cache(a, off, pts[:, 0], pts[:, 1], pts[:, 5], pts[:, 6]), d, paths, a


@pytest.mark.parametrize("name", list(mc.MEV_CASES))
def test_eval_mev_matches_oracle(name):
    names, over = mc.MEV_CASES[name]
    m = mc.mev_struct(names, over)
    for prof, rho0 in mc.MEV_MEDIA:
        sc = mc.mev_scene(prof, rho0)
        got = ffi.eval_mev(sc, m, MEV["r"], MEV["th"], MEV["g"])
        ref = po.mev_point(sc, m, MEV["r"], MEV["th"], MEV["g"])
        assert np.all(rel(got[:, 0], ref[:, 0]) <= 1e-13), (name, prof, "T")
        assert np.all(rel(got[:, 1], ref[:, 1]) <= 1e-13), (name, prof, "Ye")
        assert np.all(rel(got[:, 2], ref[:, 2]) <= 1e-12), (name, prof, "kappa")
        assert np.all(rel(got[:, 3], ref[:, 3]) <= 1e-11), (name, prof, "j")


@pytest.mark.parametrize("name", list(mc.MEV_CASES))
@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_FP32K])
def test_replay_mev_on_reference_rays(name, precision):
    names, over = mc.MEV_CASES[name]
    m = mc.mev_struct(names, over)
    cache, d, paths, a = ref_paths()
    for si, (prof, rho0, spec) in enumerate((("gaussian", 1e12, "monochromatic"),
                                             ("powerlaw_funnel_envelope", 1e3, "powerlaw_cutoff"))):
        sc = mc.mev_scene(prof, rho0, "funnel_wall", spec)
        rp = ffi.Replay(cache, sc, E, S, precision)
        rp.set_mev(m)
        rp.run()
        got = rp.fetch_mev().stack()
        rp.close()
        ref = MEV[f"{name}_rt{si}"][:, 3:8]   # the reference's own values
        diff = close_columns(got, ref, (name, si))
        assert not diff.any(), (name, si, got[diff, 3], ref[diff, 3])


def dense_scene():
    sc = workloads.c2_scene()
    sc.medium.rho0 = 1e12
    sc.medium.envelope_rho0 = 1e10
    return sc


@pytest.mark.parametrize("mev", [abi.mev(), abi.mev(spectral="fermi_dirac_band", flavor="nu_e", ye="funnel_proton_rich"),
                                 abi.mev(model="toy", thermal="collapsar_inner_hot", enu=25.0)])
def test_replay_mev_on_gpu_camera(mev):
    cam = workloads.c2_camera(nx=64, ny=64)
    sc = dense_scene()
    cache = ffi.trace_cache(cam)
    rp = ffi.Replay(cache, sc, E, S, abi.PREC_FP64)
    rp.set_mev(mev)
    rp.run()
    got = rp.fetch_mev().stack()
    rp.close()
    ref = po.rt_from_cache_mev(cache.view(), sc, E, S, mev)
    diff = close_columns(got, ref, "gpu camera")
    assert diff.mean() <= 0.01, diff.sum()
    assert (ref[:, 3] > 0).mean() > 0.2   # the dense medium puts neutrinospheres on many rays


def test_tracer_mev_equals_replay_of_its_cache():
    cam = workloads.c2_camera(nx=48, ny=48)
    sc = dense_scene()
    m = abi.mev(spectral="fermi_dirac_band")
    t = ffi.Tracer(cam, sc, E, S, abi.PREC_FP64)
    t.set_mev(m)
    t.run()
    img, st = t.fetch()
    got = t.fetch_mev().stack()
    t.close()
    cache = ffi.trace_cache(cam)
    rp = ffi.Replay(cache, sc, E, S, abi.PREC_FP64)
    rp.set_mev(m)
    rp.run()
    want = rp.fetch_mev().stack()
    rp.close()
    diff = close_columns(got, want, "tracer vs replay")
    assert diff.mean() <= 0.01, diff.sum()
    # without MeV the columns keep the reference's defaults
    t = ffi.Tracer(cam, sc, E, S, abi.PREC_FP64)
    t.run()
    none = t.fetch_mev().stack()
    t.close()
    assert np.all(none == np.array([0.0, 1.0, 0.0, -1.0, 1.0]))


def test_mev_needs_the_split_form():
    cam = workloads.c2_camera(nx=16, ny=16)
    t = ffi.Tracer(cam, dense_scene(), E, S, abi.PREC_FP64, path=abi.PATH_FUSED)
    t.set_mev(abi.mev())
    with pytest.raises(ffi.HadrosError):
        t.run()
    t.close()


def test_cli_mev_columns_match_reference_rows(tmp_path):
    d = np.load(os.path.join(GOLD, "kgeo_small.npz"))
    cache, _, paths, a = ref_paths()
    rays = [(int(rid), int(i), int(j), int(cap), al, be, p)
            for (rid, i, j, cap, n), (al, be), p in zip(d["rays"], d["alpha_beta"], paths)]
    nx, ny = int(float(d["argv"][4])), int(float(d["argv"][5]))
    kgeo.write(str(tmp_path / "ref_cache.bin"), nx, ny, a, rays)
    os.makedirs(tmp_path / "output" / "images")
    workloads.write_sigma_file(str(tmp_path / "sigma.dat"), *workloads.synthetic_sigma())
    for name in mc.MEV_CASES:
        argv = ["sigma.dat" if x == "SIGMA" else str(tmp_path / "ref_cache.bin") if x == "CACHE" else x
                for x in MEV[name + "_argv"]]
        r = subprocess.run([os.path.join(ROOT, "hadros_amd", "bin", "compute_kerr_image_from_cache")] + argv,
                           cwd=tmp_path, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, (name, r.stderr)
        saved = re.search(r"Saved: (\S+)", r.stdout).group(1)
        assert saved == str(MEV[name + "_filename"][0]), name
        lines = open(tmp_path / saved).read().splitlines()
        hdr = [ln for ln in lines if ln.startswith("#")]
        for x, y in zip(hdr, MEV[name + "_header"]):
            if not y.startswith("# MEV_CUDA_STATUS"):
                assert x == y, (name, x, y)
        rows = np.array([[float(v) for v in ln.split()] for ln in lines if not ln.startswith("#")])
        ref = MEV[name + "_rows"]
        assert np.array_equal(rows[:, [0, 1, 2, 3, 7, 11, 12]], ref[:, [0, 1, 2, 3, 7, 11, 12]]), name
        for c in (4, 5, 6, 8, 9, 10):
            scale = np.maximum(np.abs(ref[:, c]), 1e-300)
            assert np.all(np.abs(rows[:, c] - ref[:, c]) <= 2e-8 * scale + 1e-300), (name, c)
