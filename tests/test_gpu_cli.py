"""GPU: the drop-in CLIs against the reference's own outputs.

- compute_kerr_geodesics (GPU, REF mode) vs the reference's KGEO cache
  (tests/golden/kgeo_small.npz): header, ray order and headers, captured
  flags; point counts equal on >= 75% of rays (ulp differences between the
  host libm and the device flip rare step accept decisions); where equal, the
  sampled points agree to 1e-6.
- compute_kerr_image_from_cache (GPU replay) on a cache bit-identical to the
  reference's (written from the oracle's paths, which reproduce the
  reference bit for bit) vs the reference's image files
  (tests/golden/image_rows.npz): same file name, same header lines except the
  MeV status line, UHE columns within the %.8e print precision.
- compute_kerr_image direct mode: KIMG .bin + .dat rows vs the oracle.
"""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import kgeo
from hadros_amd import abi, workloads
from oracle import py_oracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hadros_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")


def run(prog, args, cwd):
    r = subprocess.run([os.path.join(BIN, prog)] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, (r.stdout, r.stderr)
    return r.stdout


@pytest.fixture
def work(tmp_path):
    os.makedirs(tmp_path / "output" / "images")
    os.makedirs(tmp_path / "output" / "rays")
    workloads.write_sigma_file(str(tmp_path / "sigma.dat"), *workloads.synthetic_sigma())
    return tmp_path


def test_stage1_matches_reference_cache(work):
    d = np.load(os.path.join(GOLD, "kgeo_small.npz"))
    argv = list(d["argv"])
    out = run("compute_kerr_geodesics", argv + ["cache.bin"], work)
    assert "Saved: cache.bin" in out and "Camera grid: 4 x 4" in out
    (magic, ver, nx, ny, a), rays = kgeo.read(str(work / "cache.bin"))
    assert [magic, ver, nx, ny] == list(d["header"]) and a == float(argv[0])
    same = 0
    off = d["sample_offsets"]
    for k, (rid, i, j, cap, al, be, pts) in enumerate(rays):
        g = d["rays"][k]
        assert (rid, i, j, cap) == tuple(g[:4])
        assert (al, be) == tuple(d["alpha_beta"][k])
        if pts.shape[0] == g[4]:
            same += 1
            smp = d["samples"][off[k]:off[k + 1]]
            idx = smp[:, 0].astype(int)
            assert np.allclose(pts[idx], smp[:, 1:], rtol=1e-6, atol=1e-8)
    assert same >= 12, same


def test_stage2_matches_reference_image_files(work):
    d = np.load(os.path.join(GOLD, "kgeo_small.npz"))
    im = np.load(os.path.join(GOLD, "image_rows.npz"))
    a, r_obs, th, fov, nx, ny, r_max, h = [float(x) for x in d["argv"]]
    cam = abi.camera(a_spin=a, r_obs=r_obs, theta_obs_deg=th, fov_deg=fov, nx=int(nx), ny=int(ny), r_max=r_max, h=h,
                     mode=abi.MODE_REF)
    rays = []
    for (rid, i, j, cap, n), (al, be) in zip(d["rays"], d["alpha_beta"]):
        pts, cp, ex, at = po.trace_pixel(cam, int(i), int(j))
        assert pts.shape[0] == n
        rays.append((int(rid), int(i), int(j), int(cap), al, be, pts))
    kgeo.write(str(work / "ref_cache.bin"), int(nx), int(ny), a, rays)
    names = sorted({k[:-5] for k in im.files if k.endswith("_argv")})
    assert len(names) == 5
    for name in names:
        argv = ["sigma.dat" if x == "SIGMA" else str(work / "ref_cache.bin") if x == "CACHE" else x
                for x in im[name + "_argv"]]
        out = run("compute_kerr_image_from_cache", argv, work)
        saved = re.search(r"Saved: (\S+)", out).group(1)
        assert saved == str(im[name + "_filename"][0]), name
        assert "Loaded rays: 16" in out and "Processed rays: 16" in out
        lines = open(work / saved).read().splitlines()
        hdr = [ln for ln in lines if ln.startswith("#")]
        ref_hdr = list(im[name + "_header"])
        assert len(hdr) == len(ref_hdr)
        for x, y in zip(hdr, ref_hdr):
            if y.startswith("# MEV_CUDA_STATUS"):
                continue
            assert x == y, (name, x, y)
        rows = np.array([[float(v) for v in ln.split()] for ln in lines if not ln.startswith("#")])
        ref = im[name + "_rows"]
        assert np.array_equal(rows[:, [0, 1, 2, 3, 7]], ref[:, [0, 1, 2, 3, 7]])
        # UHE columns and the MeV columns (the reference's default MeV channel)
        for c in (4, 5, 6, 8, 9, 10):
            scale = np.maximum(np.abs(ref[:, c]), 1e-300)
            assert np.all(np.abs(rows[:, c] - ref[:, c]) <= 2e-8 * scale + 1e-300), (name, c)
        assert np.array_equal(rows[:, [11, 12]], ref[:, [11, 12]]), name


def test_direct_mode_kimg_and_rows(work):
    args = ["1e9", "0.9", "3", "1e1", "10", "5", "0.25", "6", "1.5", "15", "2", "1e12", "1", "10", "1", "70",
            "1000", f"{workloads.fov_for_half_width(20.0, 1000.0):.17g}", "16", "12", "1100", "1.0", "32", "128",
            "20000", "1e-8", "powerlaw_funnel_envelope", "2", "0.9", "20", "1e-1", "2.5", "4", "60", "1e-3",
            "funnel_wall", "20", "1", "2", "0.1", "1", "-1", "0", "100", "sigma.dat", "powerlaw_cutoff", "2",
            "1e12", "1e5", "1e12", "8", "cfg", "fp64"]
    out = run("compute_kerr_image", args, work)
    bins = re.findall(r"Saved: (\S+\.bin)", out)
    assert len(bins) == 1
    raw = open(work / bins[0], "rb").read()
    hdr = struct.unpack_from("<7i4d", raw, 0)
    assert hdr[0] == 0x4B494D47 and hdr[1] == 1 and hdr[3:6] == (16, 12, 13)
    rows = np.frombuffer(raw, "<f8", offset=60).reshape(16 * 12, 13)
    # j-major order
    assert np.array_equal(rows[:, 1], np.repeat(np.arange(12), 16)) and np.array_equal(rows[:, 0],
                                                                                      np.tile(np.arange(16), 12))
    cam = abi.camera(a_spin=0.9, r_obs=1000.0, theta_obs_deg=70.0, fov_deg=float(args[17]), nx=16, ny=12,
                     r_max=1100.0, h=1.0, tol=1e-8, mode=abi.MODE_CFG, max_steps=20000)
    sc = abi.scene(enu=1e9, m_bh=3.0, rho0=1e1, r0=10.0, sigma_r=5.0, h_over_r=0.25, source_r=6.0, source_sigma=1.5,
                   source_theta_deg=15.0, source_powerlaw=2.0, source_emax=1e12, source_norm=1.0,
                   profile="powerlaw_funnel_envelope", radial_power=2.0, funnel_depletion=0.9, funnel_theta_deg=20.0,
                   envelope_rho0=1e-1, envelope_alpha=2.5, r_min=4.0, r_max=60.0, rho_floor=1e-3,
                   source="funnel_wall", source_funnel_theta_deg=20.0, source_q=1.0, source_s=2.0,
                   source_gradient_dr=0.1, source_gradient_dtheta_deg=1.0, source_rho_ref=-1.0,
                   source_cutoff_min=0.0, source_cutoff_max=100.0, spectrum="powerlaw_cutoff", gamma=2.0, ecut=1e12,
                   e_min=1e5, e_max=1e12, n_bins=8)
    e, _, s = workloads.synthetic_sigma()
    ref, _ = po.trace_image(cam, sc, e, s)
    k = (rows[:, 0].astype(int) * 12 + rows[:, 1].astype(int))
    assert np.array_equal(rows[:, 7].astype(int), ref.captured[k])
    assert np.abs(rows[:, 5] - ref.p[k]).max() < 1e-6
    assert abs(rows[:, 6].sum() - ref.I.sum()) <= 1e-6 * abs(ref.I.sum())
