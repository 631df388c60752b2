"""CPU: the drop-in CLIs' argument handling and error exits, forced from the
host (apps/compute_kerr_image_from_cache.cpp:365-406,672-700): missing cache,
bad magic, bad version, negative npoints, truncated points, unknown names,
bad sigma table -- all exit 1 with the reference's message before any GPU use."""
import os
import subprocess

import numpy as np
import pytest

import kgeo
from hadros_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hadros_amd", "bin")


def run(prog, args, cwd):
    return subprocess.run([os.path.join(BIN, prog)] + [str(a) for a in args], cwd=cwd, capture_output=True,
                          text=True, timeout=60)


@pytest.fixture
def work(tmp_path):
    os.makedirs(tmp_path / "output" / "images")
    os.makedirs(tmp_path / "output" / "rays")
    workloads.write_sigma_file(str(tmp_path / "sigma.dat"), *workloads.synthetic_sigma())
    return tmp_path


def image_args(cache, profile="gaussian", sigma="sigma.dat"):
    return ["1e9", "0.9", "3", "1e-2", "10", "5", "0.25", "6", "1.5", "15", "2", "1e12", "1", "10", "1", "80", sigma,
            profile, "2", "0.9", "20", "1e-1", "2.5", "4", "60", "1e-3", "60", "1", cache]


def good_rays():
    pts = np.tile(np.array([[50.0, 1.4, 1, 2, 3, 0.1, 1.01]]), (3, 1))
    return [(0, 0, 0, 0, -0.1, -0.1, pts), (1, 0, 1, 1, -0.1, 0.1, pts[:2])]


@pytest.mark.parametrize("case,msg", [
    ("missing", "Could not open geodesic cache"),
    ("magic", "Invalid binary geodesic cache: bad magic number"),
    ("version", "Unsupported binary geodesic cache version: 2"),
])
def test_stage2_header_errors(work, case, msg):
    cache = str(work / "c.bin")
    if case == "magic":
        kgeo.write(cache, 1, 2, 0.9, good_rays(), magic=0x12345678)
    elif case == "version":
        kgeo.write(cache, 1, 2, 0.9, good_rays(), version=2)
    r = run("compute_kerr_image_from_cache", image_args(cache), work)
    assert r.returncode == 1 and msg in r.stderr, (r.stdout, r.stderr)


def test_stage2_negative_npoints(work):
    cache = str(work / "c.bin")
    with open(cache, "wb") as f:
        f.write(kgeo.HDR.pack(kgeo.MAGIC, 1, 1, 1, 0.9))
        f.write(kgeo.RAY.pack(0, 0, 0, 0, -5, 0.0, 0.0))
    r = run("compute_kerr_image_from_cache", image_args(cache), work)
    assert r.returncode == 1 and "Invalid ray with negative number of points" in r.stderr


def test_stage2_truncated(work):
    cache = str(work / "c.bin")
    kgeo.write(cache, 1, 2, 0.9, good_rays())
    data = open(cache, "rb").read()
    open(cache, "wb").write(data[:-20])
    r = run("compute_kerr_image_from_cache", image_args(cache), work)
    assert r.returncode == 1 and "Unexpected end of file while reading points" in r.stderr


def test_stage2_unknown_names_and_bad_table(work):
    cache = str(work / "c.bin")
    kgeo.write(cache, 1, 2, 0.9, good_rays())
    r = run("compute_kerr_image_from_cache", image_args(cache, profile="donut"), work)
    assert r.returncode == 1 and "Unknown density profile 'donut'" in r.stderr
    (work / "bad.dat").write_text("1e3 1 1\n")
    r = run("compute_kerr_image_from_cache", image_args(cache, sigma="bad.dat"), work)
    assert r.returncode == 1 and "at least two valid data lines" in r.stderr
    r = run("compute_kerr_image_from_cache", image_args(cache) + ["ringworld"], work)
    assert r.returncode == 1 and "Unknown UHE source model 'ringworld'" in r.stderr


@pytest.mark.parametrize("pos,value,msg", [
    (45, "quantum", "Unknown MEV_MODEL 'quantum'. Use toy or physical."),
    (46, "nu_tau", "Unknown MEV_FLAVOR 'nu_tau'. Use nu_e, anti_nu_e, or nu_x."),
    (52, "lukewarm", "Unknown MEV_THERMAL_PROFILE 'lukewarm'."),
    (53, "salty", "Unknown MEV_YE_PROFILE 'salty'."),
    (62, "rainbow", "Unknown MEV_SPECTRAL_MODE 'rainbow'. Use monochromatic or fermi_dirac_band."),
])
def test_stage2_unknown_mev_names(work, pos, value, msg):
    """argv 45-71 name the MeV channel (apps/compute_kerr_image_from_cache.cpp:201-305); the
    reference's parsers reject unknown names with these messages (src/mev_neutrino_physics.cpp:36-233)."""
    cache = str(work / "c.bin")
    kgeo.write(cache, 1, 2, 0.9, good_rays())
    argv = image_args(cache) + ["inner_ring", "20", "1", "2", "0.1", "1", "-1", "0", "100", "monochromatic", "2",
                                "1e12", "1e5", "1e12", "8"]
    argv += ["physical", "anti_nu_e", "1", "1", "1", "1", "1", "inner_hot_torus", "neutron_rich_torus", "6", "0.1",
             "0.2", "0.25", "0.55", "0.45", "0.01", "0.6", "monochromatic"]
    assert len(argv) == 62
    argv[pos - 1] = value
    r = run("compute_kerr_image_from_cache", argv, work)
    assert r.returncode == 1 and msg in r.stderr, (r.stdout, r.stderr)


def test_cli_without_gpu_fails_loudly(work):
    from conftest import gpu_available

    if gpu_available():
        pytest.skip("GPU present")
    r = run("compute_kerr_geodesics", ["0.9", "60", "80", "25", "2", "2", "120", "0.1", "out.bin"], work)
    assert r.returncode == 1 and "no HIP device" in r.stderr


def test_opacity_surface_cli_errors(work):
    """apps/extract_opacity_surface.cpp:95-104: invalid grid and energy outside
    the table exit 1 with the reference's messages before any GPU use; an
    unknown profile exits 1 too."""
    base = ["1e5", "3", "sigma.dat", "gaussian", "1e-2", "10", "5", "0.25", "2", "0", "15", "0", "2.5", "4", "60",
            "1e-99", "1"]
    for grid in (["1", "1200"], ["181", "3"]):
        r = run("extract_opacity_surface", base + grid + ["out.dat"], work)
        assert r.returncode == 1 and "Invalid opacity-surface grid parameters." in r.stderr
    r = run("extract_opacity_surface", base[:4] + base[4:14] + ["3"] + base[15:] + ["181", "100", "out.dat"], work)
    assert r.returncode == 1 and "Invalid opacity-surface grid parameters." in r.stderr   # R_max <= R_min
    r = run("extract_opacity_surface", ["1e13"] + base[1:] + ["181", "100", "out.dat"], work)
    assert r.returncode == 1 and "Energy outside sigma table range." in r.stderr
    r = run("extract_opacity_surface", base[:3] + ["donut"] + base[4:] + ["181", "100", "out.dat"], work)
    assert r.returncode == 1 and "donut" in r.stderr
