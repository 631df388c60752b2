"""GPU: the scan driver (hadros_amd/scans.py) through the C-ABI.

- On a cache bit-identical to the reference's 4x4 KGEO cache, each of the
  reference's image argv sets (tests/golden/image_rows.npz) run as a scan
  case gives the metrics that the reference's scan would take from its own
  .dat rows (scripts/run_robustness_scans.py:168-183), within the rows' %.8e
  print precision.
- The full scan (47 cases over 10 GPU-traced 16x16 caches, the reference's
  GBW table) matches the oracle's replay of the same caches: every metric to
  rel 1e-9 (valid ray counts equal), and the same anomaly flags.
"""
import os

import numpy as np
import pytest

from hadros_amd import abi, ffi, scans
from oracle import py_oracle as po

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _gbw():
    d = np.load(os.path.join(GOLD, "sigma_GBW.npz"))
    return d["e"], d["s"]


def _close(a, b, tol):
    if isinstance(b, (int, np.integer)):
        return a == b
    return abs(a - b) <= tol * max(abs(b), 1e-300)


def test_cases_match_reference_image_rows():
    from hadros_amd import workloads

    d = np.load(os.path.join(GOLD, "kgeo_small.npz"))
    im = np.load(os.path.join(GOLD, "image_rows.npz"))
    a, r_obs, th, fov, nx, ny, r_max, h = [float(x) for x in d["argv"]]
    cam = abi.camera(a_spin=a, r_obs=r_obs, theta_obs_deg=th, fov_deg=fov, nx=int(nx), ny=int(ny), r_max=r_max, h=h,
                     mode=abi.MODE_REF)
    rays = [po.trace_pixel(cam, int(i), int(j))[0] for (rid, i, j, cap, n) in d["rays"]]
    off = np.zeros(len(rays) + 1, np.int64)
    off[1:] = np.cumsum([p.shape[0] for p in rays])
    allp = np.concatenate(rays)
    cache = ffi.Cache(a, off, allp[:, 0], allp[:, 1], allp[:, 5], allp[:, 6],
                      captured=np.array([int(c) for (_, _, _, c, _) in d["rays"]], np.int32))
    e, _, s = workloads.synthetic_sigma()     # the fixture images used the synthetic table
    sc = scans.Scanner(e, s)
    names = sorted({k[:-5] for k in im.files if k.endswith("_argv")})
    assert len(names) == 5
    for name in names:
        params = scans.params_from_argv(list(im[name + "_argv"]))
        got = sc.metrics(cache, params)
        rows = im[name + "_rows"]
        want = scans.image_metrics(rows[:, 4], rows[:, 5], rows[:, 6], rows[:, 7])
        for k, v in want.items():
            assert _close(got[k], v, 3e-8), (name, k, got[k], v)


def test_full_scan_matches_oracle(tmp_path):
    e, s = _gbw()
    sc = scans.Scanner(e, s)
    rows = scans.run_scans(sc)
    assert len(rows) == 47 and len(sc._caches) == 10
    for r in rows:
        cache = sc.cache(r["aspin"] if r["scan"] == "spin" else scans.BASE["aspin"],
                         r["cam_theta"] if r["scan"] == "inclination" else scans.BASE["cam_theta"])
        ref = po.rt_from_cache(cache.view(), scans.scene_from_params(r), e, s)
        want = scans.image_metrics(ref.tau, ref.p, ref.I, cache.captured)
        for k, v in want.items():
            assert _close(r[k], v, 1e-9), (r["scan"], r["case"], k, r[k], v)
    flags = [r["case"] for r in rows if r.get("anomaly")]
    print("anomalies:", flags)
    mt = {r["case"]: r["mean_tau"] for r in rows if r["scan"] == "energy"}
    for model in ["inner_ring", "jet_base", "funnel_wall", "density_weighted", "shock_layer"]:
        for lo, hi in (("1e4", "1e5"), ("1e5", "1e6")):
            assert (f"{model}_E{hi}" in flags) == (mt[f"{model}_E{hi}"] < mt[f"{model}_E{lo}"])
    scans.write_csv(rows, tmp_path / "scan_summary.csv")
    scans.write_markdown(rows, tmp_path / "scan_summary.md")
    assert len(open(tmp_path / "scan_summary.csv").read().splitlines()) == 48
