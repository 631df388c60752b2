"""CPU: the scan driver's host logic (hadros_amd/scans.py).

- image_metrics follows scripts/run_robustness_scans.py:168-183: means and
  sums over rays whose tau, P and I are all finite; valid rays are the
  finite, uncaptured ones (a ray with any non-finite column drops out of
  every metric, a captured ray only out of the valid count);
- rel_diff follows scripts/compare_cuda_cpu_metrics.py:33-38;
- image_argv / params_from_argv are inverse and put every parameter at the
  position compute_kerr_image_from_cache reads it from;
- the case list and the anomaly rule match the reference script (47 cases;
  "mean_tau_decreased" when an energy step lowers mean tau), checked on a
  stand-in scanner that needs no GPU.
"""
import math

import numpy as np

from hadros_amd import abi, scans


def test_image_metrics_masks():
    tau = np.array([1.0, 2.0, np.nan, 4.0, 8.0])
    p = np.array([0.5, 0.25, 0.1, np.inf, 0.125])
    I = np.array([1.0, 10.0, 100.0, 1000.0, 10000.0])
    cap = np.array([0, 1, 0, 0, 0])
    m = scans.image_metrics(tau, p, I, cap)
    # rays 2 (tau NaN) and 3 (P inf) are not finite; ray 1 is captured
    assert m["mean_tau"] == (1.0 + 2.0 + 8.0) / 3
    assert m["max_tau"] == 8.0
    assert m["mean_P_surv"] == (0.5 + 0.25 + 0.125) / 3
    assert m["total_intensity"] == 1.0 + 10.0 + 10000.0
    assert m["valid_rays"] == 2


def test_rel_diff_rules():
    assert scans.rel_diff(7, 5) == 2
    assert scans.rel_diff(0.0, 0.0) == 0.0
    assert scans.rel_diff(1e-30, 0.0) == np.inf
    assert scans.rel_diff(1.5, 1.0) == 0.5


def test_argv_round_trip_and_positions():
    p = dict(scans.BASE, density_profile="powerlaw_funnel", source_model="shock_layer", enu="1e7", source_q="0.5")
    argv = scans.image_argv(p, "cache.bin")
    assert len(argv) == 38 and argv[28] == "cache.bin"
    assert argv[0] == "1e7" and argv[17] == "powerlaw_funnel" and argv[29] == "shock_layer" and argv[31] == "0.5"
    assert scans.params_from_argv(argv) == p
    sp = dict(p, spectral_model="powerlaw_cutoff", spectral_n_bins="4")
    for k, v in scans.SPECTRAL_DEFAULTS.items():
        sp.setdefault(k, v)
    argv = scans.image_argv(sp, "c")
    assert len(argv) == 44 and argv[38] == "powerlaw_cutoff" and argv[43] == "4"
    assert scans.params_from_argv(argv) == sp


def test_scene_mapping():
    p = dict(scans.BASE, source_theta="25.0", funnel_theta="30.0", source_grad_dtheta="2.0", h="0.3")
    sc = scans.scene_from_params(p)
    assert sc.enu_obs_GeV == 1e5 and sc.medium.h_over_r == 0.3 and sc.medium.rho0 == 1e-2
    assert sc.source.theta_width_rad == 25.0 * math.pi / 180.0
    assert sc.medium.funnel_theta_rad == 30.0 * math.pi / 180.0
    assert sc.source.gradient_dtheta_rad == 2.0 * math.pi / 180.0
    assert sc.source.rho_ref == 1e-2 and sc.spectrum.kind == abi.SPECTRUM_KINDS["monochromatic"]


class _FakeScanner:
    """mean_tau rises with enu except for jet_base at 1e6."""

    def run_case(self, scan, case, spin, theta, **upd):
        p = dict(scans.BASE)
        p.update({k: str(v) for k, v in upd.items()})
        t = math.log10(float(p["enu"]))
        if p["source_model"] == "jet_base" and p["enu"] == "1e6":
            t = 0.0
        return {"scan": scan, "case": case, **p, "mean_tau": t, "max_tau": t, "mean_P_surv": 1.0,
                "total_intensity": 1.0, "valid_rays": 1, "spin_theta": (spin, theta)}


def test_case_list_and_anomaly_flag(tmp_path):
    rows = scans.run_scans(_FakeScanner())
    assert len(rows) == 47
    counts = {}
    for r in rows:
        counts[r["scan"]] = counts.get(r["scan"], 0) + 1
    assert counts == {"source_model": 5, "density_background": 4, "energy": 15, "inclination": 5, "spin": 4,
                      "parameter_stability": 14}
    flagged = [r["case"] for r in rows if r.get("anomaly")]
    assert flagged == ["jet_base_E1e6"]
    assert {r["spin_theta"] for r in rows if r["scan"] == "spin"} == {(a, "80.0") for a in ("0.0", "0.5", "0.9", "0.99")}
    scans.write_csv(rows, tmp_path / "s.csv")
    scans.write_markdown(rows, tmp_path / "s.md")
    head = open(tmp_path / "s.csv").readline().strip().split(",")
    assert head == sorted(head) and "anomaly" in head and "mean_tau" in head
    assert "Energy scan anomalies flagged: 1 cases." in open(tmp_path / "s.md").read()
