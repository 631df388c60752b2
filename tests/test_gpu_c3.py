"""GPU: the batched workload at its real shape.

All 256 cameras of the committed manifest (hadros_amd/manifests/c3_cameras.json:
resolution uniform over 64^2..512^2, spin, inclination, the 8 density
profiles at rho0 1e8..1e12 g/cm^3, the 5 sources x 3 spectra, 8 bins on
1e5..1e12 GeV), about 25.5 M rays, in one autotuned persistent launch of the
default (split) form at the manifest's own densities:

- every camera's image equals the plain per-camera tracer's (hadros_tracer_*,
  same form) bit for bit, and the batch's point and attempt totals are the
  sums of the per-camera runs';
- every 64 x 64 camera of the manifest matches the FP64 CPU oracle under the
  per-ray bars of the parity tests (image_metrics.compare_images).
"""
import json
import os

import numpy as np
import pytest

from hadros_amd import abi, ffi, workloads
from image_metrics import compare_images
from oracle import py_oracle as po

pytestmark = pytest.mark.gpu

E, _, S = workloads.synthetic_sigma()
FIELDS = ("tau", "p", "I", "captured", "npoints", "flags", "alpha", "beta", "bin_tau", "bin_p", "bin_i")
MANIFEST = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hadros_amd", "manifests",
                        "c3_cameras.json")


def _same(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _manifest():
    with open(MANIFEST) as f:
        man = json.load(f)
    man = man["cameras"] if isinstance(man, dict) else man
    assert len(man) == 256
    return man


def test_full_manifest_batch_equals_plain_and_oracle():
    man = _manifest()
    cams = [workloads.manifest_camera(m) for m in man]
    scenes = [workloads.manifest_scene(m) for m in man]
    b = ffi.Batch(cams, scenes, E, S, abi.PREC_FP64)
    try:
        blk, spc = b.autotune()
        b.run()
        ms = b.last_ms()
        st = b.stats()
        n_rays = b.info()["n_rays"]
        assert n_rays == sum(c.nx * c.ny for c in cams)
        print(f"batch: {len(cams)} cameras, {n_rays} rays, block {blk}, steps/chunk {spc}, {ms:.1f} ms", flush=True)
        tot_pts = tot_att = 0
        reports = []
        for c, (cam, sc) in enumerate(zip(cams, scenes)):
            img, pst = ffi.trace_rt(cam, sc, E, S, abi.PREC_FP64)
            got = b.fetch(c, pst.n_bins)
            for f in FIELDS:
                assert _same(getattr(got, f), getattr(img, f)), (c, f)
            tot_pts += pst.total_points
            tot_att += pst.total_attempts
            if cam.nx == 64:
                ref, _ = po.trace_image(cam, sc, E, S)
                reports.append(compare_images(got, ref, check=False,
                                              label=f"camera {c} ({man[c]['profile']}, "
                                                    f"rho0 1e{man[c]['log10_rho0']:.2f})"))
        assert (st.total_points, st.total_attempts) == (tot_pts, tot_att)
        assert len(reports) >= 8
        bad = [(k, r["violations"]) for k, r in enumerate(reports) if r["violations"]]
        assert not bad, bad
        print(f"oracle-checked cameras: {len(reports)}, max rel tau {max(r['max_rel_tau'] for r in reports):.2e}, "
              f"max gate fraction {max(r['gate'] for r in reports):.4f}, "
              f"max chaotic fraction {max(r['chaotic'] for r in reports):.4f}", flush=True)
    finally:
        b.close()
