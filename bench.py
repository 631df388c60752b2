#!/usr/bin/env python3
"""Benchmark of the fused per-pixel tracer on one MI355X.

A step is one pass of the hot path over one camera: trace every pixel's
Kerr null geodesic (FP64 RKF45, rtol 1e-8, <= 20000 steps), integrate the
DIS optical depth and the 8-bin UHE transfer along it, and write the image.
Workload: the 512x512 spin-0.99 / 80 deg / powerlaw_funnel_envelope /
powerlaw_cutoff / funnel_wall camera (hadros_amd.workloads.c2_*).  Inputs are
parameters (no HBM-resident input); outputs stay in HBM until the timed
region ends.

Prints ONE JSON line (rank 0).  `value` is rays/s of the plain FP64 tracer;
beside it: the precision form (FP64 geodesics + FP32-Kahan tau/I), the
batched 256-camera persistent launch, the surrogate-served 4096-point scan
against the same scan traced in full, and the FP64 CPU oracle on the host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rays/s on 1 MI355X, plain vs batched+surrogate, at ≤1e-3 rel err vs FP64"

# Algorithmic work per unit (fixed constants; the operation count of the
# analytic method, not of this implementation's instruction stream):
#   per RKF45 attempt: 6 right-hand sides x F_RHS + F_COMB stage/err flops
#   per emitted path point: F_PT (dl metric, redshift, rho, source) + n_bins x F_BIN
ALGO = {"F_RHS": 69, "F_COMB": 390, "F_PT": 150, "F_BIN": 20,
        # F_PT split between the split form's kernels: dl metric + redshift in
        # the geodesic kernel, density + source factor in the transfer kernel
        "F_PT_GEO": 60, "F_PT_RT": 90}
PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 (vector and matrix), AMD datasheet
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0   # MI355X dense BF16 MFMA
# surrogate forward per row: 2 (14 x 256 + 3 x 256 x 256 + 256 x 16) FLOP
MLP_FLOP_PER_ROW = 2 * (14 * 256 + 3 * 256 * 256 + 256 * 16)


def tracer_flops(attempts, points, nbins):
    return attempts * (6 * ALGO["F_RHS"] + ALGO["F_COMB"]) + points * (ALGO["F_PT"] + nbins * ALGO["F_BIN"])


def geodesic_flops(attempts, points):
    return attempts * (6 * ALGO["F_RHS"] + ALGO["F_COMB"]) + points * ALGO["F_PT_GEO"]


def transfer_flops(points, nbins):
    return points * (ALGO["F_PT_RT"] + nbins * ALGO["F_BIN"])


# HBM bytes a transfer pass must move: 32 B per point read, and per ray its
# point count or range (8 B) and outputs: tau, P, I and the per-bin tau, P, I
def transfer_bytes(points, rays, nbins):
    return 32.0 * points + rays * (8.0 + 8.0 * (3 + 3 * max(nbins, 1)))


# geodesic kernel of the split form: the 32-byte points written, and per ray
# its outputs (alpha, beta, captured, npoints, flags) and stored-point count
def geodesic_bytes(points, rays):
    return 32.0 * points + 32.0 * rays


RIDGE_FLOP_PER_BYTE = PEAK_FP64_TFLOPS * 1e12 / (PEAK_HBM_GBS * 1e9)


def kernel_roofline(name, ms, flop, nbytes):
    """The roofline of one kernel: bound "fp64" (FP64 VALU issue; the MI355X
    vector FP64 peak equals its matrix FP64 peak, 78.6 TFLOP/s) when its
    algorithmic intensity (FLOP per HBM byte) is above the ridge, else "hbm"."""
    sec = ms * 1e-3
    intensity = flop / nbytes
    out = {"kernel": name, "ms": float(ms), "flop": flop, "algorithmic_bytes": nbytes,
           "intensity_flop_per_byte": intensity, "tflops": flop / sec / 1e12, "gbs": nbytes / sec / 1e9}
    if intensity >= RIDGE_FLOP_PER_BYTE:
        out.update(bound="fp64", achieved=out["tflops"], peak=PEAK_FP64_TFLOPS, unit="TFLOP/s")
    else:
        out.update(bound="hbm", achieved=out["gbs"], peak=PEAK_HBM_GBS, unit="GB/s")
    out["frac"] = out["achieved"] / out["peak"]
    return out


def image_errors(img, ref):
    """Per-image error of a GPU image against the FP64 oracle's: rel(tau)
    where tau_ref > 0, |dP|, rel(P) where P_ref > 1e-300, rel(sum I)."""
    import numpy as np

    m = ref.tau > 0
    rt = np.abs(img.tau[m] - ref.tau[m]) / ref.tau[m] if m.any() else np.zeros(1)
    dp = np.abs(img.p - ref.p)
    mp = ref.p > 1e-300
    rp = dp[mp] / ref.p[mp] if mp.any() else np.zeros(1)
    si, sr = float(img.I.sum()), float(ref.I.sum())
    return {"max_rel_tau": float(rt.max()), "p99_rel_tau": float(np.percentile(rt, 99)),
            "frac_rel_tau_over_1e-6": float(np.mean(rt > 1e-6)), "max_abs_dP": float(dp.max()),
            "max_rel_P": float(rp.max()), "rel_sum_I": abs(si - sr) / abs(sr) if sr != 0 else abs(si),
            "mean_P": float(ref.p.mean()), "sum_I": sr,
            "captured_mismatch": int(np.sum(img.captured != ref.captured))}


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform's."""
    import platform

    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(cam_fn, scene_fn, nx, min_seconds=10.0, gpu_images=None, opacity=False):
    """Time the FP64 CPU oracle (same algorithm, OpenMP schedule(dynamic)) on a
    band of image rows through the image centre, growing the band until it
    takes >= min_seconds.  With `gpu_images` ({label: (scene, {precision:
    Image})}) the oracle also traces each full image once and the GPU images'
    errors against it are reported."""
    from oracle import py_oracle
    from hadros_amd import workloads

    cam = cam_fn(nx=nx, ny=nx)
    sc = scene_fn()
    e, _, s = workloads.synthetic_sigma()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    rows = 4
    while True:
        i0 = nx // 2 - rows // 2
        k0, k1 = i0 * nx, (i0 + rows) * nx
        t0 = time.perf_counter()
        py_oracle.trace_image(cam, sc, e, s, nthreads=threads, k0=k0, k1=k1)
        dt = time.perf_counter() - t0
        if dt >= min_seconds or rows >= nx:
            break
        rows = min(nx, max(rows * 2, int(rows * min_seconds / max(dt, 1e-3) * 1.1)))
    reps = 1
    if dt < min_seconds:   # the whole image is shorter than the sample: repeat it
        reps = int(min_seconds / max(dt, 1e-3)) + 1
        t0 = time.perf_counter()
        for _ in range(reps):
            py_oracle.trace_image(cam, sc, e, s, nthreads=threads, k0=k0, k1=k1)
        dt = time.perf_counter() - t0
    out = {"value": reps * (k1 - k0) / dt, "unit": "rays/s", "cores": threads, "kind": "port",
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
           "sample": f"{reps} x {k1 - k0} rays = image rows [{i0},{i0 + rows}) of the {nx}x{nx} camera, "
                     f"FP64 oracle (hadros_oracle, CFG mode), {dt:.1f} s"}
    if opacity:
        m = opacity_medium()
        out["opacity_surface"] = {}
        for nth, nr in OPACITY_GRIDS:
            t0 = time.perf_counter()
            py_oracle.opacity_surface(m, e, s, 1.0e7, 3.0, 1.0, nr, nth)
            dt = time.perf_counter() - t0
            out["opacity_surface"][f"{nth}x{nr}"] = {"seconds": dt, "nodes_per_s": nth * nr / dt, "cores": 1}
    if gpu_images:
        errs = {}
        for label, (scene, imgs) in gpu_images.items():
            ref, _ = py_oracle.trace_image(cam, scene, e, s, nthreads=threads)
            errs[label] = {prec: image_errors(img, ref) for prec, img in imgs.items()}
        out["errors_vs_oracle"] = errs
    return out


def batched_leg(e, s):
    """The 256-camera batch (manifest hadros_amd/manifests/c3_cameras.json): one
    persistent launch, autotuned workgroup size and steps per refill."""
    import numpy as np

    from hadros_amd import abi, ffi, workloads

    man = workloads.c3_manifest()
    cams = [workloads.manifest_camera(m) for m in man]
    scenes = [workloads.manifest_scene(m) for m in man]
    b = ffi.Batch(cams, scenes, e, s, abi.PREC_FP64)
    blk, spc = b.autotune()
    b.run()
    b.last_ms()
    ms = []
    for _ in range(2):
        b.run()
        ms.append(b.last_ms())
    st = b.stats()
    info = b.info()
    b.close()
    best = float(np.min(ms))
    return {"value": info["n_rays"] / (best * 1e-3), "unit": "rays/s", "images_per_s": len(cams) / (best * 1e-3),
            "cameras": len(cams), "rays": info["n_rays"], "kernel_ms": best, "block": blk, "steps_per_chunk": spc,
            "grid_blocks": info["grid"], "points": st.total_points, "attempts": st.total_attempts,
            "algorithmic_tflops": tracer_flops(st.total_attempts, st.total_points, 8) / (best * 1e-3) / 1e12}


def replay_leg(cam, sc, e, s, steps=5):
    """Transfer over a resident geodesic cache (image-from-cache) of the same
    camera: trace the cache (CFG mode), then time the replay kernel.
    Algorithmic HBM bytes: 32 B per point (r, theta, dl, redshift) + per ray
    8 B offset + (3 + 3 n_bins) x 8 B outputs."""
    import numpy as np

    from hadros_amd import abi, ffi

    t0 = time.perf_counter()
    cache = ffi.trace_cache(cam)
    t_cache = time.perf_counter() - t0
    pts, rays = int(cache.offset[-1]), cache.n_rays
    out = {"points": pts, "rays": rays, "cache_seconds_incl_d2h": t_cache}
    for name, prec in (("fp64", abi.PREC_FP64), ("fp32k", abi.PREC_FP32K)):
        rp = ffi.Replay(cache, sc, e, s, prec)
        rp.run()
        rp.last_ms()
        ms = []
        for _ in range(steps):
            rp.run()
            ms.append(rp.last_ms())
        _, st = rp.fetch()
        rp.close()
        kms = float(np.mean(ms))
        alg = transfer_bytes(pts, rays, st.n_bins)
        out[name] = {"kernel_ms": kms, "rays_per_s": rays / (kms * 1e-3), "achieved": alg / (kms * 1e-3) / 1e9,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": alg / (kms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                     "bound": "hbm", "algorithmic_bytes": alg}
    return out


OPACITY_GRIDS = ((181, 1200), (2048, 16384))   # (ntheta, nr): the reference default and a fine grid


def opacity_medium():
    """The 512x512 workload's medium, 1e-8 as dense, on [R_min, R_max] = [3, 50]."""
    from hadros_amd import workloads

    m = workloads.c2_scene().medium
    m.rho0 *= 1e-8
    m.envelope_rho0 *= 1e-8
    return m


def opacity_leg(e, s):
    """Opacity surfaces (apps/extract_opacity_surface.cpp) at E = 1e7 GeV, tau_surface 1:
    host-to-host call time of hadros_opacity_surface (allocation, one launch,
    copies) per grid."""
    from hadros_amd import ffi

    m = opacity_medium()
    out = {}
    for nth, nr in OPACITY_GRIDS:
        ffi.opacity_surface(m, e, s, 1.0e7, 3.0, 1.0, nr, nth)
        t0 = time.perf_counter()
        o = ffi.opacity_surface(m, e, s, 1.0e7, 3.0, 1.0, nr, nth)
        dt = time.perf_counter() - t0
        out[f"{nth}x{nr}"] = {"seconds": dt, "nodes_per_s": nth * nr / dt, "crossings": int(o["crossing"].sum())}
    return out


def _scan_report(S, P, rep, tw, tms, kept, metric_true, tol=1e-3):
    """Served-vs-traced errors of a scan: per pixel on the kept images, and the
    per-image mean survival on every served (not fallen-back) image."""
    import numpy as np

    failed = set(rep.pop("failed").tolist())
    served = np.array([i for i in range(len(metric_true)) if i not in failed], dtype=int)
    errs = [float(np.abs(P[i].cpu().numpy() - ref).max()) for i, ref in kept.items() if i not in failed]
    fr = [float(np.mean(np.abs(P[i].cpu().numpy() - ref) > tol)) for i, ref in kept.items() if i not in failed]
    m, mt = rep.pop("metric"), metric_true
    rep.pop("metric_se", None)
    rel = np.abs(m[served] - mt[served]) / np.maximum(mt[served], 1e-300) if served.size else np.zeros(0)
    return {"speedup_vs_all_traced": tw / rep["seconds"]["total"], "fallback_rate": rep["fallback_rate"],
            "rule": rep["rule"], "ensemble": rep["ensemble"], "scan_seconds": rep["seconds"],
            "images_per_s": rep["images_per_s"], "traced_rays": rep["traced_rays"],
            "spot_rays_per_image": rep["spot_rays_per_image"],
            "served_vs_traced": {"images": len(errs), "max_abs_dP": max(errs) if errs else None,
                                 "frac_pixel_bins_over_1e-3": float(np.mean(fr)) if fr else None},
            "served_mean_survival": {"images": int(served.size),
                                     "max_rel_err": float(rel.max()) if rel.size else None,
                                     "p99_rel_err": float(np.percentile(rel, 99)) if rel.size else None,
                                     "frac_over_1e-3": float(np.mean(rel > tol)) if rel.size else None}}


def surrogate_leg():
    """Train a two-network ensemble on 4096 sampled cameras x 32 x 32 traced
    rays (the opaque family of the benchmark configuration), then serve the
    4096-point Sobol scan (256 x 256 x 8 bins per point, 1 % spot-checked,
    full-trace fallback above |dP| 1e-3 at a spot pixel or between the two
    networks) and trace the same scan in full for comparison."""
    import numpy as np
    import torch

    from hadros_amd import surrogate as S

    t0 = time.perf_counter()
    X, Y, gen_ms = S.generate(4096, 32, seed=1)
    Xt, Yt, _ = S.generate(200, 32, seed=2)
    t_gen = time.perf_counter() - t0
    nets = S.train_ensemble(X, Y, k=2, epochs=20, seed=0)
    sur = nets[0]
    ev = S.evaluate(sur, Xt, Yt)
    # fused network kernel alone: 4M rows, survival form, HIP events on the launch stream
    xb = torch.from_numpy(np.ascontiguousarray(np.tile(Xt, (21, 1))[:1 << 22])).cuda()
    pb = torch.empty((xb.shape[0], S.N_BINS), dtype=torch.float32, device="cuda")
    sur.net.survival(xb, pb)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(5):
        sur.net.survival(xb, pb)
    ev1.record()
    torch.cuda.synchronize()
    mlp_ms = ev0.elapsed_time(ev1) / 5
    mlp_tflops = xb.shape[0] * MLP_FLOP_PER_ROW / (mlp_ms * 1e-3) / 1e12
    del xb, pb
    prm = S.sobol_points(4096)
    S.scan(nets, S.subset(prm, np.arange(64)))     # warm-up of the scan path
    P, rep = S.scan(nets, prm)
    keep = set(int(i) for i in np.random.default_rng(5).choice(4096, 32, replace=False))
    tw, tms, kept, mt = S.traced_scan(prm, keep=keep)
    px = 4096 * 256 * 256
    out = {"value": px / rep["seconds"]["total"], "unit": "pixels/s (256x256x8-bin images)",
           "all_traced": {"value": px / tw, "seconds": tw, "kernel_ms": tms}}
    out.update(_scan_report(S, P, rep, tw, tms, kept, mt))
    del P
    torch.cuda.empty_cache()
    out.update({"training": {"rays": int(X.shape[0]), "trace_ms": gen_ms, "data_seconds": t_gen, "networks": len(nets),
                             "steps": sum(n.train_steps for n in nets),
                             "seconds": sum(n.train_seconds for n in nets)},
                "test_error": ev,
                "mlp_kernel": {"ms_per_4M_rows": mlp_ms, "achieved": mlp_tflops, "peak": PEAK_BF16_TFLOPS,
                               "unit": "TFLOP/s", "frac": mlp_tflops / PEAK_BF16_TFLOPS, "bound": "mfma",
                               "flop_per_row": MLP_FLOP_PER_ROW}})
    return out


def surrogate_translucent_leg():
    """The same scan over a translucent medium (powerlaw_funnel_envelope at
    rho0 1-1e3 g/cm^3, survival spread over (0, 1)): a two-network ensemble
    trained on 8192 cameras x 32 x 32 rays of that family serves the 4096
    Sobol points under rule "pixel" (|dP| 1e-3 at every spot pixel) and rule
    "metric" (each image's mean survival to 1e-3 relative), against one
    all-traced scan."""
    import numpy as np
    import torch

    from hadros_amd import surrogate as S

    t0 = time.perf_counter()
    X, Y, gen_ms = S.generate(8192, 32, seed=11, family="translucent")
    Xt, Yt, _ = S.generate(200, 32, seed=12, family="translucent")
    t_gen = time.perf_counter() - t0
    nets = S.train_ensemble(X, Y, k=2, epochs=40, seed=10)
    ev = [S.evaluate(n, Xt, Yt) for n in nets]
    prm = S.sobol_points(4096, family="translucent")
    S.scan(nets, S.subset(prm, np.arange(64)))     # warm-up of the scan path
    keep = set(int(i) for i in np.random.default_rng(6).choice(4096, 32, replace=False))
    tw, tms, kept, mt = S.traced_scan(prm, keep=keep)
    out = {"family": f"{S.TRANSLUCENT_PROFILE}, log10 rho0 {S.TRANSLUCENT_RHO_RANGE}",
           "mean_survival_range": [float(mt.min()), float(mt.max())], "mean_survival_mean": float(mt.mean()),
           "all_traced": {"seconds": tw, "kernel_ms": tms}}
    for rule in ("pixel", "metric"):
        P, rep = S.scan(nets, prm, rule=rule)
        out[rule] = _scan_report(S, P, rep, tw, tms, kept, mt)
        del P
        torch.cuda.empty_cache()
    out["training"] = {"rays": int(X.shape[0]), "trace_ms": gen_ms, "data_seconds": t_gen, "networks": len(nets),
                       "seconds": sum(n.train_seconds for n in nets)}
    out["test_error"] = ev
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 256-camera leg")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate training + scan leg")
    ap.add_argument("--no-replay", action="store_true", help="skip the cache-replay leg")
    ap.add_argument("--no-opacity", action="store_true", help="skip the opacity-surface leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--pmc-target", nargs="?", const="tracer", choices=["tracer", "replay", "batch"],
                    help="for rocprofv3 passes: run the plain tracer as one band (or the cache replay, or the 256-camera batch "
                         "at block 128 and 64 steps per refill, without the autotuner) twice, then fill and read a "
                         "1 GiB device buffer twice each with the library's streaming kernels (calibration of "
                         "WRITE_SIZE / FETCH_SIZE), print nothing else")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the per-ray work has no "
              "cross-GPU exchange, so N GPUs would each trace their own cameras.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from hadros_amd import abi, ffi, workloads

    torch.cuda.set_device(0)
    dev_name, cus, _ = ffi.device_info()
    e, _, s = workloads.synthetic_sigma()
    cam = workloads.c2_camera(nx=args.size, ny=args.size)
    sc = workloads.c2_scene()
    npix = cam.nx * cam.ny
    if args.pmc_target:
        if args.pmc_target == "tracer":
            # one geodesic launch and one transfer launch per run (the kernels' own counters)
            t = ffi.Tracer(cam, sc, e, s, abi.PREC_FP64, args.block, segments=1)
        elif args.pmc_target == "replay":
            t = ffi.Replay(ffi.trace_cache(cam), sc, e, s, abi.PREC_FP64)
        else:
            man = workloads.c3_manifest()
            t = ffi.Batch([workloads.manifest_camera(m) for m in man], [workloads.manifest_scene(m) for m in man],
                          e, s, abi.PREC_FP64)
            t.set_tuning(128, 64)
        for _ in range(2):
            t.run()
            t.last_ms()
        t.close()
        # calibration: the library's streaming kernels over 1 GiB (larger than
        # the 256 MB last-level cache, so every byte crosses HBM): hd::k_stream_fill
        # writes 1048576 KiB, hd::k_stream_read reads 1048576 KiB
        x = torch.empty(128 * 1024 * 1024, dtype=torch.float64, device="cuda")
        part = torch.empty(abi.STREAM_GRID, dtype=torch.float64, device="cuda")
        for _ in range(2):
            ffi.stream_probe(1, x, value=1.0)
        for _ in range(2):
            ffi.stream_probe(0, x, part)
        torch.cuda.synchronize()
        assert float(part.sum()) == float(x.numel())
        return

    def timed(precision, scene=sc, segments=0):
        t = ffi.Tracer(cam, scene, e, s, precision, args.block, segments=segments)
        for _ in range(args.warmup):
            t.run()
            t.last_ms()   # each warmup run settled: the chunk pool has grown to the image's points
        torch.cuda.synchronize()
        # timed: the K images queued back to back on the launch stream, one
        # synchronization at each end.  Every run traces the whole image, and
        # the pool holds every run's points as it held the warmup's (a camera's
        # rays store the same points on every run); the last run is settled.
        t0 = time.perf_counter()
        for _ in range(args.steps):
            t.run()
        t.last_ms()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        n_pts = t.fetch()[1].total_points
        # each run's device time and the bands' spans (HIP events), untimed
        kern, parts = [], []
        for _ in range(args.steps):
            t.run()
            kern.append(t.last_ms())
            parts.append(t.kernel_ms())
        img, st = t.fetch()
        assert st.total_points == n_pts, "runs of one camera stored different point counts"
        t.close()
        return wall, kern, np.mean(np.array(parts), axis=0), img, st

    # the headline: the default schedule, whose tile-row bands overlap one
    # band's transfer with the others' geodesic launches
    wall, kern, (geo_span, rt_span), img, st = timed(abi.PREC_FP64)
    ms_step = 1e3 * wall / args.steps
    value = args.steps * npix / wall
    kavg = float(np.mean(kern))
    # each kernel priced on its own launch: the same image as one geodesic
    # launch, then one transfer launch (HIP events between the two on the
    # launch stream); the bands give the same image bit for bit
    _, _, (geo_ms, rt_ms), img1, _ = timed(abi.PREC_FP64, segments=1)
    assert np.array_equal(img1.tau, img.tau) and np.array_equal(img1.I, img.I), "segment counts disagree"
    kernels = {}
    if rt_ms > 0:
        kernels = {
            "geodesic": kernel_roofline("hd::k_trace_points<CFG>", geo_ms,
                                        geodesic_flops(st.total_attempts, st.total_points),
                                        geodesic_bytes(st.total_points, npix)),
            "transfer": kernel_roofline("hd::k_replay<FP64, chunks>", rt_ms,
                                        transfer_flops(st.total_points, st.n_bins),
                                        transfer_bytes(st.total_points, npix, st.n_bins)),
        }
        dom = max(kernels.values(), key=lambda k: k["ms"])
    else:
        flops = tracer_flops(st.total_attempts, st.total_points, st.n_bins)
        dom = kernel_roofline("hd::k_trace_rt<CFG, FP64, 8>", kavg, flops,
                              geodesic_bytes(st.total_points, npix))
    dom_name = dom["kernel"]
    # the whole plain path (both kernels, or the fused one) under the same FLOP count
    whole_flop = tracer_flops(st.total_attempts, st.total_points, st.n_bins)

    wall32, kern32, _, img32, st32 = timed(abi.PREC_FP32K)
    # the same camera through a medium 1e-8 as dense (rho0 100 g/cm^3): the
    # 512x512 workload's medium is opaque at every pixel, this one is not
    thin = workloads.c2_scene()
    thin.medium.rho0 *= 1e-8
    thin.medium.envelope_rho0 *= 1e-8
    thin_imgs = {}
    wall_t, kern_t, (geo_ts, rt_ts), thin_imgs["fp64"], st_t = timed(abi.PREC_FP64, thin)
    _, _, (geo_t, rt_t), _, _ = timed(abi.PREC_FP64, thin, segments=1)
    thin_imgs["fp32k"], _ = ffi.trace_rt(cam, thin, e, s, abi.PREC_FP32K)
    thin_leg = {"value": args.steps * npix / wall_t, "unit": "rays/s", "kernel_ms": float(np.mean(kern_t)),
                "geodesic_ms": float(geo_t), "transfer_ms": float(rt_t),
                "banded_spans_ms": {"geodesic": float(geo_ts), "transfer": float(rt_ts)},
                "mean_P": float(thin_imgs["fp64"].p.mean()), "sum_I": float(thin_imgs["fp64"].I.sum()),
                "points_per_ray": st_t.total_points / npix}
    if rt_t > 0:
        thin_leg["transfer"] = kernel_roofline("hd::k_replay<FP64, chunks>", rt_t,
                                               transfer_flops(st_t.total_points, st_t.n_bins),
                                               transfer_bytes(st_t.total_points, npix, st_t.n_bins))
    m = (img.tau > 0) & (img.tau < 700)
    err_tau = float(np.max(np.abs(img32.tau[m] - img.tau[m]) / img.tau[m])) if m.any() else 0.0
    err_p = float(np.max(np.abs(img32.p - img.p)))
    si = float(img.I.sum())
    err_i = abs(float(img32.I.sum()) - si) / abs(si) if si != 0 else abs(float(img32.I.sum()))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain FP64 tracer, {cam.nx}x{cam.ny} camera, spin 0.99, incl 80 deg, "
                               "powerlaw_funnel_envelope, powerlaw_cutoff 8 bins, funnel_wall, CFG mode rtol 1e-8",
                   "rays_per_step": npix, "device": dev_name, "block": st.block_size},
        # the compute roof of the one-line contract is "mfma"; the kernel's is FP64 VALU issue
        "roofline": {"bound": "mfma" if dom["bound"] == "fp64" else dom["bound"], "achieved": dom["achieved"],
                     "peak": dom["peak"], "unit": dom["unit"],
                     "frac": dom["frac"], "traffic": None, "kernel": dom_name, "kernel_ms": dom["ms"],
                     "kernel_bound": dom["bound"],
                     "peak_kind": "fp64 VALU issue (no MFMA; MI355X vector FP64 peak = matrix FP64 peak, "
                                  "78.6 TFLOP/s)" if dom["bound"] == "fp64" else "hbm3e",
                     "whole_path": {"ms": kavg, "flop": whole_flop, "tflops": whole_flop / (kavg * 1e-3) / 1e12,
                                    "frac_fp64": whole_flop / (kavg * 1e-3) / 1e12 / PEAK_FP64_TFLOPS},
                     "ridge_flop_per_byte": RIDGE_FLOP_PER_BYTE, "tracer_ms": kavg,
                     "algorithmic_flop_per_launch": dom["flop"], "algorithmic_bytes_per_launch": dom["algorithmic_bytes"],
                     "points_per_launch": st.total_points, "attempts_per_launch": st.total_attempts,
                     "kernels": kernels,
                     # the timed schedule: the spans of the tile-row bands' overlapping
                     # launches (start to the last geodesic end; first transfer start to the end)
                     "banded": {"segments": abi.TRACER_MAX_SEGMENTS, "geodesic_span_ms": float(geo_span),
                                "transfer_span_ms": float(rt_span), "step_ms": kavg}},
        "precision_form": {"value": args.steps * npix / wall32, "unit": "rays/s", "dtype": "f64 geodesic + f32 Kahan",
                           "kernel_ms": float(np.mean(kern32)), "max_rel_tau_vs_fp64": err_tau,
                           "max_abs_p_vs_fp64": err_p, "rel_sum_i_vs_fp64": err_i},
        "image": {"captured": st.n_captured, "exhausted": st.n_exhausted, "mean_P": float(img.p.mean()),
                  "sum_I": si, "points_per_ray": st.total_points / npix},
        # the same camera and path through the medium 1e-8 as dense: survival
        # and intensity are non-zero, so no emission term is skipped
        "thin_medium": thin_leg,
    }
    if not args.no_batch:
        out["batched"] = batched_leg(e, s)
    if not args.no_replay:
        out["replay"] = replay_leg(cam, sc, e, s)
    if not args.no_opacity:
        out["opacity_surface"] = opacity_leg(e, s)
    if not args.no_surrogate:
        out["surrogate_scan"] = surrogate_leg()
        out["surrogate_scan_translucent"] = surrogate_translucent_leg()
    tr = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tr):
        # HBM bytes per launch of each kernel, from rocprofv3 FETCH_SIZE / WRITE_SIZE passes
        t = json.load(open(tr))
        if t.get("workload") == out["config"]["workload"]:
            per = t.get("kernels", {})
            if dom_name in per:
                out["roofline"]["traffic"] = per[dom_name]["bytes_per_launch"]
            for k in out["roofline"]["kernels"].values():
                if k["kernel"] in per:
                    k["traffic"] = per[k["kernel"]]["bytes_per_launch"]
            out["roofline"]["traffic_source"] = t["source"]
            # hardware counters of the same launches (rocprofv3 --pmc): FP64 FLOPs
            # executed, VALU issue and active lanes, priced over this run's kernel times
            for k in out["roofline"]["kernels"].values():
                c = t.get("valu", {}).get(k["kernel"])
                if c:
                    k["hw_fp64_flop"] = c["fp64_flop"]
                    k["hw_tflops"] = c["fp64_flop"] / (k["ms"] * 1e-3) / 1e12
                    k["hw_frac_fp64"] = k["hw_tflops"] / PEAK_FP64_TFLOPS
                    k["valu_issue"] = c["valu_issue"]
                    k["active_lanes"] = c["active_lanes"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(workloads.c2_camera, workloads.c2_scene, args.size, args.cpu_seconds,
                                           {"workload": (sc, {"fp64": img, "fp32k": img32}),
                                            "rho0_x1e-8": (thin, thin_imgs)}, opacity=not args.no_opacity)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
