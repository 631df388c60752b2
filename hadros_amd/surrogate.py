"""The learned surrogate of the per-pixel path and the scan that serves it.

Features (14 per ray): alpha, beta (tan-plane pixel coordinates), spin,
cos(inclination), profile one-hot (8), log10 rho0, radial slope.
Targets (16 per ray): log10 tau and survival P for each of the 8 energy bins
of the powerlaw_cutoff spectrum (1e5-1e12 GeV).

Training rays come from the batched GPU tracer (FP64 geodesics, CFG
integrator) on cameras of the 512x512 workload's medium/source/spectrum
conventions with sampled spin, inclination, profile, rho0 and slope; the
network is trained by the library's own BF16-MFMA forward/backward kernels.

The scan renders parameter points with the network, traces a fixed 1 %
pixel sub-lattice of every image exactly, and falls back to the full plain
tracer for every image whose spot check disagrees by more than the survival
tolerance.  The exact tracer is thus both the fallback and the judge of the
network's error.
"""
import math
import time

import numpy as np

from . import abi, ffi, workloads

N_FEAT = 14
N_BINS = 8
N_OUT = 2 * N_BINS
LOG_TAU_FLOOR = -12.0

SLOPE_RANGE = (1.0, 3.0)     # radial power-law slope of the disk
RHO_RANGE = (8.0, 12.0)      # log10 rho0 [g/cm^3]
INC_RANGE = (5.0, 85.0)      # inclination [deg]
SPIN_RANGE = (0.0, 0.998)
HALF_WIDTH = 20.0            # camera half-width [r_g] at r_obs = 1000
FOV_JITTER = (0.9, 1.1)      # training cameras: half-width scale, so pixel centres cover the plane densely


def sample_params(rng, n):
    return {
        "spin": rng.uniform(*SPIN_RANGE, n),
        "inc": rng.uniform(*INC_RANGE, n),
        "profile": rng.integers(0, 8, n),
        "log_rho0": rng.uniform(*RHO_RANGE, n),
        "slope": rng.uniform(*SLOPE_RANGE, n),
        "fov_scale": rng.uniform(*FOV_JITTER, n),
    }


def fov_deg(scale=1.0):
    return workloads.fov_for_half_width(HALF_WIDTH * float(scale), 1000.0)


def camera_for(spin, inc, res, scale=1.0, mode=abi.MODE_CFG):
    return abi.camera(a_spin=float(spin), r_obs=1000.0, theta_obs_deg=float(inc), fov_deg=fov_deg(scale),
                      nx=res, ny=res, r_max=1100.0, h=1.0, mode=mode)


def scene_for(profile, log_rho0, slope):
    """The 512x512 workload's medium / funnel_wall source / powerlaw_cutoff
    spectrum with the given profile kind, density scale and slope."""
    sc = workloads.c2_scene()
    rho0 = 10.0 ** float(log_rho0)
    sc.medium.kind = int(profile)
    sc.medium.rho0 = rho0
    sc.medium.envelope_rho0 = 1.0e-2 * rho0
    sc.medium.radial_power = float(slope)
    return sc


def _scales(params):
    return params.get("fov_scale", np.ones(len(params["spin"])))


def pixel_coords(res, stride=1, offset=0, scale=1.0):
    """(alpha, beta) tan-plane coordinates of the traced pixels, pixel
    k = p * n_sub + q with i = offset + stride p, j = offset + stride q
    (src/kerr_camera.cpp:112-116, 196-200)."""
    th = math.tan(0.5 * fov_deg(scale) * math.pi / 180.0)
    idx = np.arange(offset, res, stride)
    u = (2.0 * (idx + 0.5) / res - 1.0) * th
    a, b = np.meshgrid(u, u, indexing="ij")
    return a.ravel(), b.ravel()


def features(params, res, stride=1, offset=0):
    """[n_cams * n_pix, 14] float32 features, camera-major."""
    al, be = pixel_coords(res, stride, offset)     # at scale 1
    npix = al.size
    n = len(params["spin"])
    th1 = math.tan(0.5 * fov_deg(1.0) * math.pi / 180.0)
    ths = np.array([math.tan(0.5 * fov_deg(c) * math.pi / 180.0) for c in _scales(params)]) / th1
    X = np.zeros((n, npix, N_FEAT), np.float32)
    X[:, :, 0] = al[None, :] * ths[:, None]
    X[:, :, 1] = be[None, :] * ths[:, None]
    X[:, :, 2] = params["spin"][:, None]
    X[:, :, 3] = np.cos(np.radians(params["inc"]))[:, None]
    X[np.arange(n), :, 4 + params["profile"]] = 1.0
    X[:, :, 12] = params["log_rho0"][:, None]
    X[:, :, 13] = params["slope"][:, None]
    return X.reshape(n * npix, N_FEAT)


def targets(bin_tau, bin_p):
    t = np.empty((bin_tau.shape[0], N_OUT), np.float32)
    t[:, :N_BINS] = np.log10(np.maximum(bin_tau, 10.0 ** LOG_TAU_FLOOR))
    t[:, N_BINS:] = bin_p
    return t


def make_batch(params, res, stride=1, offset=0, precision=abi.PREC_FP64, tuning=None):
    e, _, s = workloads.synthetic_sigma()
    cams = [camera_for(a, i, res, c) for a, i, c in zip(params["spin"], params["inc"], _scales(params))]
    scenes = [scene_for(p, r, sl) for p, r, sl in zip(params["profile"], params["log_rho0"], params["slope"])]
    b = ffi.Batch(cams, scenes, e, s, precision, stride, offset)
    if tuning:
        b.set_tuning(*tuning)
    return b


def trace_params(params, res, stride=1, offset=0, precision=abi.PREC_FP64, tuning=None,
                 fields=("bin_tau", "bin_p")):
    """Batched trace of every parameter point; returns (Image of all rays
    with `fields` filled, kernel ms)."""
    b = make_batch(params, res, stride, offset, precision, tuning)
    try:
        b.run()
        ms = b.last_ms()
        return b.fetch_all(N_BINS, fields), ms
    finally:
        b.close()


def generate(n_cams, res, seed, precision=abi.PREC_FP64):
    """Training / test rays: (X [N x 14], Y [N x 16], trace ms)."""
    rng = np.random.default_rng(seed)
    params = sample_params(rng, n_cams)
    img, ms = trace_params(params, res, precision=precision)
    return features(params, res), targets(img.bin_tau, img.bin_p), ms


class Surrogate:
    """Trained network + standardisation; predictions as (log10 tau, P) per bin."""

    def __init__(self, seed=0, train_batch=8192, infer_batch=1 << 21, head="tau"):
        self.net = ffi.MLP(N_FEAT, 256, 4, N_OUT, train_batch, infer_batch, seed)
        self.batch = train_batch
        self.head = head     # "tau": P = exp(-10^log10tau); "p": the survival outputs, clipped to [0, 1]

    def fit(self, X, Y, epochs=8, lr=1e-3, lr_final=1e-4, seed=0, log=None):
        import torch

        xm, xs = X.mean(0), X.std(0)
        xs = np.where(xs > 0, xs, 1.0)
        ym, ys = Y.mean(0), Y.std(0)
        ys = np.where(ys > 1e-6, ys, 1.0)
        self.net.set_norm(xm, xs, ym, ys)
        Xd = torch.from_numpy(np.ascontiguousarray(X, np.float32)).cuda()
        Yd = torch.from_numpy(np.ascontiguousarray(Y, np.float32)).cuda()
        n = Xd.shape[0]
        steps_per_epoch = n // self.batch
        total = epochs * steps_per_epoch
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        hist = []
        t0 = time.perf_counter()
        step = 0
        for ep in range(epochs):
            perm = torch.randperm(n, device="cuda", generator=g)
            Xp, Yp = Xd[perm], Yd[perm]
            ep_loss = float("nan")
            for k in range(steps_per_epoch):
                # cosine decay from lr to lr_final
                cur = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * step / max(total - 1, 1)))
                sl = slice(k * self.batch, (k + 1) * self.batch)
                # the loss is read back (a host synchronisation) for the epoch's last batch only
                last = k == steps_per_epoch - 1
                loss = self.net.train_step(Xp[sl], Yp[sl], cur, want_loss=last)
                if last:
                    ep_loss = loss
                step += 1
            hist.append(ep_loss)
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4e}")
        torch.cuda.synchronize()
        self.train_seconds = time.perf_counter() - t0
        self.train_steps = step
        return hist

    def predict(self, Xd):
        """Xd: torch CUDA float32 [n x 14] -> torch CUDA float32 [n x 16] (log10 tau, P)."""
        return self.net.forward(Xd)

    @staticmethod
    def survival(Yp):
        """Survival from the log10 tau outputs: P = exp(-10^lt)."""
        import torch

        return torch.exp(-torch.pow(10.0, Yp[:, :N_BINS].double())).float()

    def survival_of(self, Yp):
        if self.head == "p":
            return Yp[:, N_BINS:].clamp(0.0, 1.0)
        return self.survival(Yp)


def evaluate(sur, X, Y):
    """Errors of the network on held-out rays (host arrays)."""
    import torch

    Yp = sur.predict(torch.from_numpy(np.ascontiguousarray(X)).cuda())
    P_tau = Surrogate.survival(Yp).cpu().numpy()
    Yp = Yp.cpu().numpy()
    P_true = Y[:, N_BINS:]
    lt_true = Y[:, :N_BINS]
    dP_head = np.abs(np.clip(Yp[:, N_BINS:], 0.0, 1.0) - P_true)
    dP_tau = np.abs(P_tau - P_true)
    dlt = np.abs(Yp[:, :N_BINS] - lt_true)
    return {
        "rays": int(X.shape[0]),
        "max_abs_dP_from_tau": float(dP_tau.max()), "p99_abs_dP_from_tau": float(np.percentile(dP_tau, 99)),
        "frac_dP_le_1e-3": float(np.mean(dP_tau.max(1) <= 1e-3)),
        "max_abs_dP_head": float(dP_head.max()), "p99_abs_dP_head": float(np.percentile(dP_head, 99)),
        "mean_abs_dlog10tau": float(dlt.mean()), "p99_abs_dlog10tau": float(np.percentile(dlt, 99)),
        "frac_P_nontrivial": float(np.mean((P_true > 1e-6) & (P_true < 1 - 1e-6))),
    }


def sobol_points(n=4096, profile="powerlaw_funnel_envelope"):
    """First n points of the unscrambled 4-D Sobol sequence mapped to spin,
    inclination, log10 rho0 and slope, all with one density profile."""
    from scipy.stats import qmc

    u = qmc.Sobol(d=4, scramble=False).random(n)
    return {
        "spin": SPIN_RANGE[0] + (SPIN_RANGE[1] - SPIN_RANGE[0]) * u[:, 0],
        "inc": INC_RANGE[0] + (INC_RANGE[1] - INC_RANGE[0]) * u[:, 1],
        "profile": np.full(n, abi.PROFILE_KINDS[profile]),
        "log_rho0": RHO_RANGE[0] + (RHO_RANGE[1] - RHO_RANGE[0]) * u[:, 2],
        "slope": SLOPE_RANGE[0] + (SLOPE_RANGE[1] - SLOPE_RANGE[0]) * u[:, 3],
    }


def subset(params, idx):
    return {k: v[idx] for k, v in params.items()}


def scan(sur, params, res=256, spot_stride=10, spot_offset=5, tol=1e-3, group=256, tuning=None,
         fallback_chunk=512):
    """Render every parameter point at res x res x 8 bins with the network,
    spot-check the sub-lattice i, j = offset + stride k of each image (26 x 26
    = 676 pixels = 1.03 % at 256 x 256) against the exact batched tracer, and
    re-trace every image whose spot check has any |dP| > tol.  Everything
    stays on the GPU.  Returns (P images on the GPU [n, res*res, 8] float32,
    report)."""
    import torch

    n = len(params["spin"])
    npix = res * res
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    # 1. exact spot checks: one batched launch over every image's sub-lattice
    b = make_batch(params, res, spot_stride, spot_offset, tuning=tuning)
    try:
        b.run()
        n_spot_total = b.info()["n_rays"]
        spot = torch.empty(n_spot_total * N_BINS, dtype=torch.float64, device="cuda")
        b.fetch_all_device(bin_p=spot)
        spot_ms = b.last_ms()
    finally:
        b.close()
    n_spot = n_spot_total // n
    spot = spot.view(n, n_spot, N_BINS)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # 2. network over every pixel of every image
    P_img = torch.empty((n, npix, N_BINS), dtype=torch.float32, device="cuda")
    al, be = pixel_coords(res)
    base = torch.from_numpy(np.stack([al, be], 1).astype(np.float32)).cuda()
    prm = torch.from_numpy(features(params, 1)[:, 2:].astype(np.float32)).cuda()   # per-image features
    X = torch.empty((group, npix, N_FEAT), dtype=torch.float32, device="cuda")
    Y = torch.empty((group * npix, N_OUT), dtype=torch.float32, device="cuda") if sur.head == "p" else None
    for g0 in range(0, n, group):
        k = min(n, g0 + group) - g0
        X[:k, :, :2] = base[None]
        X[:k, :, 2:] = prm[g0:g0 + k, None, :]
        Xv = X[:k].view(k * npix, N_FEAT)
        if sur.head == "tau":   # fused kernel writes P = exp(-10^log10tau) straight into the images
            sur.net.survival(Xv, P_img[g0:g0 + k].view(k * npix, N_BINS))
        else:
            Yp = sur.net.forward(Xv, Y[:k * npix])
            P_img[g0:g0 + k] = sur.survival_of(Yp).view(k, npix, N_BINS)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    # 3. compare at the spot pixels
    sp_idx = np.arange(spot_offset, res, spot_stride)
    pix = torch.from_numpy((sp_idx[:, None] * res + sp_idx[None, :]).ravel()).cuda()
    err = (P_img[:, pix, :].double() - spot).abs().amax(dim=(1, 2))
    fail = torch.nonzero(err > tol).flatten()
    fail_h = fail.cpu().numpy()
    err_h = err.cpu().numpy()
    t3 = time.perf_counter()
    # 4. fallback: full plain tracing of the failing images
    fb_ms = 0.0
    for c0 in range(0, fail_h.size, fallback_chunk):
        idx = fail_h[c0:c0 + fallback_chunk]
        b = make_batch(subset(params, idx), res, tuning=tuning)
        try:
            b.run()
            full = torch.empty(idx.size * npix * N_BINS, dtype=torch.float64, device="cuda")
            b.fetch_all_device(bin_p=full)
            fb_ms += b.last_ms()
        finally:
            b.close()
        P_img[fail[c0:c0 + fallback_chunk]] = full.view(idx.size, npix, N_BINS).float()
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    ok = err_h <= tol
    return P_img, {
        "images": n, "res": res, "bins": N_BINS, "spot_rays_per_image": int(n_spot),
        "spot_fraction": n_spot / npix, "tolerance": tol, "head": sur.head,
        "fallback_images": int(fail_h.size), "fallback_rate": fail_h.size / n,
        "max_spot_err_accepted": float(err_h[ok].max()) if ok.any() else None,
        "median_spot_err": float(np.median(err_h)),
        "seconds": {"spot_trace": t1 - t0, "network": t2 - t1, "check": t3 - t2, "fallback": t4 - t3,
                    "total": t4 - t0},
        "kernel_ms": {"spot_trace": spot_ms, "fallback_trace": fb_ms},
        "traced_rays": int(n_spot_total + fail_h.size * npix),
        "images_per_s": n / (t4 - t0),
        "pixels_per_s": n * npix / (t4 - t0),
        "failed": fail_h,
    }


def traced_scan(params, res=256, chunk=1024, tuning=None, keep=()):
    """All-traced scan: the batched tracer over every pixel of every image, in
    chunks of `chunk` images.  Returns (wall seconds, kernel ms, {index: P
    image [res*res, 8]} for the indices in `keep`)."""
    t0 = time.perf_counter()
    n = len(params["spin"])
    total_ms = 0.0
    kept = {}
    for c0 in range(0, n, chunk):
        idx = np.arange(c0, min(n, c0 + chunk))
        b = make_batch(subset(params, idx), res, tuning=tuning)
        try:
            b.run()
            total_ms += b.last_ms()
            for q, i in enumerate(idx):
                if int(i) in keep:
                    kept[int(i)] = b.fetch(q, N_BINS).bin_p.astype(np.float32)
        finally:
            b.close()
    return time.perf_counter() - t0, total_ms, kept
