"""The learned surrogate of the per-pixel path and the scan that serves it.

Features (14 per ray): alpha, beta (tan-plane pixel coordinates), spin,
cos(inclination), profile one-hot (8), log10 rho0, radial slope.
Targets (16 per ray): log10 tau and survival P for each of the 8 energy bins
of the powerlaw_cutoff spectrum (1e5-1e12 GeV).

Training rays come from the batched GPU tracer (FP64 geodesics, CFG
integrator) on cameras of the 512x512 workload's medium/source/spectrum
conventions with sampled spin, inclination, profile, rho0 and slope; the
networks are trained by the library's own BF16-MFMA forward/backward kernels.
Two parameter families are served:

- "opaque" (the benchmark's configuration): all 8 profiles, rho0 1e8-1e12
  g/cm^3, where the medium blocks almost every ray;
- "translucent": the powerlaw_funnel_envelope profile at rho0 1-1e3 g/cm^3,
  where survival is spread over (0, 1).

The scan renders parameter points with an ensemble of independently seeded
networks, traces a fixed 1 % pixel sub-lattice of every image exactly, and
falls back to the full plain tracer for every image that fails its rule:

- rule "pixel": any spot pixel with |P_served - P_traced| > tol, or any pixel
  of the check lattice (a quarter of the pixels) where another network of
  the ensemble differs from the served one by more than tol;
- rule "metric" (the per-image survival a robustness scan reports, the mean
  over pixels of the bin-weighted survival of src/radiative_transfer.cpp:
  763-769): the network mean corrected by the spot pixels' mean residual (a
  control variate), accepted when three standard errors of that correction
  and the ensemble's disagreement on the check-lattice mean are both within
  tol times the estimate.

The exact tracer is thus both the fallback and the judge of the networks'
error.
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
RHO_RANGE = (8.0, 12.0)      # log10 rho0 [g/cm^3], the opaque family
TRANSLUCENT_RHO_RANGE = (0.0, 3.0)
TRANSLUCENT_PROFILE = "powerlaw_funnel_envelope"
INC_RANGE = (5.0, 85.0)      # inclination [deg]
SPIN_RANGE = (0.0, 0.998)
HALF_WIDTH = 20.0            # camera half-width [r_g] at r_obs = 1000
FOV_JITTER = (0.9, 1.1)      # training cameras: half-width scale, so pixel centres cover the plane densely


def sample_params(rng, n, family="opaque"):
    """Camera and medium parameters of n training cameras of a family."""
    if family == "translucent":
        prof, rho = np.full(n, abi.PROFILE_KINDS[TRANSLUCENT_PROFILE]), TRANSLUCENT_RHO_RANGE
    else:
        prof, rho = None, RHO_RANGE
    out = {
        "spin": rng.uniform(*SPIN_RANGE, n),
        "inc": rng.uniform(*INC_RANGE, n),
        "profile": rng.integers(0, 8, n) if prof is None else prof,
        "log_rho0": rng.uniform(*rho, n),
        "slope": rng.uniform(*SLOPE_RANGE, n),
        "fov_scale": rng.uniform(*FOV_JITTER, n),
    }
    return out


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


def generate(n_cams, res, seed, precision=abi.PREC_FP64, family="opaque"):
    """Training / test rays: (X [N x 14], Y [N x 16], trace ms)."""
    rng = np.random.default_rng(seed)
    params = sample_params(rng, n_cams, family)
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


def sobol_points(n=4096, profile="powerlaw_funnel_envelope", family="opaque"):
    """First n points of the unscrambled 4-D Sobol sequence mapped to spin,
    inclination, log10 rho0 and slope, all with one density profile."""
    import warnings

    from scipy.stats import qmc

    with warnings.catch_warnings():   # prefixes of a power-of-two length are balanced; others are still the sequence
        warnings.simplefilter("ignore", UserWarning)
        u = qmc.Sobol(d=4, scramble=False).random(n)
    rho = TRANSLUCENT_RHO_RANGE if family == "translucent" else RHO_RANGE
    return {
        "spin": SPIN_RANGE[0] + (SPIN_RANGE[1] - SPIN_RANGE[0]) * u[:, 0],
        "inc": INC_RANGE[0] + (INC_RANGE[1] - INC_RANGE[0]) * u[:, 1],
        "profile": np.full(n, abi.PROFILE_KINDS[profile]),
        "log_rho0": rho[0] + (rho[1] - rho[0]) * u[:, 2],
        "slope": SLOPE_RANGE[0] + (SLOPE_RANGE[1] - SLOPE_RANGE[0]) * u[:, 3],
    }


def bin_weights():
    """Normalised weights w_b / sum w of the 8 powerlaw_cutoff bins of the
    scenes (integrate_kerr_ray_spectral's bins, src/radiative_transfer.cpp:
    698-761, on the synthetic table's range): the aggregate survival of a
    pixel is sum_b w_b P_b (:763-769)."""
    e, _, _ = workloads.synthetic_sigma()
    sc = workloads.c2_scene().spectrum
    lo, hi = max(sc.e_min_GeV, e[0]), min(sc.e_max_GeV, e[-1])
    ll, lh = math.log(lo), math.log(hi)
    d = (lh - ll) / N_BINS
    w = []
    for i in range(N_BINS):
        a, b = ll + d * i, ll + d * i + d
        em = math.exp(0.5 * (a + b))
        w.append(em ** (-sc.gamma) * math.exp(-em / sc.ecut_GeV) * (math.exp(b) - math.exp(a)))
    w = np.array(w)
    return w / w.sum()


def train_ensemble(X, Y, k=2, epochs=8, seed=0, **kw):
    """k networks on the same rays, each with its own initialisation and
    minibatch order (seeds seed, seed + 1, ...)."""
    nets = []
    for i in range(k):
        sur = Surrogate(seed=seed + i, **kw)
        sur.fit(X, Y, epochs=epochs, seed=seed + i)
        nets.append(sur)
    return nets


def subset(params, idx):
    return {k: v[idx] for k, v in params.items()}


def _aggregate(p, wb):
    """Bin-weighted survival sum_b w_b P_b over the last axis (FP64; an
    elementwise product and sum, not a BLAS matrix-vector product, whose
    launch grid is bounded by the row count)."""
    return (p.double() * wb).sum(dim=-1)


def _models(sur):
    return list(sur) if isinstance(sur, (list, tuple)) else [sur]


def scan(sur, params, res=256, spot_stride=10, spot_offset=5, tol=1e-3, group=256, tuning=None,
         fallback_chunk=2048, rule="pixel", check_stride=2):
    """Render every parameter point at res x res x 8 bins with the network
    ensemble `sur` (one Surrogate or a list; the served image is the first
    member's, the others are evaluated on the check lattice of every
    check_stride-th row and column, a quarter of the pixels, and compared
    with it there), trace the sub-lattice i, j = offset + stride k of each
    image (26 x 26 = 676 pixels = 1.03 % at 256 x 256) exactly, and re-trace
    every image that fails `rule` (module docstring) in full.  Everything
    stays on the GPU.  Returns (P images on the GPU [n, res*res, 8] float32, report); the
    report's "metric" holds each image's served mean survival (exact for
    fallback images) and, under rule "metric", its standard error."""
    import torch

    if rule not in ("pixel", "metric"):
        raise ValueError(f"unknown scan rule {rule}")
    models = _models(sur)
    n = len(params["spin"])
    npix = res * res
    wb = torch.from_numpy(bin_weights()).to("cuda", torch.float64)
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
    # 2. every network over every pixel of every image; 3. the rule, per image
    sp_idx = np.arange(spot_offset, res, spot_stride)
    pix = torch.from_numpy((sp_idx[:, None] * res + sp_idx[None, :]).ravel()).cuda()
    P_img = torch.empty((n, npix, N_BINS), dtype=torch.float32, device="cuda")
    al, be = pixel_coords(res)
    base = torch.from_numpy(np.stack([al, be], 1).astype(np.float32)).cuda()
    prm = torch.from_numpy(features(params, 1)[:, 2:].astype(np.float32)).cuda()   # per-image features
    X = torch.empty((group, npix, N_FEAT), dtype=torch.float32, device="cuda")
    # further members run on the check lattice (every check_stride-th row and
    # column), where each is compared with the first member
    ck = np.arange(0, res, check_stride)
    cpix = torch.from_numpy((ck[:, None] * res + ck[None, :]).ravel()).cuda()
    ncheck = int(cpix.numel())
    Xc = torch.empty((group, ncheck, N_FEAT), dtype=torch.float32, device="cuda") if len(models) > 1 else None
    P_k = torch.empty((group, ncheck, N_BINS), dtype=torch.float32, device="cuda") if len(models) > 1 else None
    Y = torch.empty((group * npix, N_OUT), dtype=torch.float32, device="cuda") \
        if any(m.head != "tau" for m in models) else None
    fail = torch.zeros(n, dtype=torch.bool, device="cuda")
    spot_err = torch.zeros(n, dtype=torch.float64, device="cuda")
    disagree = torch.zeros(n, dtype=torch.float64, device="cuda")
    metric = torch.zeros(n, dtype=torch.float64, device="cuda")
    metric_se = torch.zeros(n, dtype=torch.float64, device="cuda")

    def run_model(m, Xv, dst, rows):
        if m.head == "tau":   # fused kernel writes P = exp(-10^log10tau) straight into the images
            m.net.survival(Xv, dst.view(rows, N_BINS))
        else:
            Yp = m.net.forward(Xv, Y[:rows])
            dst.copy_(m.survival_of(Yp).view(dst.shape))

    for g0 in range(0, n, group):
        k = min(n, g0 + group) - g0
        X[:k, :, :2] = base[None]
        X[:k, :, 2:] = prm[g0:g0 + k, None, :]
        out = P_img[g0:g0 + k]
        run_model(models[0], X[:k].view(k * npix, N_FEAT), out, k * npix)
        if len(models) > 1:
            Xc[:k] = X[:k, cpix, :]
            first = out[:, cpix, :]
            m0 = _aggregate(first, wb).mean(dim=1)
            for m in models[1:]:
                dst = P_k[:k]
                run_model(m, Xc[:k].view(k * ncheck, N_FEAT), dst, k * ncheck)
                if rule == "pixel":
                    d = (dst - first).abs().amax(dim=(1, 2)).double()
                else:
                    d = (_aggregate(dst, wb).mean(dim=1) - m0).abs()
                disagree[g0:g0 + k] = torch.maximum(disagree[g0:g0 + k], d)
        sp_t = spot[g0:g0 + k]
        sp_n = out[:, pix, :].double()
        spot_err[g0:g0 + k] = (sp_n - sp_t).abs().amax(dim=(1, 2))
        # the image's mean served survival from per-bin pixel sums (FP64 accumulation)
        mean_n = (out.sum(dim=1, dtype=torch.float64) * wb).sum(dim=-1) / npix
        d = _aggregate(sp_t, wb) - _aggregate(sp_n, wb)    # residuals at the spot pixels
        if rule == "metric":
            metric[g0:g0 + k] = mean_n + d.mean(dim=1)
            metric_se[g0:g0 + k] = d.std(dim=1) / math.sqrt(n_spot)
            fail[g0:g0 + k] = (3.0 * metric_se[g0:g0 + k] > tol * metric[g0:g0 + k]) | \
                              (disagree[g0:g0 + k] > tol * metric[g0:g0 + k])
        else:
            metric[g0:g0 + k] = mean_n
            fail[g0:g0 + k] = (spot_err[g0:g0 + k] > tol) | (disagree[g0:g0 + k] > tol)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    fail_idx = torch.nonzero(fail).flatten()
    fail_h = fail_idx.cpu().numpy()
    t3 = time.perf_counter()
    # 4. fallback: full plain tracing of the failing images, in few large
    # batches (each batch allocates and frees a chunk pool of tens of GB)
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
        full = full.view(idx.size, npix, N_BINS)
        sel = fail_idx[c0:c0 + fallback_chunk]
        # whole images along dim 0 (advanced indexing would address every element)
        P_img.index_copy_(0, sel, full.float())
        metric[sel] = _aggregate(full, wb).mean(dim=1)
        metric_se[sel] = 0.0
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    ok = ~fail.cpu().numpy()
    err_h = spot_err.cpu().numpy()
    return P_img, {
        "images": n, "res": res, "bins": N_BINS, "rule": rule, "ensemble": len(models),
        "spot_rays_per_image": int(n_spot), "spot_fraction": n_spot / npix, "tolerance": tol,
        "head": models[0].head, "fallback_images": int(fail_h.size), "fallback_rate": fail_h.size / n,
        "max_spot_err_accepted": float(err_h[ok].max()) if ok.any() else None,
        "median_spot_err": float(np.median(err_h)),
        "max_disagreement_accepted": float(disagree.cpu().numpy()[ok].max()) if ok.any() and len(models) > 1 else None,
        "seconds": {"spot_trace": t1 - t0, "network": t2 - t1, "check": t3 - t2, "fallback": t4 - t3,
                    "total": t4 - t0},
        "kernel_ms": {"spot_trace": spot_ms, "fallback_trace": fb_ms},
        "traced_rays": int(n_spot_total + fail_h.size * npix),
        "images_per_s": n / (t4 - t0),
        "pixels_per_s": n * npix / (t4 - t0),
        "failed": fail_h,
        "metric": metric.cpu().numpy(),
        "metric_se": metric_se.cpu().numpy(),
    }


def traced_scan(params, res=256, chunk=2048, tuning=None, keep=()):
    """All-traced scan: the batched tracer over every pixel of every image, in
    chunks of `chunk` images.  Returns (wall seconds, kernel ms, {index: P
    image [res*res, 8]} for the indices in `keep`, each image's mean
    aggregate survival)."""
    import torch

    t0 = time.perf_counter()
    n = len(params["spin"])
    npix = res * res
    total_ms = 0.0
    kept = {}
    metric = np.zeros(n)
    for c0 in range(0, n, chunk):
        idx = np.arange(c0, min(n, c0 + chunk))
        b = make_batch(subset(params, idx), res, tuning=tuning)
        try:
            b.run()
            total_ms += b.last_ms()
            agg = torch.empty(idx.size * npix, dtype=torch.float64, device="cuda")
            b.fetch_all_device(p_surv=agg)
            metric[idx] = agg.view(idx.size, npix).mean(dim=1).cpu().numpy()
            for q, i in enumerate(idx):
                if int(i) in keep:
                    kept[int(i)] = b.fetch(q, N_BINS).bin_p.astype(np.float32)
        finally:
            b.close()
    return time.perf_counter() - t0, total_ms, kept, metric
