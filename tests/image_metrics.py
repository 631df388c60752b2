"""Per-ray comparison of a traced image with the oracle's, under the error
metrics of the parity tests: rel(tau) where tau_ref > 0; rel(P) where
P_ref > 1e-300, else |dP|; rel(sum I) per image, relative to the image's
I-weighted mean of max(1, tau_ref) (a point's emission is attenuated by
e^{-tau before it}, so a relative tau error e moves I by up to e tau_ref, as
it moves P).  Every ray is bounded:

- rays whose point count equals the oracle's integrate the same points up to
  rounding and meet the tight bars: rel(tau) <= tol_tau, and for survival
  rel(P) <= tol_p max(1, tau_ref) where P_ref > 1e-300 (dP/P = -dtau), else
  |dP| <= tol_p;
- "gate" rays: with the same point count, a path point that lies within
  rounding (the two sides' points differ by ~1e-9 relative) of a hard
  density gate -- gaussian_shape's 4 < r < 18 and |theta - pi/2| < 0.45, the
  [r_min, r_max] window of powerlaw_shape, envelope_rho's r0 <= r <= r_max
  (src/torus_profile.cpp:181-248) -- falls on the other side of it, and tau
  takes the density jump of that one point: these meet rel(tau) <= GATE_TOL
  and the survival bar GATE_TOL, and are at most max_gate of the rays;
- "chaotic" rays, whose adaptive step sequence forked on a rounding
  difference so that the two sides integrate different points, meet
  rel(tau) <= CHAOS_TAU and |dP| <= CHAOS_P and are at most max_chaotic of
  the rays.
"""
import numpy as np


def rel(a, b, floor=1e-300):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


CHAOS_TAU, CHAOS_P = 1e-3, 1e-4
GATE_TOL = 1e-4


def survival_err(p, p_ref):
    """rel(P) where P_ref > 1e-300, |dP| elsewhere."""
    p, p_ref = np.asarray(p, float), np.asarray(p_ref, float)
    big = p_ref > 1e-300
    return np.where(big, np.abs(p - p_ref) / np.where(big, p_ref, 1.0), np.abs(p - p_ref))


def compare_images(img, ref, tol_tau=1e-6, tol_p=1e-6, tol_i=1e-6, max_chaotic=0.01, max_gate=1e-3, label="",
                   check=True):
    """Bounds every ray of a traced image against the oracle's.  Returns the
    report it prints: max, p99 and fraction over the bar of rel(tau) and of
    the survival error, the gate and chaotic ray fractions, rel(sum I), and
    "violations" (empty when every bar holds).  With check=True a violation
    raises AssertionError."""
    bad = []
    if not np.array_equal(img.captured, ref.captured):
        bad.append("captured flags differ")
    matched = img.npoints == ref.npoints
    chaotic = ~matched
    has_tau = ref.tau > 0
    rt = np.where(has_tau, rel(img.tau, ref.tau), 0.0)
    # survival bar relative to max(1, tau_ref): dP/P = -dtau, so a relative tau
    # error e is a relative P error e tau (the plain rel(P) is reported too)
    ep_plain = survival_err(img.p, ref.p)
    ep = np.where(ref.p > 1e-300, ep_plain / np.maximum(1.0, ref.tau), ep_plain)
    dp = np.abs(img.p - ref.p)
    si, sr = img.I.sum(), ref.I.sum()
    ri = abs(si - sr) / abs(sr) if sr != 0 else abs(si)
    # the sum-I bar's scale: I-weighted mean of max(1, tau_ref)
    wi = np.abs(ref.I)
    i_scale = float((wi * np.maximum(1.0, ref.tau)).sum() / wi.sum()) if wi.sum() > 0 else 1.0
    gate = matched & ((rt > tol_tau) | (ep > tol_p))
    rep = dict(rays=int(rt.size), chaotic=float(chaotic.mean()), gate=float(gate.mean()),
               max_rel_tau=float(rt.max()), p99_rel_tau=float(np.percentile(rt, 99)),
               frac_tau_over=float(np.mean(rt > tol_tau)), max_p_err=float(ep.max()),
               p99_p_err=float(np.percentile(ep, 99)), frac_p_over=float(np.mean(ep > tol_p)),
               max_rel_P=float(ep_plain.max()), rel_sum_I=float(ri), sum_I_scale=i_scale)
    if rep["chaotic"] > max_chaotic:
        bad.append(f"chaotic fraction {rep['chaotic']} > {max_chaotic}")
    if rep["gate"] > max_gate:
        bad.append(f"gate fraction {rep['gate']} > {max_gate}")
    if gate.any() and (rt[gate].max() > GATE_TOL or ep[gate].max() > GATE_TOL):
        bad.append(f"gate rays beyond {GATE_TOL}: rel tau {rt[gate].max()}, P {ep[gate].max()}")
    if chaotic.any() and (rt[chaotic].max() > CHAOS_TAU or dp[chaotic].max() > CHAOS_P):
        bad.append(f"chaotic rays beyond the bars: rel tau {rt[chaotic].max()}, |dP| {dp[chaotic].max()}")
    if ri > tol_i * i_scale:
        bad.append(f"rel sum I {ri} > {tol_i} x {i_scale}")
    nb = ref.bin_tau.shape[1] if ref.bin_tau.ndim == 2 else 0
    tight = matched & ~gate
    if nb and tight.any():
        # per bin: rel(tau) as above; survival relative to max(1, tau_ref)
        bt = np.where(ref.bin_tau > 0, rel(img.bin_tau, ref.bin_tau), 0.0)[tight]
        bp = (survival_err(img.bin_p, ref.bin_p) / np.maximum(1.0, ref.bin_tau))[tight]
        rep["bin_max_rel_tau"], rep["bin_max_p_err"] = float(bt.max()), float(bp.max())
        if bt.max() > tol_tau or bp.max() > tol_p:
            bad.append(f"per-bin errors beyond the bars: rel tau {bt.max()}, P {bp.max()}")
    rep["violations"] = bad
    print(label, rep)
    if check:
        assert not bad, (label, rep)
    return rep
