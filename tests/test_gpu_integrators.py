"""GPU: the integrator forms the kernels take, point by point, the legacy
optical depth, and the GPU's CFG rays against the exact Schwarzschild orbit.

Tolerances:
  RK4 and REF Fehlberg attempts (finite-difference right-hand side:
    device/libm ulps amplified ~1e5)                 rel 1e-8 per component
  CFG Fehlberg attempt, all stages by sincos_m       rel 1e-12
  CFG attempt with angle-addition stages (the form   rel 1e-12 where the form
    k_trace_points and the batch take)               applies; it declines (flag 0)
                                                     for stage offsets beyond 1/4
  tau_along_ray vs the reference's own function      rel 1e-12 (wavefront sum order,
                                                     table exponential)
  GPU CFG cache points, a = 0 equatorial rays        |dphi| within the oracle's own
                                                     bound at each rtol
"""
import numpy as np
import pytest

import test_oracle_integrators as TI
from hadros_amd import abi, ffi, workloads
from oracle import py_oracle as po

pytestmark = pytest.mark.gpu


def _states(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.0, 0.998)
    y = np.zeros((n, 8))
    y[:, 1] = rng.uniform(2.5, 300.0, n)
    y[:, 2] = rng.uniform(0.1, np.pi - 0.1, n)
    y[:, 3] = rng.uniform(-3, 3, n)
    y[:, 4] = -1.0
    y[:, 5] = rng.normal(size=n)
    y[:, 6] = 3 * rng.normal(size=n)
    y[:, 7] = 3 * rng.normal(size=n)
    return a, y


def _rel(d, ref):
    scale = np.abs(ref) + 1e-3 * np.abs(ref).max(0)
    return (np.abs(d - ref) / scale).max()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_step_forms_match_oracle(seed):
    L = po.lib()
    a, y = _states(400, seed)
    comp = [1, 2, 3, 5, 6]
    for h in (0.05, 0.4):
        # RK4 (the REF fallback) and a REF attempt
        g4, g5, ok = ffi.eval_step(a, ffi.STEP_RK4, h, y)
        ref = y.copy()
        for k in range(y.shape[0]):
            L.or_step_rk4(a, h, po.dptr(ref[k]))
        assert ok.all() and np.array_equal(g4, g5)
        assert _rel(g5, ref[:, comp]) < 1e-8
        r4, r5 = np.zeros_like(y), np.zeros_like(y)
        for mode, kinds, tol in ((abi.MODE_REF, (ffi.STEP_RKF45_REF,), 1e-8),
                                 (abi.MODE_CFG, (ffi.STEP_RKF45_CFG, ffi.STEP_RKF45_CFG_SHIFT), 1e-12)):
            for k in range(y.shape[0]):
                L.or_rkf45_attempt(a, mode, h, po.dptr(np.ascontiguousarray(y[k])), po.dptr(r4[k]), po.dptr(r5[k]))
            for kind in kinds:
                g4, g5, ok = ffi.eval_step(a, kind, h, y)
                assert ok.mean() > 0.9, (kind, h, ok.mean())
                assert _rel(g4[ok], r4[ok][:, comp]) < tol, (kind, h)
                assert _rel(g5[ok], r5[ok][:, comp]) < tol, (kind, h)
                if kind == ffi.STEP_RKF45_CFG:
                    assert ok.all()


def test_shift_form_declines_large_theta_offsets():
    """The angle-addition form covers stage offsets |d theta| <= 1/4; a step
    whose first stage already moves theta by more (h d theta/d lambda / 4 > 1/4)
    must decline, and one whose whole step moves theta by < 1e-3 must not."""
    a, y = _states(200, 7)
    dy = ffi.eval_rhs(a, abi.MODE_CFG, y)
    big = 1.2 / np.abs(dy[:, 2]) * 4          # a21 = h/4: first offset 1.2
    small = np.minimum(1e-4 / np.abs(dy[:, 2]), 0.1)
    for hs, want in ((big, False), (small, True)):
        for k in range(y.shape[0]):
            _, _, ok = ffi.eval_step(a, ffi.STEP_RKF45_CFG_SHIFT, float(hs[k]), y[k])
            assert bool(ok[0]) == want, (k, hs[k])


def test_tau_along_ray_matches_reference():
    d = np.load(TI.GOLD + "/legacy.npz")
    e, _, s = workloads.synthetic_sigma()
    cache = ffi.Cache(0.9, d["lt_offset"], d["lt_r"], d["lt_th"], d["lt_dl"], d["lt_z"])
    for k in range(4):
        m = TI._medium(d["lt_cases"][k], d[f"lt{k}_medium"])
        enu = float(TI.json.loads(d["lt_cases"][k])[2])
        tau, p = ffi.tau_along_ray(cache, m, e, s, enu)
        want = d[f"lt{k}_out"]
        rt = np.abs(tau - want[:, 0]) / np.maximum(want[:, 0], 1e-300)
        assert rt.max() < 1e-12, (k, rt.max())
        assert np.abs(p - want[:, 1]).max() < 1e-12 * max(1.0, want[:, 0].max())
        # and it is the oracle's restatement too
        ot, op = po.tau_along_ray(cache.view(), m, e, s, enu)
        assert (np.abs(tau - ot) / np.maximum(ot, 1e-300)).max() < 1e-12


@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
@pytest.mark.parametrize("tol,bound", [(1e-8, 3e-6), (1e-10, 5e-8)])
def test_gpu_cfg_points_follow_schwarzschild_orbits(tol, bound):
    """The GPU's CFG points (k_trace_cache: the angle-addition Fehlberg form
    and the CFG step controller) on a = 0 equatorial rays lie on the exact orbit."""
    nx = 21
    cam = abi.camera(a_spin=0.0, r_obs=1000.0, theta_obs_deg=90.0, fov_deg=workloads.fov_for_half_width(20.0, 1000.0),
                     nx=nx, ny=nx, r_max=1100.0, h=1.0, mode=abi.MODE_CFG, tol=tol)
    cache = ffi.trace_cache(cam)
    j = nx // 2
    worst, checked = 0.0, 0
    for i in range(nx):
        y0 = np.zeros(8)
        po.lib().or_initial_state(cam, i, j, po.dptr(y0))
        if y0[7] == 0.0:
            continue
        b = -y0[7] / y0[4]
        sgn = np.copysign(1.0, y0[7])
        k = i * nx + j
        q = slice(cache.offset[k], cache.offset[k + 1])
        r = cache.r[q]
        phi = np.unwrap(np.arctan2(cache.y[q], cache.x[q]))
        kmin = int(np.argmin(r))
        for m in range(0, r.size, max(1, r.size // 30)):
            ex = TI.exact_orbit_phi(b, r[m], m > kmin)
            if ex is not None:
                worst = max(worst, abs(phi[m] - sgn * ex))
                checked += 1
    print(f"rtol {tol:g}: max |dphi| {worst:.2e} over {checked} points")
    assert checked > 300 and worst <= bound
