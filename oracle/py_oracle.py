"""TEST INFRASTRUCTURE ONLY: ctypes loader for the CPU FP64 oracle
(oracle/_build/libhadros_oracle.so) and, where it was built, the compiled
reference harness (oracle/_ref/libhadros_ref.so).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module.  The product path (hadros_amd, the CLIs) never does.
"""
import ctypes as C
import os

import numpy as np

from hadros_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libhadros_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_SO = os.path.join(REF_DIR, "libhadros_ref.so")

_dp = C.POINTER(C.c_double)
_lib = None
_ref = None


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run make -C oracle)")
        L = C.CDLL(ORACLE_SO)
        L.or_horizon.restype = C.c_double
        L.or_lapse.restype = C.c_double
        L.or_omega.restype = C.c_double
        L.or_hamiltonian.restype = C.c_double
        L.or_sigma.restype = C.c_double
        L.or_rho.restype = C.c_double
        L.or_spectral_weight.restype = C.c_double
        L.or_source_weight.restype = C.c_double
        L.or_emissivity.restype = C.c_double
        for f in ("or_horizon",):
            getattr(L, f).argtypes = [C.c_double]
        for f in ("or_lapse", "or_omega"):
            getattr(L, f).argtypes = [C.c_double, C.c_double, C.c_double]
        L.or_metric.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.or_inverse_metric.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.or_rhs.argtypes = [C.c_double, C.c_int, _dp, _dp]
        L.or_hamiltonian.argtypes = [C.c_double, _dp]
        L.or_step_ref.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.or_step_rk4.argtypes = [C.c_double, C.c_double, _dp]
        L.or_rkf45_attempt.argtypes = [C.c_double, C.c_int, C.c_double, _dp, _dp, _dp]
        L.or_trace_states.restype = C.c_int64
        L.or_tau_along_ray.argtypes = [C.POINTER(abi.CacheView), C.POINTER(abi.Medium), _dp, _dp, C.c_int, C.c_double,
                                       _dp, _dp]
        L.or_trace_states.argtypes = [C.POINTER(abi.Camera), C.c_int, C.c_int, _dp, C.c_int64]
        L.or_initial_state.argtypes = [C.POINTER(abi.Camera), C.c_int, C.c_int, _dp]
        L.or_trace_pixel.restype = C.c_int64
        L.or_trace_pixel.argtypes = [C.POINTER(abi.Camera), C.c_int, C.c_int, _dp, C.c_int64,
                                     C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64)]
        L.or_sigma.argtypes = [_dp, _dp, C.c_int, C.c_double]
        L.or_rho.argtypes = [C.POINTER(abi.Medium), C.c_double, C.c_double]
        L.or_mev_point.argtypes = [C.POINTER(abi.Scene), C.POINTER(abi.Mev), C.c_double, C.c_double, C.c_double,
                                   _dp]
        L.or_mev_point.restype = None
        L.or_rt_from_cache_mev.argtypes = [C.POINTER(abi.CacheView), C.POINTER(abi.Scene), _dp, _dp, C.c_int,
                                           C.POINTER(abi.Mev), C.c_int, _dp]
        L.or_opacity_surface.argtypes = [C.POINTER(abi.Medium), _dp, _dp, C.c_int, C.c_double, C.c_double,
                                         C.c_double, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.POINTER(C.c_int)]
        L.or_spectral_weight.argtypes = [C.POINTER(abi.Scene), C.c_double]
        L.or_source_weight.argtypes = [C.POINTER(abi.Scene), C.c_double, C.c_double]
        L.or_emissivity.argtypes = [C.POINTER(abi.Scene), C.c_double, C.c_double, C.c_double]
        L.or_bins.argtypes = [C.POINTER(abi.Scene), _dp, C.c_int, _dp, _dp, _dp]
        L.or_integrate_ray.argtypes = [C.POINTER(abi.Scene), _dp, _dp, C.c_int, C.c_double, _dp,
                                       C.c_int64, _dp, _dp, _dp, _dp]
        L.or_trace_image.argtypes = [C.POINTER(abi.Camera), C.POINTER(abi.Scene), _dp, _dp, C.c_int,
                                     C.c_int, C.POINTER(abi.ImageOut), C.POINTER(abi.Stats)]
        L.or_trace_image_range.argtypes = [C.POINTER(abi.Camera), C.POINTER(abi.Scene), _dp, _dp, C.c_int,
                                           C.c_int, C.c_int64, C.c_int64, C.POINTER(abi.ImageOut),
                                           C.POINTER(abi.Stats)]
        L.or_rt_from_cache.argtypes = [C.POINTER(abi.CacheView), C.POINTER(abi.Scene), _dp, _dp, C.c_int,
                                       C.c_int, C.POINTER(abi.ImageOut)]
        _lib = L
    return _lib


def have_ref():
    return os.path.exists(REF_SO)


def ref():
    """The compiled reference harness (only where /root/reference was built)."""
    global _ref
    if _ref is None:
        if not have_ref():
            raise RuntimeError("reference harness not built (make -C oracle ref)")
        R = C.CDLL(REF_SO)
        R.ref_metric.argtypes = [C.c_double, C.c_double, C.c_double, _dp, _dp]
        R.ref_lapse_omega.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        R.ref_rhs.argtypes = [C.c_double, _dp, _dp]
        R.ref_hamiltonian.restype = C.c_double
        R.ref_hamiltonian.argtypes = [C.c_double, _dp]
        R.ref_step_adaptive.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        R.ref_step_rk4.argtypes = [C.c_double, C.c_double, _dp]
        R.ref_tau_along_ray.argtypes = [C.c_char_p, C.c_char_p, _dp, C.c_double, _dp, _dp, _dp, _dp,
                                        C.POINTER(C.c_long), C.c_long, _dp]
        R.ref_trace_pixel.restype = C.c_long
        R.ref_trace_pixel.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_double, C.c_double, C.c_int, C.c_int, _dp, C.c_long,
                                      C.POINTER(C.c_int), _dp]
        R.ref_sigma.argtypes = [C.c_char_p, _dp, _dp, C.c_long]
        R.ref_mev_point.argtypes = [C.c_char_p, _dp, C.POINTER(C.c_char_p), _dp, _dp, _dp, _dp, _dp, C.c_long]
        R.ref_integrate_mev.argtypes = [C.c_char_p, C.c_char_p, _dp, C.c_char_p, _dp, C.c_char_p, _dp, C.c_double,
                                        C.c_double, C.c_double, C.POINTER(C.c_char_p), _dp, _dp, C.c_long, _dp]
        R.ref_rho.argtypes = [C.c_char_p, _dp, _dp, _dp, _dp, C.c_long]
        R.ref_emissivity.argtypes = [C.c_char_p, _dp, C.c_char_p, _dp, C.c_char_p, _dp, _dp, _dp, _dp,
                                     _dp, _dp, _dp, C.c_long]
        R.ref_integrate.argtypes = [C.c_char_p, C.c_char_p, _dp, C.c_char_p, _dp, C.c_char_p, _dp,
                                    C.c_double, C.c_double, C.c_double, _dp, C.c_long, _dp]
        _ref = R
    return _ref


# ------------------------------------------------------------------ helpers
class Image:
    """Host arrays for one camera's outputs (pixel k = i*ny + j)."""

    def __init__(self, npix, nbins=abi.MAX_BINS):
        self.tau = np.zeros(npix)
        self.p = np.zeros(npix)
        self.I = np.zeros(npix)
        self.captured = np.zeros(npix, np.int32)
        self.npoints = np.zeros(npix, np.int32)
        self.flags = np.zeros(npix, np.int32)
        self.alpha = np.zeros(npix)
        self.beta = np.zeros(npix)
        self.bin_tau = np.zeros(npix * nbins)
        self.bin_p = np.zeros(npix * nbins)
        self.bin_i = np.zeros(npix * nbins)

    def out(self):
        o = abi.ImageOut()
        o.tau, o.p_surv, o.intensity = dptr(self.tau), dptr(self.p), dptr(self.I)
        o.captured, o.npoints, o.flags = iptr(self.captured), iptr(self.npoints), iptr(self.flags)
        o.alpha, o.beta = dptr(self.alpha), dptr(self.beta)
        o.bin_tau, o.bin_p, o.bin_i = dptr(self.bin_tau), dptr(self.bin_p), dptr(self.bin_i)
        return o

    def trim_bins(self, nb):
        n = self.tau.size
        nb = max(nb, 0)
        self.bin_tau = self.bin_tau[: n * nb].reshape(n, nb)
        self.bin_p = self.bin_p[: n * nb].reshape(n, nb)
        self.bin_i = self.bin_i[: n * nb].reshape(n, nb)
        return self


def trace_image(cam, scene, sig_e, sig_s, nthreads=0, k0=None, k1=None):
    """Oracle trace + transfer of a camera; returns (Image, Stats)."""
    L = lib()
    npix = cam.nx * cam.ny
    img = Image(npix)
    st = abi.Stats()
    o = img.out()
    sig_e = np.ascontiguousarray(sig_e, np.float64)
    sig_s = np.ascontiguousarray(sig_s, np.float64)
    if k0 is None:
        rc = L.or_trace_image(C.byref(cam), C.byref(scene), dptr(sig_e), dptr(sig_s), sig_e.size, nthreads,
                              C.byref(o), C.byref(st))
    else:
        rc = L.or_trace_image_range(C.byref(cam), C.byref(scene), dptr(sig_e), dptr(sig_s), sig_e.size,
                                    nthreads, k0, k1, C.byref(o), C.byref(st))
    if rc:
        raise RuntimeError(f"or_trace_image failed: {rc}")
    return img.trim_bins(st.n_bins), st


def trace_pixel(cam, i, j, cap=400000):
    L = lib()
    pts = np.zeros((cap, 7))
    cp, ex, at = C.c_int(), C.c_int(), C.c_int64()
    n = L.or_trace_pixel(C.byref(cam), i, j, dptr(pts), cap, C.byref(cp), C.byref(ex), C.byref(at))
    return pts[: min(n, cap)].copy(), cp.value, ex.value, at.value


def trace_states(cam, i, j, cap=400000):
    """[n x 8] state (t, r, theta, phi, p_t, p_r, p_theta, p_phi) at every emitted point of a ray."""
    L = lib()
    st = np.zeros((cap, 8))
    n = L.or_trace_states(C.byref(cam), i, j, dptr(st), cap)
    return st[: min(n, cap)].copy()


def tau_along_ray(view, medium, sig_e, sig_s, enu_inf):
    """Legacy per-ray tau and survival of a cache view (optical_depth::tau_along_ray)."""
    L = lib()
    sig_e = np.ascontiguousarray(sig_e, np.float64)
    sig_s = np.ascontiguousarray(sig_s, np.float64)
    tau = np.zeros(view.n_rays)
    p = np.zeros(view.n_rays)
    rc = L.or_tau_along_ray(C.byref(view), C.byref(medium), dptr(sig_e), dptr(sig_s), sig_e.size, enu_inf,
                            dptr(tau), dptr(p))
    assert rc == 0, rc
    return tau, p


def bins(scene, sig_e):
    L = lib()
    sig_e = np.ascontiguousarray(sig_e, np.float64)
    e = np.zeros(abi.MAX_BINS)
    de = np.zeros(abi.MAX_BINS)
    w = np.zeros(abi.MAX_BINS)
    nb = L.or_bins(C.byref(scene), dptr(sig_e), sig_e.size, dptr(e), dptr(de), dptr(w))
    nb = max(nb, 0)
    return e[:nb], de[:nb], w[:nb]


def rt_from_cache(view, scene, sig_e, sig_s, nthreads=0):
    L = lib()
    n = view.n_rays
    img = Image(n)
    o = img.out()
    sig_e = np.ascontiguousarray(sig_e, np.float64)
    sig_s = np.ascontiguousarray(sig_s, np.float64)
    rc = L.or_rt_from_cache(C.byref(view), C.byref(scene), dptr(sig_e), dptr(sig_s), sig_e.size, nthreads,
                            C.byref(o))
    if rc:
        raise RuntimeError(f"or_rt_from_cache failed: {rc}")
    nb = bins(scene, sig_e)[0].size
    return img.trim_bins(nb)


def rt_from_cache_mev(view, scene, sig_e, sig_s, mev, nthreads=0):
    """Oracle MeV columns per ray of a cache: [n_rays x 5] (tau_mev, P_surv_mev,
    I_obs_mev, r_neutrinosphere_rg, leakage_mev)."""
    L = lib()
    sig_e = np.ascontiguousarray(sig_e, np.float64)
    sig_s = np.ascontiguousarray(sig_s, np.float64)
    out = np.zeros((view.n_rays, 5))
    rc = L.or_rt_from_cache_mev(C.byref(view), C.byref(scene), dptr(sig_e), dptr(sig_s), sig_e.size, C.byref(mev),
                                nthreads, dptr(out))
    if rc:
        raise RuntimeError(f"or_rt_from_cache_mev failed: {rc}")
    return out


def mev_point(scene, mev, r, th, g):
    """Oracle (T_MeV, Ye, kappa, j) per point, [n x 4]."""
    L = lib()
    r, th, g = (np.atleast_1d(np.asarray(x, np.float64)) for x in (r, th, g))
    out = np.zeros((r.size, 4))
    for k in range(r.size):
        L.or_mev_point(C.byref(scene), C.byref(mev), r[k], th[k], g[k], dptr(out[k]))
    return out


def integrate_ray(scene, sig_e, sig_s, a_bh, pts):
    L = lib()
    pts = np.ascontiguousarray(pts, np.float64)
    sig_e = np.ascontiguousarray(sig_e, np.float64)
    sig_s = np.ascontiguousarray(sig_s, np.float64)
    agg = np.zeros(3)
    bt = np.zeros(abi.MAX_BINS)
    bp = np.zeros(abi.MAX_BINS)
    bi = np.zeros(abi.MAX_BINS)
    L.or_integrate_ray(C.byref(scene), dptr(sig_e), dptr(sig_s), sig_e.size, a_bh, dptr(pts), pts.shape[0],
                       dptr(agg), dptr(bt), dptr(bp), dptr(bi))
    return agg, bt, bp, bi


def opacity_surface(medium, sig_e, sig_s, enu, m_bh, tau_surface, nr, ntheta):
    """Oracle opacity surface: dict of theta, r_tau, tau_total, cls, crossing."""
    L = lib()
    sig_e = np.ascontiguousarray(sig_e, np.float64)
    sig_s = np.ascontiguousarray(sig_s, np.float64)
    o = {k: np.zeros(ntheta) for k in ("theta", "r_tau", "tau_total", "cls")}
    cr = np.zeros(ntheta, np.int32)
    rc = L.or_opacity_surface(C.byref(medium), dptr(sig_e), dptr(sig_s), sig_e.size, enu, m_bh, tau_surface, nr,
                              ntheta, dptr(o["theta"]), dptr(o["r_tau"]), dptr(o["tau_total"]), dptr(o["cls"]),
                              cr.ctypes.data_as(C.POINTER(C.c_int)))
    if rc:
        raise RuntimeError(f"or_opacity_surface failed: {rc}")
    o["crossing"] = cr
    return o
