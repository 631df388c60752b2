"""ctypes binding of libhadros.so (the C-ABI of include/hadros.h).

This is the Python face of the drop-in boundary: the reference's Python
drivers reach the hot path only through `subprocess` calls of the CLIs
(scripts/run_robustness_scans.py:60-70,147-165); here the same path is a
library call.  There is no CPU fallback: if the library is missing or no
gfx950 device is present, every compute call raises HadrosError.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhadros.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
_lib = None

# every symbol include/hadros.h declares
EXPORTS = [
    "hadros_last_error", "hadros_abi_version", "hadros_device_info", "hadros_camera_defaults",
    "hadros_scene_defaults", "hadros_parse_profile", "hadros_parse_source", "hadros_parse_spectrum",
    "hadros_profile_name", "hadros_source_name", "hadros_spectrum_name", "hadros_sigma_load",
    "hadros_sigma_check", "hadros_trace_rt", "hadros_trace_rt_noalloc", "hadros_tracer_create", "hadros_tracer_set_block", "hadros_tracer_set_path",
    "hadros_tracer_set_segments",
    "hadros_tracer_run", "hadros_tracer_last_ms", "hadros_tracer_kernel_ms", "hadros_tracer_fetch",
    "hadros_tracer_destroy",
    "hadros_rt_from_cache", "hadros_replay_create", "hadros_replay_run", "hadros_replay_last_ms",
    "hadros_replay_fetch", "hadros_replay_destroy", "hadros_trace_cache", "hadros_trace_cache_range", "hadros_eval_rho",
    "hadros_eval_sigma", "hadros_eval_emissivity", "hadros_eval_rhs", "hadros_eval_metric", "hadros_eval_mev",
    "hadros_eval_step", "hadros_tau_along_ray", "hadros_eval_transfer_point", "hadros_stream_probe",
    "hadros_tracer_set_mev", "hadros_tracer_fetch_mev", "hadros_replay_set_mev", "hadros_replay_fetch_mev",
    "hadros_mev_defaults", "hadros_parse_mev_model", "hadros_parse_mev_flavor", "hadros_parse_mev_thermal",
    "hadros_parse_mev_ye", "hadros_parse_mev_spectral",
    "hadros_batch_create", "hadros_batch_create_sub", "hadros_batch_fetch_all", "hadros_batch_fetch_all_device",
    "hadros_batch_set_tuning", "hadros_batch_set_path", "hadros_batch_set_pool", "hadros_batch_autotune",
    "hadros_batch_run", "hadros_batch_last_ms",
    "hadros_batch_fetch", "hadros_batch_stats", "hadros_batch_info", "hadros_batch_destroy",
    "hadros_mlp_create", "hadros_mlp_n_params", "hadros_mlp_get_params", "hadros_mlp_set_params",
    "hadros_mlp_set_norm", "hadros_mlp_forward", "hadros_mlp_survival", "hadros_mlp_set_infer_path",
    "hadros_mlp_grad", "hadros_mlp_train_step", "hadros_mlp_destroy", "hadros_opacity_surface",
]


class HadrosError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hadros error {code}: {msg}")
        self.code = code


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def lib():
    """Load libhadros.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HadrosError(abi.ERR_DEVICE, f"{LIB_PATH} not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    L.hadros_last_error.restype = C.c_char_p
    for f in ("hadros_profile_name", "hadros_source_name", "hadros_spectrum_name"):
        getattr(L, f).restype = C.c_char_p
        getattr(L, f).argtypes = [C.c_int32]
    L.hadros_device_info.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.hadros_camera_defaults.argtypes = [C.POINTER(abi.Camera)]
    L.hadros_scene_defaults.argtypes = [C.POINTER(abi.Scene)]
    for f in ("hadros_parse_profile", "hadros_parse_source", "hadros_parse_spectrum"):
        getattr(L, f).argtypes = [C.c_char_p, C.POINTER(C.c_int32)]
    L.hadros_sigma_load.argtypes = [C.c_char_p, _dp, _dp, _dp, C.c_int, C.POINTER(C.c_int)]
    L.hadros_sigma_check.argtypes = [_dp, _dp, C.c_int]
    L.hadros_trace_rt.argtypes = [C.POINTER(abi.Camera), C.POINTER(abi.Scene), _dp, _dp, C.c_int, C.c_int,
                                  C.POINTER(abi.ImageOut), C.POINTER(abi.Stats)]
    L.hadros_trace_rt_noalloc.argtypes = [C.POINTER(abi.Camera), C.POINTER(abi.Scene), _dp, _dp, C.c_int, C.c_int,
                                          C.POINTER(abi.ImageOut), C.c_void_p, C.c_void_p]
    L.hadros_tracer_create.argtypes = [C.POINTER(abi.Camera), C.POINTER(abi.Scene), _dp, _dp, C.c_int, C.c_int,
                                       C.POINTER(C.c_void_p)]
    L.hadros_tracer_set_block.argtypes = [C.c_void_p, C.c_int]
    L.hadros_tracer_set_path.argtypes = [C.c_void_p, C.c_int]
    L.hadros_tracer_set_segments.argtypes = [C.c_void_p, C.c_int]
    L.hadros_tracer_run.argtypes = [C.c_void_p, C.c_void_p]
    L.hadros_tracer_last_ms.argtypes = [C.c_void_p, _dp]
    L.hadros_tracer_kernel_ms.argtypes = [C.c_void_p, _dp, _dp]
    L.hadros_tracer_fetch.argtypes = [C.c_void_p, C.POINTER(abi.ImageOut), C.POINTER(abi.Stats)]
    L.hadros_tracer_destroy.argtypes = [C.c_void_p]
    L.hadros_rt_from_cache.argtypes = [C.POINTER(abi.CacheView), C.POINTER(abi.Scene), _dp, _dp, C.c_int, C.c_int,
                                       C.POINTER(abi.ImageOut), C.POINTER(abi.Stats)]
    L.hadros_replay_create.argtypes = [C.POINTER(abi.CacheView), C.POINTER(abi.Scene), _dp, _dp, C.c_int, C.c_int,
                                       C.POINTER(C.c_void_p)]
    L.hadros_replay_run.argtypes = [C.c_void_p, C.c_void_p]
    L.hadros_replay_last_ms.argtypes = [C.c_void_p, _dp]
    L.hadros_replay_fetch.argtypes = [C.c_void_p, C.POINTER(abi.ImageOut), C.POINTER(abi.Stats)]
    L.hadros_replay_destroy.argtypes = [C.c_void_p]
    L.hadros_trace_cache.argtypes = [C.POINTER(abi.Camera), _lp, _ip, _dp, C.c_int64]
    L.hadros_trace_cache_range.argtypes = [C.POINTER(abi.Camera), C.c_int64, C.c_int64, _lp, _ip, _dp, C.c_int64]
    L.hadros_batch_create.argtypes = [C.POINTER(abi.Camera), C.POINTER(abi.Scene), C.c_int, _dp, _dp, C.c_int,
                                      C.c_int, C.POINTER(C.c_void_p)]
    L.hadros_batch_create_sub.argtypes = [C.POINTER(abi.Camera), C.POINTER(abi.Scene), C.c_int, _dp, _dp, C.c_int,
                                          C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.hadros_batch_fetch_all.argtypes = [C.c_void_p, C.POINTER(abi.ImageOut)]
    L.hadros_batch_fetch_all_device.argtypes = [C.c_void_p, C.POINTER(abi.ImageOut), C.c_void_p]
    L.hadros_batch_set_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.hadros_batch_set_path.argtypes = [C.c_void_p, C.c_int]
    L.hadros_batch_set_pool.argtypes = [C.c_void_p, C.c_double]
    L.hadros_batch_autotune.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.hadros_batch_run.argtypes = [C.c_void_p, C.c_void_p]
    L.hadros_batch_last_ms.argtypes = [C.c_void_p, _dp]
    L.hadros_batch_fetch.argtypes = [C.c_void_p, C.c_int, C.POINTER(abi.ImageOut)]
    L.hadros_batch_stats.argtypes = [C.c_void_p, C.POINTER(abi.Stats)]
    L.hadros_batch_info.argtypes = [C.c_void_p, _lp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.hadros_batch_destroy.argtypes = [C.c_void_p]
    _fp = C.POINTER(C.c_float)
    L.hadros_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64,
                                    C.POINTER(C.c_void_p)]
    L.hadros_mlp_n_params.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    L.hadros_mlp_get_params.argtypes = [C.c_void_p, _fp, C.c_int64]
    L.hadros_mlp_set_params.argtypes = [C.c_void_p, _fp, C.c_int64]
    L.hadros_mlp_set_norm.argtypes = [C.c_void_p, _fp, _fp, _fp, _fp]
    L.hadros_mlp_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    L.hadros_mlp_survival.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    L.hadros_mlp_set_infer_path.argtypes = [C.c_void_p, C.c_int]
    L.hadros_mlp_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, _fp, _fp]
    L.hadros_mlp_train_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_float, _fp]
    L.hadros_mlp_destroy.argtypes = [C.c_void_p]
    L.hadros_eval_rho.argtypes = [C.POINTER(abi.Medium), _dp, _dp, _dp, C.c_int64]
    L.hadros_opacity_surface.argtypes = [C.POINTER(abi.Medium), _dp, _dp, C.c_int, C.c_double, C.c_double,
                                         C.c_double, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _ip]
    L.hadros_eval_sigma.argtypes = [_dp, _dp, C.c_int, _dp, _dp, C.c_int64]
    L.hadros_eval_emissivity.argtypes = [C.POINTER(abi.Scene), _dp, _dp, _dp, _dp, C.c_int64]
    L.hadros_eval_mev.argtypes = [C.POINTER(abi.Scene), C.POINTER(abi.Mev), _dp, _dp, _dp, _dp, C.c_int64]
    for f in ("hadros_tracer_set_mev", "hadros_replay_set_mev"):
        getattr(L, f).argtypes = [C.c_void_p, C.POINTER(abi.Mev)]
    for f in ("hadros_tracer_fetch_mev", "hadros_replay_fetch_mev"):
        getattr(L, f).argtypes = [C.c_void_p, C.POINTER(abi.MevOut)]
    L.hadros_mev_defaults.argtypes = [C.POINTER(abi.Mev)]
    L.hadros_mev_defaults.restype = None
    for f in ("model", "flavor", "thermal", "ye", "spectral"):
        getattr(L, "hadros_parse_mev_" + f).argtypes = [C.c_char_p, C.POINTER(C.c_int32)]
    L.hadros_eval_rhs.argtypes = [C.c_double, C.c_int, _dp, _dp, C.c_int64]
    L.hadros_eval_step.argtypes = [C.c_double, C.c_int, C.c_double, _dp, _dp, C.c_int64]
    L.hadros_tau_along_ray.argtypes = [C.POINTER(abi.CacheView), C.POINTER(abi.Medium), _dp, _dp, C.c_int, C.c_double,
                                       _dp, _dp]
    L.hadros_eval_metric.argtypes = [C.c_double, _dp, _dp, _dp, C.c_int64]
    L.hadros_stream_probe.argtypes = [C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_double, C.c_void_p]
    L.hadros_eval_transfer_point.argtypes = [C.POINTER(abi.Scene), _dp, _dp, C.c_int, C.c_double, C.c_int, _dp, _dp,
                                             _dp, _dp, _dp, C.c_int64]
    _lib = L
    return L


def check(rc):
    if rc != abi.OK:
        raise HadrosError(rc, lib().hadros_last_error().decode(errors="replace"))


def device_info():
    L = lib()
    buf = C.create_string_buffer(256)
    cu, arch = C.c_int(), C.c_int()
    check(L.hadros_device_info(buf, 256, C.byref(cu), C.byref(arch)))
    return buf.value.decode(), cu.value, arch.value


def parse_profile(name):
    k = C.c_int32()
    check(lib().hadros_parse_profile(name.encode(), C.byref(k)))
    return k.value


def parse_source(name):
    k = C.c_int32()
    check(lib().hadros_parse_source(name.encode(), C.byref(k)))
    return k.value


def parse_spectrum(name):
    k = C.c_int32()
    check(lib().hadros_parse_spectrum(name.encode(), C.byref(k)))
    return k.value


def sigma_load(path):
    """SigmaTable(path): returns (E_GeV, sigma_GeV^-2, sigma_cm2) arrays."""
    L = lib()
    n = C.c_int()
    check(L.hadros_sigma_load(path.encode(), None, None, None, 0, C.byref(n)))
    e, g, s = np.zeros(n.value), np.zeros(n.value), np.zeros(n.value)
    check(L.hadros_sigma_load(path.encode(), dptr(e), dptr(g), dptr(s), n.value, C.byref(n)))
    return e, g, s


class Image:
    """Host arrays of one camera's (or cache's) outputs, pixel k = i*ny + j."""

    def __init__(self, npix, nbins):
        self.nbins = max(nbins, 0)
        self.tau = np.zeros(npix)
        self.p = np.zeros(npix)
        self.I = np.zeros(npix)
        self.captured = np.zeros(npix, np.int32)
        self.npoints = np.zeros(npix, np.int32)
        self.flags = np.zeros(npix, np.int32)
        self.alpha = np.zeros(npix)
        self.beta = np.zeros(npix)
        self.bin_tau = np.zeros((npix, self.nbins))
        self.bin_p = np.zeros((npix, self.nbins))
        self.bin_i = np.zeros((npix, self.nbins))

    def out(self):
        o = abi.ImageOut()
        o.tau, o.p_surv, o.intensity = dptr(self.tau), dptr(self.p), dptr(self.I)
        o.captured, o.npoints, o.flags = iptr(self.captured), iptr(self.npoints), iptr(self.flags)
        o.alpha, o.beta = dptr(self.alpha), dptr(self.beta)
        if self.nbins:
            o.bin_tau, o.bin_p, o.bin_i = dptr(self.bin_tau), dptr(self.bin_p), dptr(self.bin_i)
        return o


def _sig(sig_e, sig_s):
    return np.ascontiguousarray(sig_e, np.float64), np.ascontiguousarray(sig_s, np.float64)


class MevImage:
    """The five MeV columns per ray (hadros_mev_out) as numpy arrays."""

    FIELDS = ("tau", "p_surv", "intensity", "r_neutrinosphere", "leakage")

    def __init__(self, n):
        for f in self.FIELDS:
            setattr(self, f, np.zeros(n))

    def out(self):
        o = abi.MevOut()
        for f in self.FIELDS:
            setattr(o, f, getattr(self, f).ctypes.data)
        return o

    def stack(self):
        """[n x 5] in .dat column order (tau_mev P_surv_mev I_obs_mev r_neutrinosphere_rg leakage_mev)."""
        return np.stack([getattr(self, f) for f in self.FIELDS], 1)


def _fetch_mev(fn, h, n):
    img = MevImage(n)
    o = img.out()
    check(fn(h, C.byref(o)))
    return img


class Tracer:
    """Re-usable plan of hadros_trace_rt (device buffers allocated once)."""

    def __init__(self, cam, scene, sig_e, sig_s, precision=abi.PREC_FP64, block=0, path=abi.PATH_AUTO,
                 segments=0):
        L = lib()
        self._se, self._ss = _sig(sig_e, sig_s)
        self.cam = cam
        h = C.c_void_p()
        check(L.hadros_tracer_create(C.byref(cam), C.byref(scene), dptr(self._se), dptr(self._ss), self._se.size,
                                     precision, C.byref(h)))
        self.h = h
        if block:
            check(L.hadros_tracer_set_block(self.h, block))
        if path != abi.PATH_AUTO:
            check(L.hadros_tracer_set_path(self.h, path))
        if segments:
            check(L.hadros_tracer_set_segments(self.h, segments))

    def set_segments(self, n):
        """Tile-row bands of the split form (0 = automatic; hadros_tracer_set_segments)."""
        check(lib().hadros_tracer_set_segments(self.h, n))

    def run(self, stream=None):
        check(lib().hadros_tracer_run(self.h, stream))

    def last_ms(self):
        ms = C.c_double()
        check(lib().hadros_tracer_last_ms(self.h, C.byref(ms)))
        return ms.value

    def set_mev(self, mev):
        """MeV channel for the next runs (abi.mev(...)), or None to turn it off."""
        check(lib().hadros_tracer_set_mev(self.h, C.byref(mev) if mev is not None else None))

    def fetch_mev(self):
        return _fetch_mev(lib().hadros_tracer_fetch_mev, self.h, self.cam.nx * self.cam.ny)

    def kernel_ms(self):
        """(geodesic ms, transfer ms) of the last run: with one segment the two
        launches' durations, with more the spans of the overlapping bands
        (hadros_tracer_kernel_ms); the fused form reports its one kernel as
        the first."""
        g, r = C.c_double(), C.c_double()
        check(lib().hadros_tracer_kernel_ms(self.h, C.byref(g), C.byref(r)))
        return g.value, r.value

    def fetch(self):
        st = abi.Stats()
        check(lib().hadros_tracer_fetch(self.h, None, C.byref(st)))
        img = Image(self.cam.nx * self.cam.ny, st.n_bins)
        o = img.out()
        check(lib().hadros_tracer_fetch(self.h, C.byref(o), C.byref(st)))
        return img, st

    def close(self):
        if self.h:
            lib().hadros_tracer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def trace_rt(cam, scene, sig_e, sig_s, precision=abi.PREC_FP64, block=0, path=abi.PATH_AUTO):
    t = Tracer(cam, scene, sig_e, sig_s, precision, block, path)
    try:
        t.run()
        return t.fetch()
    finally:
        t.close()


class Cache:
    """A geodesic cache as SoA numpy arrays (what a KGEO file holds)."""

    def __init__(self, a_bh, offset, r, theta, dl, redshift, captured=None, pixel_i=None, pixel_j=None,
                 alpha=None, beta=None, x=None, y=None, z=None, nx=0, ny=0):
        self.a_bh = float(a_bh)
        self.offset = np.ascontiguousarray(offset, np.int64)
        self.r = np.ascontiguousarray(r, np.float64)
        self.theta = np.ascontiguousarray(theta, np.float64)
        self.dl = np.ascontiguousarray(dl, np.float64)
        self.redshift = np.ascontiguousarray(redshift, np.float64)
        n = self.offset.size - 1
        self.captured = captured if captured is not None else np.zeros(n, np.int32)
        self.pixel_i, self.pixel_j = pixel_i, pixel_j
        self.alpha, self.beta = alpha, beta
        self.x, self.y, self.z = x, y, z
        self.nx, self.ny = nx, ny

    @property
    def n_rays(self):
        return self.offset.size - 1

    def view(self):
        v = abi.CacheView()
        v.n_rays = self.n_rays
        v.a_bh = self.a_bh
        v.offset = self.offset.ctypes.data_as(_lp)
        v.r, v.theta, v.dl, v.redshift = dptr(self.r), dptr(self.theta), dptr(self.dl), dptr(self.redshift)
        return v


class Replay:
    """Re-usable plan of hadros_rt_from_cache with the cache resident in HBM."""

    def __init__(self, cache, scene, sig_e, sig_s, precision=abi.PREC_FP64):
        L = lib()
        self._se, self._ss = _sig(sig_e, sig_s)
        self.n = cache.n_rays
        self._v = cache.view()
        self._keep = cache
        h = C.c_void_p()
        check(L.hadros_replay_create(C.byref(self._v), C.byref(scene), dptr(self._se), dptr(self._ss),
                                     self._se.size, precision, C.byref(h)))
        self.h = h

    def run(self, stream=None):
        check(lib().hadros_replay_run(self.h, stream))

    def last_ms(self):
        ms = C.c_double()
        check(lib().hadros_replay_last_ms(self.h, C.byref(ms)))
        return ms.value

    def set_mev(self, mev):
        """MeV channel for the next runs (abi.mev(...)), or None to turn it off."""
        check(lib().hadros_replay_set_mev(self.h, C.byref(mev) if mev is not None else None))

    def fetch_mev(self):
        return _fetch_mev(lib().hadros_replay_fetch_mev, self.h, self.n)

    def fetch(self):
        st = abi.Stats()
        check(lib().hadros_replay_fetch(self.h, None, C.byref(st)))
        img = Image(self.n, st.n_bins)
        o = img.out()
        o.captured = o.npoints = o.flags = None
        o.alpha = o.beta = None
        check(lib().hadros_replay_fetch(self.h, C.byref(o), C.byref(st)))
        return img, st

    def close(self):
        if self.h:
            lib().hadros_replay_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rt_from_cache(cache, scene, sig_e, sig_s, precision=abi.PREC_FP64):
    p = Replay(cache, scene, sig_e, sig_s, precision)
    try:
        p.run()
        return p.fetch()
    finally:
        p.close()


def trace_cache(cam):
    """GPU geodesic cache of a camera: returns a Cache with x, y, z as well."""
    L = lib()
    npix = cam.nx * cam.ny
    off = np.zeros(npix + 1, np.int64)
    cap = np.zeros(npix, np.int32)
    check(L.hadros_trace_cache(C.byref(cam), off.ctypes.data_as(_lp), iptr(cap), None, 0))
    total = int(off[-1])
    pts = np.zeros((max(total, 1), 7))
    check(L.hadros_trace_cache(C.byref(cam), off.ctypes.data_as(_lp), iptr(cap), dptr(pts), total))
    pts = pts[:total]
    k = np.arange(npix)
    i, j = k // cam.ny, k % cam.ny
    th = np.tan(0.5 * cam.fov_deg * np.pi / 180.0)
    alpha = (2.0 * (i + 0.5) / cam.nx - 1.0) * th
    beta = (2.0 * (j + 0.5) / cam.ny - 1.0) * th
    return Cache(cam.a_spin, off, pts[:, 0], pts[:, 1], pts[:, 5], pts[:, 6], captured=cap, pixel_i=i, pixel_j=j,
                 alpha=alpha, beta=beta, x=pts[:, 2], y=pts[:, 3], z=pts[:, 4], nx=cam.nx, ny=cam.ny)


class Batch:
    """Persistent multi-camera tracer (hadros_batch_*): one launch for all cameras."""

    def __init__(self, cams, scenes, sig_e, sig_s, precision=abi.PREC_FP64, stride=1, offset=0,
                 path=abi.PATH_AUTO):
        L = lib()
        self._se, self._ss = _sig(sig_e, sig_s)
        n = len(cams)
        self.cams = (abi.Camera * n)(*cams)
        self.scenes = (abi.Scene * n)(*scenes)
        h = C.c_void_p()
        check(L.hadros_batch_create_sub(self.cams, self.scenes, n, dptr(self._se), dptr(self._ss), self._se.size,
                                        precision, stride, offset, C.byref(h)))
        self.h = h
        self.n = n
        self.stride, self.offset = stride, offset
        if path != abi.PATH_AUTO:
            self.set_path(path)

    def set_path(self, path):
        """abi.PATH_AUTO, PATH_FUSED or PATH_SPLIT (hadros_batch_set_path)."""
        check(lib().hadros_batch_set_path(self.h, path))

    def set_pool(self, max_bytes):
        """Bound of the split form's chunk pool in bytes (0 = default)."""
        check(lib().hadros_batch_set_pool(self.h, float(max_bytes)))

    ALL_FIELDS = ("tau", "p_surv", "intensity", "captured", "npoints", "flags", "bin_tau", "bin_p", "bin_i")

    def fetch_all(self, nbins, fields=ALL_FIELDS):
        """Outputs of every ray of the batch (camera-major); only `fields`
        (hadros_image_out member names) are copied, the others stay zero."""
        n, nb = self.info()["n_rays"], self.stats().n_bins
        if nbins is not None and nbins != nb:
            raise ValueError(f"batch stores {nb} bins per ray, asked for {nbins}")
        img = Image(n, nb)
        o = img.out()
        o.alpha = o.beta = None
        for f in self.ALL_FIELDS:
            if f not in fields:
                setattr(o, f, None)
        check(lib().hadros_batch_fetch_all(self.h, C.byref(o)))
        return img

    def fetch_all_device(self, **dst):
        """Device-to-device copy of whole-batch outputs into caller tensors:
        keyword = hadros_image_out member name, value = contiguous torch CUDA
        tensor of the matching size and dtype (float64, int32 for the ints)."""
        import torch

        n, nb = self.info()["n_rays"], self.stats().n_bins
        o = abi.ImageOut()
        for k, t in dst.items():
            if k not in self.ALL_FIELDS:
                raise ValueError(f"unknown output field {k}")
            want = n * nb if k.startswith("bin_") else n
            dt = torch.int32 if k in ("captured", "npoints", "flags") else torch.float64
            if not (t.is_cuda and t.is_contiguous() and t.dtype == dt and t.numel() == want):
                raise ValueError(f"{k}: need a contiguous CUDA {dt} tensor of {want} elements")
            setattr(o, k, C.cast(C.c_void_p(t.data_ptr()), _ip if dt == torch.int32 else _dp))
        check(lib().hadros_batch_fetch_all_device(self.h, C.byref(o), None))
        return dst

    def set_tuning(self, block, steps_per_chunk):
        check(lib().hadros_batch_set_tuning(self.h, block, steps_per_chunk))

    def autotune(self):
        b, s = C.c_int(), C.c_int()
        check(lib().hadros_batch_autotune(self.h, C.byref(b), C.byref(s)))
        return b.value, s.value

    def info(self):
        n, b, s, g = C.c_int64(), C.c_int(), C.c_int(), C.c_int()
        check(lib().hadros_batch_info(self.h, C.byref(n), C.byref(b), C.byref(s), C.byref(g)))
        return {"n_rays": n.value, "block": b.value, "steps_per_chunk": s.value, "grid": g.value}

    def run(self, stream=None):
        check(lib().hadros_batch_run(self.h, stream))

    def last_ms(self):
        ms = C.c_double()
        check(lib().hadros_batch_last_ms(self.h, C.byref(ms)))
        return ms.value

    def stats(self):
        st = abi.Stats()
        check(lib().hadros_batch_stats(self.h, C.byref(st)))
        return st

    def cam_bins(self, cam):
        """Bins stored per ray for camera `cam` (the library's rule: 1 for a
        monochromatic scene, 0 when the spectrum's range misses the sigma
        table, else n_bins)."""
        sp = self.scenes[cam].spectrum
        if sp.kind == abi.SPECTRUM_KINDS["mono"]:
            return 1
        lo, hi = max(sp.e_min_GeV, self._se[0]), min(sp.e_max_GeV, self._se[-1])
        return max(sp.n_bins, 1) if (lo > 0.0 and hi > lo) else 0

    def fetch(self, cam, nbins=None):
        c = self.cams[cam]
        nbc = self.cam_bins(cam)
        if nbins is not None and nbins != nbc:
            raise ValueError(f"camera {cam} stores {nbc} bins per ray, asked for {nbins}")
        sub = lambda n: max(0, (n - self.offset + self.stride - 1) // self.stride)  # noqa: E731
        img = Image(sub(c.nx) * sub(c.ny), nbc)
        o = img.out()
        check(lib().hadros_batch_fetch(self.h, cam, C.byref(o)))
        return img

    def close(self):
        if self.h:
            lib().hadros_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    """The BF16-MFMA surrogate network (hadros_mlp_*).  Data arguments are
    torch CUDA float32 tensors (device memory plumbing only)."""

    def __init__(self, n_in=14, width=256, depth=4, n_out=16, train_batch=8192, infer_batch=1 << 20, seed=0):
        L = lib()
        h = C.c_void_p()
        check(L.hadros_mlp_create(n_in, width, depth, n_out, train_batch, infer_batch, seed, C.byref(h)))
        self.h = h
        self.n_in, self.width, self.depth, self.n_out = n_in, width, depth, n_out
        n = C.c_int64()
        check(L.hadros_mlp_n_params(self.h, C.byref(n)))
        self.n_params = n.value

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        check(lib().hadros_mlp_get_params(self.h, fptr(p), p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(lib().hadros_mlp_set_params(self.h, fptr(p), p.size))

    def set_norm(self, x_mean, x_std, y_mean, y_std):
        a = [np.ascontiguousarray(v, np.float32) for v in (x_mean, x_std, y_mean, y_std)]
        self._norm = a
        check(lib().hadros_mlp_set_norm(self.h, *(fptr(v) for v in a)))

    def forward(self, x, y=None, stream=None):
        import torch

        assert x.is_cuda and x.dtype == torch.float32 and x.is_contiguous() and x.shape[1] == self.n_in
        if y is None:
            y = torch.empty((x.shape[0], self.n_out), dtype=torch.float32, device=x.device)
        assert y.is_cuda and y.dtype == torch.float32 and y.is_contiguous() and y.numel() == x.shape[0] * self.n_out
        check(lib().hadros_mlp_forward(self.h, C.c_void_p(x.data_ptr()), x.shape[0], C.c_void_p(y.data_ptr()),
                                       stream))
        return y

    def survival(self, x, p=None, stream=None):
        """P = exp(-10^y_b) of the first n_out/2 outputs (log10 tau per bin)."""
        import torch

        assert x.is_cuda and x.dtype == torch.float32 and x.is_contiguous() and x.shape[1] == self.n_in
        nb = self.n_out // 2
        if p is None:
            p = torch.empty((x.shape[0], nb), dtype=torch.float32, device=x.device)
        assert p.is_cuda and p.dtype == torch.float32 and p.is_contiguous() and p.numel() == x.shape[0] * nb
        check(lib().hadros_mlp_survival(self.h, C.c_void_p(x.data_ptr()), x.shape[0], C.c_void_p(p.data_ptr()),
                                        stream))
        return p

    def set_infer_path(self, path):
        check(lib().hadros_mlp_set_infer_path(self.h, path))

    def grad(self, x, t):
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        check(lib().hadros_mlp_grad(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()), x.shape[0],
                                    fptr(g), C.byref(loss)))
        return g, loss.value

    def train_step(self, x, t, lr, want_loss=True):
        """One Adam step; returns the batch loss, or None with want_loss=False
        (then nothing is read back and the step stays asynchronous)."""
        loss = C.c_float()
        check(lib().hadros_mlp_train_step(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()), x.shape[0],
                                          lr, C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def close(self):
        if self.h:
            lib().hadros_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def opacity_surface(medium, sig_e, sig_s, enu, m_bh, tau_surface, nr, ntheta):
    """hadros_opacity_surface: dict of theta, r_tau, tau_total, cls (ntheta
    doubles each) and crossing (int32)."""
    se, ss = _sig(sig_e, sig_s)
    o = {k: np.zeros(ntheta) for k in ("theta", "r_tau", "tau_total", "cls")}
    cr = np.zeros(ntheta, np.int32)
    check(lib().hadros_opacity_surface(C.byref(medium), dptr(se), dptr(ss), se.size, enu, m_bh, tau_surface, nr,
                                       ntheta, dptr(o["theta"]), dptr(o["r_tau"]), dptr(o["tau_total"]),
                                       dptr(o["cls"]), iptr(cr)))
    o["crossing"] = cr
    return o


def eval_rho(medium, r, th):
    r = np.ascontiguousarray(r, np.float64)
    th = np.ascontiguousarray(th, np.float64)
    out = np.zeros(r.size)
    check(lib().hadros_eval_rho(C.byref(medium), dptr(r), dptr(th), dptr(out), r.size))
    return out


def eval_sigma(sig_e, sig_s, e):
    se, ss = _sig(sig_e, sig_s)
    e = np.ascontiguousarray(e, np.float64)
    out = np.zeros(e.size)
    check(lib().hadros_eval_sigma(dptr(se), dptr(ss), se.size, dptr(e), dptr(out), e.size))
    return out


def eval_emissivity(scene, r, th, e):
    r, th, e = (np.ascontiguousarray(x, np.float64) for x in (r, th, e))
    out = np.zeros(r.size)
    check(lib().hadros_eval_emissivity(C.byref(scene), dptr(r), dptr(th), dptr(e), dptr(out), r.size))
    return out


def eval_mev(scene, mev, r, th, g):
    """(T_MeV, Ye, kappa, j) per point, [n x 4] (hadros_eval_mev)."""
    r, th, g = (np.ascontiguousarray(x, np.float64) for x in (r, th, g))
    out = np.zeros((r.size, 4))
    check(lib().hadros_eval_mev(C.byref(scene), C.byref(mev), dptr(r), dptr(th), dptr(g), dptr(out), r.size))
    return out


def eval_rhs(a, mode, y):
    y = np.ascontiguousarray(y, np.float64).reshape(-1, 8)
    out = np.zeros_like(y)
    check(lib().hadros_eval_rhs(a, mode, dptr(y), dptr(out), y.shape[0]))
    return out


STEP_RK4, STEP_RKF45_REF, STEP_RKF45_CFG, STEP_RKF45_CFG_SHIFT = 0, 1, 2, 3


def eval_transfer_point(scene, sig_e, sig_s, a_bh, precision, r, th, dl, z):
    """The transfer kernel's terms at points (r, theta, dl, redshift)
    (hadros_eval_transfer_point): a dict of ok, z, lnz, g3Adl, nbar_dl, rho
    [n] and dtau, emis [n x MAX_BINS] (emis is the log-form term for FP32K)."""
    r, th, dl, z = (np.ascontiguousarray(x, np.float64) for x in (r, th, dl, z))
    sig_e = np.ascontiguousarray(sig_e, np.float64)
    sig_s = np.ascontiguousarray(sig_s, np.float64)
    out = np.zeros((r.size, abi.EVAL_TRANSFER_COLS))
    check(lib().hadros_eval_transfer_point(C.byref(scene), dptr(sig_e), dptr(sig_s), sig_e.size, a_bh, precision,
                                           dptr(r), dptr(th), dptr(dl), dptr(z), dptr(out), r.size))
    nb = abi.MAX_BINS
    return {"ok": out[:, 0] != 0.0, "z": out[:, 1], "lnz": out[:, 2], "g3Adl": out[:, 3], "nbar_dl": out[:, 4],
            "rho": out[:, 5], "dtau": out[:, 6:6 + nb], "emis": out[:, 6 + nb:6 + 2 * nb]}


def stream_probe(kind, buf, partial=None, value=1.0):
    """Launch a streaming kernel of known traffic on torch's current stream
    (hadros_stream_probe): kind 0 reads the float64 device tensor buf once
    into partial (abi.STREAM_GRID doubles), kind 1 fills buf with value."""
    import torch

    assert buf.dtype == torch.float64 and buf.is_cuda and buf.is_contiguous()
    pp = None if partial is None else C.c_void_p(partial.data_ptr())
    if partial is not None:
        assert partial.dtype == torch.float64 and partial.numel() >= abi.STREAM_GRID
    check(lib().hadros_stream_probe(kind, C.c_void_p(buf.data_ptr()), buf.numel(), pp, value,
                                    C.c_void_p(torch.cuda.current_stream().cuda_stream)))


def eval_step(a, kind, h, y):
    """One integrator step of each state (hadros_eval_step): (y4 [n x 5], y5 [n x 5], ok [n])
    with components r, theta, phi, p_r, p_theta."""
    y = np.ascontiguousarray(y, np.float64).reshape(-1, 8)
    out = np.zeros((y.shape[0], 11))
    check(lib().hadros_eval_step(a, kind, h, dptr(y), dptr(out), y.shape[0]))
    return out[:, :5], out[:, 5:10], out[:, 10] != 0.0


def tau_along_ray(cache, medium, sig_e, sig_s, enu_inf):
    """The legacy per-ray optical depth and survival of a cache (hadros_tau_along_ray)."""
    se, ss = _sig(sig_e, sig_s)
    tau = np.zeros(cache.n_rays)
    p = np.zeros(cache.n_rays)
    v = cache.view()
    check(lib().hadros_tau_along_ray(C.byref(v), C.byref(medium), dptr(se), dptr(ss), se.size, float(enu_inf),
                                     dptr(tau), dptr(p)))
    return tau, p


def eval_metric(a, r, th):
    r = np.ascontiguousarray(r, np.float64)
    th = np.ascontiguousarray(th, np.float64)
    out = np.zeros((r.size, 10))
    check(lib().hadros_eval_metric(a, dptr(r), dptr(th), dptr(out), r.size))
    return out
