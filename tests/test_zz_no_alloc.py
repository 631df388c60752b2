"""GPU: the allocation-free entry point hadros_trace_rt_noalloc.

It runs the fused kernel with its scene block, sigma table and counters in
the library's static device scratch, and writes into caller-owned device
buffers.  Its outputs must equal the plain tracer's fused path
(hadros_tracer_* with HADROS_PATH_FUSED) bit for bit: the same kernel on the
same inputs.  Two calls in a row, the second with a smaller camera, fewer
bins and a shorter sigma table, check that nothing of the first call's
scratch leaks into the second."""
import ctypes as C

import numpy as np
import pytest

from hadros_amd import abi, ffi, workloads
import scenes as S

pytestmark = pytest.mark.gpu

GUARD = 64   # extra elements after each output, which must stay untouched


def _same(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _noalloc(cam, scene, sig_e, sig_s, precision, nbins):
    """One hadros_trace_rt_noalloc call into fresh torch device buffers; returns host copies."""
    import torch

    dev = torch.device("cuda:0")
    npix = cam.nx * cam.ny
    f64 = dict(dtype=torch.float64, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    bufs = {
        "tau": torch.full((npix + GUARD,), -7.0, **f64),
        "p_surv": torch.full((npix + GUARD,), -7.0, **f64),
        "intensity": torch.full((npix + GUARD,), -7.0, **f64),
        "alpha": torch.full((npix + GUARD,), -7.0, **f64),
        "beta": torch.full((npix + GUARD,), -7.0, **f64),
        "captured": torch.full((npix + GUARD,), -7, **i32),
        "npoints": torch.full((npix + GUARD,), -7, **i32),
        "flags": torch.full((npix + GUARD,), -7, **i32),
        "bin_tau": torch.full((npix * nbins + GUARD,), -7.0, **f64),
        "bin_p": torch.full((npix * nbins + GUARD,), -7.0, **f64),
        "bin_i": torch.full((npix * nbins + GUARD,), -7.0, **f64),
    }
    counters = torch.full((4,), 12345, dtype=torch.int64, device=dev)
    o = abi.ImageOut()
    for k, t in bufs.items():
        ptr = C.c_void_p(t.data_ptr())
        setattr(o, k, C.cast(ptr, C.POINTER(C.c_int32) if t.dtype == torch.int32 else C.POINTER(C.c_double)))
    se = np.ascontiguousarray(sig_e, np.float64)
    ss = np.ascontiguousarray(sig_s, np.float64)
    stream = torch.cuda.current_stream(dev).cuda_stream
    ffi.check(ffi.lib().hadros_trace_rt_noalloc(C.byref(cam), C.byref(scene), ffi.dptr(se), ffi.dptr(ss), se.size,
                                                 precision, C.byref(o), C.c_void_p(counters.data_ptr()),
                                                 C.c_void_p(stream)))
    torch.cuda.synchronize(dev)
    host = {k: t.cpu().numpy() for k, t in bufs.items()}
    for k, v in host.items():
        assert np.all(v[-GUARD:] == -7), f"{k}: the call wrote past its output"
    return host, counters.cpu().numpy()


def _check_against_fused(cam, scene, sig_e, sig_s, precision):
    ref, st = ffi.trace_rt(cam, scene, sig_e, sig_s, precision, path=abi.PATH_FUSED)
    nb = st.n_bins
    got, cnt = _noalloc(cam, scene, sig_e, sig_s, precision, nb)
    npix = cam.nx * cam.ny
    pairs = (("tau", ref.tau), ("p_surv", ref.p), ("intensity", ref.I), ("alpha", ref.alpha),
             ("beta", ref.beta), ("captured", ref.captured), ("npoints", ref.npoints), ("flags", ref.flags))
    for k, want in pairs:
        assert _same(got[k][:npix], want), k
    for k, want in (("bin_tau", ref.bin_tau), ("bin_p", ref.bin_p), ("bin_i", ref.bin_i)):
        assert _same(got[k][:npix * nb], want.reshape(-1)), k
    assert list(cnt) == [st.total_points, st.total_attempts, st.n_captured, st.n_exhausted]
    return ref, st


@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_FP32K])
def test_noalloc_equals_fused_tracer_twice(precision):
    e, _, s = workloads.synthetic_sigma()
    big = workloads.c2_scene()                # the 512x512 workload's scene, 8 bins,
    big.medium.rho0 *= 1e-8                   # 1e-8 as dense: survival non-trivial
    big.medium.envelope_rho0 *= 1e-8
    # first call: 24 x 20 camera, 8 bins, 64-row table
    ref1, st1 = _check_against_fused(S.cfg_camera(nx=24, ny=20, a=0.95, theta_deg=75.0), big, e, s, precision)
    assert st1.n_bins == 8 and 0.0 < ref1.p.mean() < 1.0
    # second call, smaller in every size: 12 x 9 camera, one bin, 33-row table
    small = S.c1_lowdensity()
    ref2, st2 = _check_against_fused(S.cfg_camera(nx=12, ny=9, a=0.5, theta_deg=40.0), small, e[::2], s[::2],
                                     precision)
    assert st2.n_bins == 1 and st2.total_points > 0


def test_noalloc_ref_mode_and_default_stream():
    e, _, s = workloads.synthetic_sigma()
    _check_against_fused(S.small_ref_camera(nx=4, ny=4, h=0.2), S.parity_scenes()[1][1], e, s, abi.PREC_FP64)


def test_noalloc_rejects_oversized_table_and_null_output():
    e = np.logspace(4, 12, abi.MAX_SIGMA + 1)
    s = 7.84e-36 * e ** 0.363
    cam = S.cfg_camera(nx=8, ny=8)
    sc = S.c1_lowdensity()
    o = abi.ImageOut()
    L = ffi.lib()
    rc = L.hadros_trace_rt_noalloc(C.byref(cam), C.byref(sc), ffi.dptr(e), ffi.dptr(s), e.size, abi.PREC_FP64,
                                   C.byref(o), None, None)
    assert rc == abi.ERR_ARG
    rc = L.hadros_trace_rt_noalloc(C.byref(cam), C.byref(sc), ffi.dptr(e), ffi.dptr(s), 64, abi.PREC_FP64, None,
                                   None, None)
    assert rc == abi.ERR_ARG
