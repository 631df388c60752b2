"""bench.py's per-kernel roofline: the bound follows the kernel's algorithmic
intensity against the MI355X ridge (78.6 TFLOP/s FP64 over 8 TB/s), and the
fraction is against that bound's peak (no GPU needed)."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_ridge(bench):
    assert bench.RIDGE_FLOP_PER_BYTE == pytest.approx(78.6e12 / 8000e9)


def test_geodesic_kernel_is_fp64_bound(bench):
    attempts, points, rays = 42366154, 33540730, 262144
    flop = bench.geodesic_flops(attempts, points)
    assert flop == attempts * (6 * 69 + 390) + points * 60
    nbytes = bench.geodesic_bytes(points, rays)
    assert nbytes == 32.0 * points + 32.0 * rays
    r = bench.kernel_roofline("geo", 2.0, flop, nbytes)
    assert r["bound"] == "fp64" and r["unit"] == "TFLOP/s" and r["peak"] == 78.6
    assert r["achieved"] == pytest.approx(flop / 2.0e-3 / 1e12)
    assert r["frac"] == pytest.approx(r["achieved"] / 78.6)


def test_transfer_kernel_is_hbm_bound(bench):
    points, rays, nb = 33540730, 262144, 8
    flop = bench.transfer_flops(points, nb)
    nbytes = bench.transfer_bytes(points, rays, nb)
    assert nbytes == 32.0 * points + rays * (8.0 + 8.0 * (3 + 3 * nb))
    r = bench.kernel_roofline("rt", 2.1, flop, nbytes)
    assert r["intensity_flop_per_byte"] < bench.RIDGE_FLOP_PER_BYTE
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["frac"] == pytest.approx(nbytes / 2.1e-3 / 1e9 / 8000.0)
