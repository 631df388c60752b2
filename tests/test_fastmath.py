"""The kernels' own math routines (hadros_amd/csrc/fastmath.h), built for the
host from the same header and checked against long-double references: the
transfer kernels' table-driven exponential (within 2.5 ulp of the correctly
rounded value over the finite range, including subnormal results, exact at the
special values) and logarithm (within 2 ulp over the whole range, subnormal
arguments included, exact at 1, 2, 0.75 and the special values), and the
geodesic kernels' sine/cosine, their stage form and the densities' cosine of
a polar angle."""
import math
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exp_check(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("fastmath") / "exp_tab_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "hadros_amd", "csrc"),
                    "-o", exe, os.path.join(ROOT, "tests", "cpp", "exp_tab_check.cpp")], check=True)
    return exe


def test_exp_tab_ulp_and_specials(exp_check):
    out = subprocess.run([exp_check, "2000000"], check=True, capture_output=True, text=True).stdout.split("\n")
    max_ulp = float(out[0].split()[1])
    assert max_ulp < 2.5, out[0]
    sp = out[1].split()[1:]
    assert float.fromhex(sp[0]) == 1.0
    assert float.fromhex(sp[1]) == math.exp(1.0)
    assert float.fromhex(sp[2]) == 0.0 and float.fromhex(sp[3]) == 0.0
    assert abs(float.fromhex(sp[4]) / math.exp(709.0) - 1.0) < 5e-16
    assert sp[5] == "1"


@pytest.fixture(scope="module")
def sincos_check(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("fastmath") / "sincos_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "hadros_amd", "csrc"),
                    "-o", exe, os.path.join(ROOT, "tests", "cpp", "sincos_check.cpp")], check=True)
    return exe


def test_sincos_and_stage_shift(sincos_check):
    """sincos_m (both evaluation forms) within 1.5 x 2^-53 of the long-double
    sine and cosine over +-4 pi; the stage form (angle addition from the first
    stage's sin/cos, offsets up to 1/4) within 2.5 x 2^-53; offsets beyond 1/4
    and NaN reported out of range."""
    out = subprocess.run([sincos_check], check=True, capture_output=True, text=True).stdout.split()
    e_fused, e_plain, e_shift, any_out = float(out[1]), float(out[2]), float(out[3]), int(out[4])
    assert e_fused < 1.5 and e_plain < 1.5, out
    assert e_shift < 2.5, out
    assert any_out == 0
    # cos_polar (the densities' cos theta) over [0, pi], near pi/2 and outside
    assert out[5] == "polar" and float(out[6]) < 2.5, out


@pytest.fixture(scope="module")
def log_check(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("fastmath") / "log_tab_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "hadros_amd", "csrc"),
                    "-o", exe, os.path.join(ROOT, "tests", "cpp", "log_tab_check.cpp")], check=True)
    return exe


def test_log_tab_ulp_and_specials(log_check):
    out = subprocess.run([log_check, "2000000"], check=True, capture_output=True, text=True).stdout.split("\n")
    max_ulp = float(out[0].split()[1])
    assert max_ulp < 2.0, out[0]
    sp = out[1].split()[1:]
    assert float.fromhex(sp[0]) == 0.0
    assert float.fromhex(sp[1]) == math.log(2.0)
    assert sp[2] == "-inf" and sp[3] == "inf"
    assert sp[4] == "1" and sp[5] == "1"
    assert float.fromhex(sp[6]) == math.log(0.75)
