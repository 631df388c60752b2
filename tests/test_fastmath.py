"""The transfer kernels' table-driven exponential (hadros_amd/csrc/fastmath.h),
built for the host from the same header and checked against the long-double
exponential: within 2.5 ulp of the correctly rounded value over the finite
range, including subnormal results, and exact at the special values."""
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
