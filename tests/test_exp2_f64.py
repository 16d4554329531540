"""The FP64 exponential of the soft-min passes (rwig_amd/csrc/exp2_f64.hpp),
built for the host from the same source and checked against long-double
exp2l: under one ulp over [-1022, 60], exact powers of two at integers, 0 at
and below the subnormal range's end and at -inf, NaN kept."""
import math
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def exp2_report(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not installed")
    exe = str(tmp_path_factory.mktemp("exp2") / "exp2_f64_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-o", exe,
                    os.path.join(HERE, "exp2_f64_check.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {k: float(v) for k, v in (line.split() for line in out.splitlines())}


def test_exp2_f64_is_within_one_ulp(exp2_report):
    assert exp2_report["max_ulp"] < 1.0


def test_exp2_f64_subnormal_results(exp2_report):
    assert exp2_report["max_subnormal_err_in_min_subnormals"] <= 1.0
    assert exp2_report["at_-1074"] == 5e-324


def test_exp2_f64_special_arguments(exp2_report):
    assert exp2_report["at_0"] == 1.0
    assert exp2_report["at_-1"] == 0.5
    assert exp2_report["at_-1075.5"] == 0.0
    assert exp2_report["at_-2000"] == 0.0
    assert exp2_report["at_-inf"] == 0.0
    assert exp2_report["at_nan_is_nan"] == 1
    assert math.isclose(exp2_report["at_1023_5"], 2.0 ** 1023 * math.sqrt(2.0), rel_tol=2e-16)
