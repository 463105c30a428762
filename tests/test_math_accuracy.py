"""Accuracy of the device logarithm / arc tangent used for cell indices
(galaxy_amd/csrc/galhip_math.h), tested through a host build of the same
header (IEEE basic operations only, no contraction: host and device return
the same bits).  Pins:
  - gh_log_ge1 within 1.5 ulp of double, gh_atan2 within 3 ulp, against
    64-bit-mantissa logl / atan2l;
  - the FP32 values the reference forms from them (u = logf(1+r)/alpha,
    t = atan2 * na / 2 pi rounded to float) agree with those formed from
    glibc's double log / atan2 on every sample;
  - IEEE signed-zero results of atan2.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("math") / "math_accuracy")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "math_accuracy.cpp")], check=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2])
def test_log_atan2_accuracy(harness, seed):
    out = subprocess.run([harness, "2000000", str(seed)], check=True,
                         capture_output=True, text=True).stdout.split()
    n, max_log, max_atan, mis_log, mis_u, mis_t, zeros = out
    assert float(max_log) <= 1.5, max_log
    assert float(max_atan) <= 3.0, max_atan
    assert int(mis_log) == 0 and int(mis_u) == 0 and int(mis_t) == 0, out
    assert zeros == "1"
