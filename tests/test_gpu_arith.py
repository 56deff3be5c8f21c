"""The pair terms' shortened correctly rounded square root and division
(sqrtRn, divRn in libfluid_amd/csrc/lf_internal.hpp) give the same bits as
hipcc's own sqrtf and FP32 division on the inputs they are used for:
tests/cpp/arith_check.hip checks every float from 2^-96 up for the square
root, the bound below 2^-96, and 2^32 quotients of the spiky term's operands
(1e-4 <= h <= 1e4) plus 2^32 random ones, bitwise on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "libfluid_amd", "bin", "arith_check")


def test_shortened_sqrt_and_division_match_ieee():
    assert os.path.exists(BIN), "build first: make -C libfluid_amd"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 4
    assert all((" mismatches 0 " in l) or (" above 2^-47 0 " in l) for l in lines), r.stdout
