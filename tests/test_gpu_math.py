"""The short correctly rounded square root and quotient of the pair kernels (sqrtCR, divCR in
tsp_device.h) against the compiler's general correctly rounded lowering, bit for bit, on the GPU:
0 and every float from 2^-96 to +inf for the square root (its domain); every dist in [h, 2h] for the gradient factor's
quotient at particle sizes 0.0008..1; 2^32 random operand pairs over divCR's stated range.
The program is tests/cpp/math_selftest.hip, built into turbospharticles_amd/lib/.
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_short_exact_sqrt_and_quotient_match_the_general_lowering():
    print("case: math_selftest (sqrt exhaustive, quotient ranges, random pairs)", flush=True)
    exe = os.path.join(ROOT, "turbospharticles_amd", "lib", "math_selftest")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS" in r.stdout
