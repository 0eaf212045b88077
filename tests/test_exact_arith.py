"""Exact-arithmetic identities the kernels rely on (CPU only)."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_u16_division_equals_ieee_quotient_exhaustively(tmp_path):
    """div_u16 (three FP32 operations with a host-rounded reciprocal) is the
    correctly rounded quotient for every uint16 numerator and range."""
    exe = str(tmp_path / "exact_div_u16")
    subprocess.run(["gcc", "-O2", "-mfma", "-fopenmp", "-ffp-contract=off", os.path.join(HERE, "exact_div_u16.c"),
                    "-o", exe, "-lm"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "0", out.stdout


@pytest.mark.parametrize("ratio", [2.0 ** k for k in range(0, 40, 3)])
def test_power_of_two_ratio_scaling_equals_division(ratio):
    """|z| * (1/ratio) equals |z| / ratio when ratio is a power of two (the
    loss stencil's scaled magnitude), over magnitudes spanning the float range."""
    rng = np.random.default_rng(0)
    x = (rng.random(100000) * 10.0 ** rng.integers(-30, 30, 100000)).astype(np.float32)
    r = np.float32(ratio)
    assert np.array_equal(x * (np.float32(1) / r), x / r)
