"""Host-side code without a GPU: the warm-start training's input
normalisation (NumPy arrays and torch tensors), the SVG restatement's float
formatting and the solver's reciprocal quotient compiled for the CPU."""
import os

import numpy as np
import pytest


def test_fit_norm_numpy_and_torch_agree():
    import torch
    from apbf_amd import learned as Lm
    rng = np.random.default_rng(2)
    x = (rng.standard_normal((5000, Lm.FEATURES)) * rng.uniform(0.1, 10, Lm.FEATURES)
         + rng.uniform(-5, 5, Lm.FEATURES)).astype(np.float32)
    x[:, 5] = 3.25  # a constant feature: zero spread, scale 1
    t = rng.standard_normal(5000).astype(np.float32) * 0.7
    a = Lm.fit_norm(x, t)
    b = Lm.fit_norm(torch.from_numpy(x), torch.from_numpy(t))
    assert a.dtype == b.dtype == np.float32 and a.shape == b.shape == (2 * Lm.FEATURES + 1,)
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-6)
    assert a[5] == np.float32(3.25) and a[Lm.FEATURES + 5] == 1.0
    np.testing.assert_allclose(a[:Lm.FEATURES], x.astype(np.float64).mean(0), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(a[-1], np.sqrt(np.mean(t.astype(np.float64) ** 2)), rtol=1e-6)


def test_svg_float_format_matches_std_format():
    """The SVG restatement's float formatting against libstdc++'s
    std::to_chars(float) outputs (the "{}" of std::format), recorded from a
    g++ 11 build: shortest round-trip digits, fixed or scientific by length."""
    from svg_text import fmt32
    cases = {0.1: "0.1", 1e-7: "1e-07", 123456789.0: "123456792", 1.5: "1.5", -0.25: "-0.25",
             1e20: "1e+20", 3.0: "3", 100.0: "100", 1234567.0: "1234567", 1e-4: "1e-04", 0.0: "0",
             0.001: "0.001", 60.125: "60.125", -60.0: "-60", 0.7: "0.7", 2e-3: "0.002", 1e5: "1e+05",
             1e6: "1e+06", 12345.678: "12345.678", 123.456: "123.456", 5e-5: "5e-05", 0.05: "0.05",
             -1e-10: "-1e-10", 3.4e38: "3.4e+38", 1.17549435e-38: "1.1754944e-38", 262144.0: "262144",
             16777217.0: "16777216", 0.030000001: "0.030000001"}
    for v, want in cases.items():
        assert fmt32(v) == want, (v, fmt32(v), want)


def test_reciprocal_quotient_equals_ieee_division(tmp_path):
    """The solver divides by inverse masses through the reciprocal and one
    fused correction (quotient_by_recip); compiled for the CPU, it gives the
    IEEE quotient for 64 divisors, each with every dividend significand
    (5.4e8 cases, divisor exponents -40..40, dividend exponents -80..80)."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc absent")
    exe = str(tmp_path / "quotient_check")
    src = os.path.join(os.path.dirname(__file__), "quotient_check.c")
    flags = ["-O2", "-ffp-contract=off"]
    if subprocess.run([gcc, *flags, "-march=native", "-o", exe, src, "-lm"]).returncode != 0:
        subprocess.run([gcc, *flags, "-o", exe, src, "-lm"], check=True)
    out = subprocess.run([exe, "64"], capture_output=True, text=True, check=True, timeout=300).stdout.split()
    bad, total = int(out[0]), int(out[1])
    assert total == 64 * (1 << 23)
    assert bad == 0, bad
