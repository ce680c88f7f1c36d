"""Host side of the warm-start training: the input normalisation, from NumPy
arrays and from torch tensors (reduced on their own device)."""
import numpy as np


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
