"""Test infrastructure only: a NumPy restatement of the warm-start network of
libapbf_hip (16-64-64-1, ReLU), with the BF16 rounding of every matrix-core
operand the HIP kernels use (inputs, weights, hidden activations, and the
pre-activation gradients of the backward pass) and everything else in float64.
Used by tests to check the GPU forward pass and gradients.

With exact=True every BF16 rounding is switched off and all arithmetic is
float64; tests/test_oracle_pins.py checks that gradient against central
differences of `loss`. `_MUTATION` (default 0) selects a deliberately wrong
backward pass, only to show that check failing."""
import numpy as np

FEATURES, HIDDEN = 16, 64
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 1024, 1088, 5184, 5248, 5312
_MUTATION = 0


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _exact(a):
    return np.asarray(a, np.float64)


def unpack(p, dtype=np.float32):
    p = np.asarray(p, dtype)
    return (p[OFF_W1:OFF_B1].reshape(HIDDEN, FEATURES), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(HIDDEN, HIDDEN),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, norm, x, exact=False):
    rnd = _exact if exact else (lambda a: bf16(np.asarray(a, np.float32)))
    w1, b1, w2, b2, w3, b3 = unpack(p, np.float64 if exact else np.float32)
    mean, scale, out_scale = norm[:16], norm[16:32], norm[32]
    if exact:
        xn = (np.asarray(x, np.float64) - np.asarray(mean, np.float64)) * np.asarray(scale, np.float64)
    else:
        xn = ((np.asarray(x, np.float32) - mean.astype(np.float32)) * scale.astype(np.float32)).astype(np.float32)
    xb = rnd(xn)
    z1 = xb @ rnd(w1).T + b1
    h1b = rnd(np.maximum(z1, 0.0))
    z2 = h1b @ rnd(w2).T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ np.asarray(w3, np.float64) + float(b3)
    return y * float(out_scale), dict(xb=xb, z1=z1, h1b=h1b, z2=z2, h2=h2, y=y)


def loss(p, norm, x, t, exact=True):
    """0.5 * mean((y - t / out_scale)^2) of the network output y."""
    _, c = forward(p, norm, x, exact=exact)
    d = c["y"] - np.asarray(t, np.float64) / float(norm[32])
    return 0.5 * np.mean(d * d)


def gradient(p, norm, x, t, exact=False):
    """Gradient of 0.5 * mean((y - t / out_scale)^2) and the loss."""
    rnd = _exact if exact else (lambda a: bf16(np.asarray(a, np.float32)))
    w1, b1, w2, b2, w3, b3 = unpack(p, np.float64 if exact else np.float32)
    _, c = forward(p, norm, x, exact=exact)
    n = len(t)
    d = c["y"] - np.asarray(t, np.float64) / float(norm[32])
    loss_v = 0.5 * np.mean(d * d)
    dy = d / n
    g = np.zeros(len(p), np.float64)
    g[OFF_W3:OFF_B3] = c["h2"].T @ dy
    g[OFF_B3] = dy.sum()
    dz2 = np.outer(dy, np.asarray(w3, np.float64)) * (c["z2"] > 0)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz2b = rnd(dz2)
    g[OFF_W2:OFF_B2] = (dz2b.T @ c["h1b"]).reshape(-1)
    dh1 = dz2b @ (rnd(w2).T if _MUTATION == 2 else rnd(w2))
    dz1 = dh1 if _MUTATION == 1 else dh1 * (c["z1"] > 0)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W1:OFF_B1] = (rnd(dz1).T @ c["xb"]).reshape(-1)
    return g, loss_v
