"""The planar build (DIMENSIONS 2, cpu_gpu_shared_config.h:2) of the oracle,
pinned by what the mathematics fixes: the 2D z-curve against an independent
bit interleave, the 2D hash-grid walk against brute force, and the 2D kernel
normalisations (unit area) by numerical integration. The planar rest
density, transfer volumes and split factor are pinned in
test_oracle_pins.py. CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _naive_hash_2d(p, mn, mx, res):
    out = []
    for q in p:
        g = []
        for d in range(2):
            x = np.float32(np.float32(q[d]) / np.float32(262144.0))
            v = np.float32((x - np.float32(mn[d])) / (np.float32(mx[d]) - np.float32(mn[d])) * np.float32(1 << res))
            g.append(0 if not v > 0 else (0xFFFFFFFF if v >= 4294967296.0 else int(v)))
        h = 0
        for i in range(res):
            h |= ((g[0] >> i) & 1) << (2 * i)
            h |= ((g[1] >> i) & 1) << (2 * i + 1)
        out.append(h & 0xFFFFFFFF)
    return np.array(out, np.uint32)


def test_planar_hash_bit_interleave():
    """calculate_position_hash.comp:29-36 with DIMENSIONS 2: bit i of x at 2i,
    of y at 2i + 1; z plays no part."""
    rng = np.random.default_rng(5)
    p = np.zeros((400, 4), np.int32)
    p[:, :3] = (rng.uniform(-2, 18, (400, 3)) * 262144).astype(np.int32)
    mn, mx = (0.0, 0.0, 0.0), (16.0, 8.0, 4.0)
    for res in (1, 4, 7, 12):
        got = O.position_hash(p, mn, mx, res, dims=2)
        assert np.array_equal(got, _naive_hash_2d(p, mn, mx, res))
        q = p.copy()
        q[:, 2] = rng.integers(-(1 << 30), 1 << 30, len(q))
        assert np.array_equal(O.position_hash(q, mn, mx, res, dims=2), got)
    q = np.zeros((4, 4), np.int32)
    q[:, :3] = np.array([[1.5, 0.5, 9.5], [0.5, 1.5, 0.5], [1.5, 1.5, 3.5], [2.5, 0.5, 0.5]]) * 262144
    assert O.position_hash(q, (0, 0, 0), (16, 16, 16), 4, dims=2).tolist() == [1, 2, 3, 4]
    # every cell of a 2^res x 2^res grid gets its own hash below 4^res
    res = 5
    c = (np.stack(np.meshgrid(np.arange(32), np.arange(32), indexing="ij"), -1).reshape(-1, 2) + 0.5) * 262144
    q = np.zeros((len(c), 4), np.int32)
    q[:, :2] = c
    h = O.position_hash(q, (0, 0, 0), (32, 32, 32), res, dims=2)
    assert np.array_equal(np.sort(h), np.arange(1 << (2 * res), dtype=np.uint32))


def _planar_cloud(n_side, seed, layers=1):
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(np.arange(n_side), np.arange(n_side), np.arange(layers), indexing="ij"),
                 -1).reshape(-1, 3)
    pw = 0.25 + 0.5 * g + rng.uniform(-0.05, 0.05, g.shape)
    p = np.zeros((len(pw), 4), np.int32)
    p[:, :3] = np.trunc(pw * 262144.0).astype(np.int32)
    return p


@pytest.mark.parametrize("res,scale,layers", [(2, 1.0, 1), (4, 1.5, 1), (6, 1.0, 3), (3, 1.0, 2)])
def test_planar_green_equals_brute_force(res, scale, layers):
    """neighborhood_green.comp:35-86 with DIMENSIONS 2: the walk covers the
    x-y cells of the search square in z cell 0 and tests 3D distances, so
    its pairs are exactly the brute-force pairs, also for particles off the
    plane (layers > 1)."""
    mn, mx = (0.0, 0.0, 0.0), (12.0, 12.0, 12.0)
    p = _planar_cloud(22, res, layers)
    h = O.position_hash(p, mn, mx, res, dims=2)
    order = np.argsort(h, kind="stable")
    p, h = p[order], h[order]
    rng = np.random.default_rng(9)
    rad = rng.uniform(0.5, 1.2, len(p)).astype(np.float32)
    s, e = O.find_value_ranges(h, 1 << (2 * res))
    green = O.neighborhood_green(p, rad, s, e, scale, mn, mx, res, 2_000_000, dims=2)
    brute = O.neighborhood_brute_force(p, rad, scale, 2_000_000)
    key = lambda a: a[np.lexsort((a[:, 1], a[:, 0]))]
    assert len(green) > 0
    assert np.array_equal(key(green), key(brute))


def _integrate_2d(kid, kw, n=801):
    """Midpoint sum of W over the square [-kw, kw]^2 (the support, or 7 sigma
    of the Gaussian) at z = 0."""
    ext = kw * (1.0 if kid in (3, 4) else 2.2)
    xs = (np.arange(n) + 0.5) / n * 2 * ext - ext
    da = (2 * ext / n) ** 2
    total = 0.0
    for x in xs[::1]:
        col = 0.0
        for y in xs:
            h, _ = O.kernel(kid, [x, y, 0.0], kw, dims=2)
            col += h
        total += col
    return total * da


@pytest.mark.parametrize("kid,kw", [(1, 4.0), (1, 2.5), (3, 3.0)])
def test_planar_kernels_have_unit_area(kid, kw):
    """kernels.glsl:48-104 in 2D: the Gaussian (height 0.6 / (kw/2)^2,
    variance from pow(height, 2/D)) and the cone (3 / (pi h^2)) integrate to
    1 over the plane; with the 3D constants they would not."""
    n = 161 if kid == 1 else 201
    area = _integrate_2d(kid, kw, n)
    assert abs(area - 1.0) < 2e-3, area


def _spike_from_gradient(h, dims, n=4000):
    """The quadratic spike's height rebuilt from its gradient magnitude,
    W(r) = integral from r to h of |grad W|, on a radial midpoint grid. (The
    reference's height, kernels.glsl:62-70, takes pow of a negative base
    inside the support, which is undefined, so its gradient carries the
    kernel.)"""
    rs = (np.arange(n) + 0.5) / n * h
    g = np.array([np.linalg.norm(O.kernel(4, [r, 0.0, 0.0], h, dims=dims)[1]) for r in rs])
    w = np.cumsum(g[::-1])[::-1] * (h / n)
    return rs, w


@pytest.mark.parametrize("dims,h", [(2, 2.0), (2, 0.7), (3, 1.5)])
def test_quadratic_spike_gradient_normalisation(dims, h):
    """kernels.glsl:72-82: the spike gradient -2 a (h - r) r / |r| with
    a = 6 / (pi h^4) in 2D and 15 / (2 pi h^5) in 3D belongs to a kernel of
    unit area (volume): the height rebuilt from it integrates to 1."""
    rs, w = _spike_from_gradient(h, dims)
    dr = h / len(rs)
    shell = 2 * math.pi * rs if dims == 2 else 4 * math.pi * rs ** 2
    total = float(np.sum(w * shell) * dr)
    assert abs(total - 1.0) < 3e-3, total


@pytest.mark.parametrize("kid", [3])
def test_kernels_have_unit_volume_in_3d(kid):
    """The 3D constant of the cone (3 / (pi h^3)): the radial integral of
    W 4 pi r^2 is 1."""
    h = 2.0
    n = 20000
    rs = (np.arange(n) + 0.5) / n * h
    w = np.array([O.kernel(kid, [r, 0.0, 0.0], h, dims=3)[0] for r in rs[::10]])
    vol = float(np.sum(w * 4 * math.pi * rs[::10] ** 2) * (h / n) * 10)
    assert abs(vol - 1.0) < 2e-3, vol


def test_planar_gradients_are_the_height_derivative():
    """The planar gradients (kernels.glsl:54-97) are -dW/dr along r, by
    central differences of the planar heights."""
    for kid, kw in ((1, 4.0), (3, 3.0)):
        for rr in (0.3, 0.7, 1.3):
            rv = np.array([rr * 0.6, rr * 0.8, 0.0])
            eps = 1e-3
            hp, _ = O.kernel(kid, rv * (1 + eps / rr), kw, dims=2)
            hm, _ = O.kernel(kid, rv * (1 - eps / rr), kw, dims=2)
            _, g = O.kernel(kid, rv, kw, dims=2)
            dwdr = (hp - hm) / (2 * eps)
            want = dwdr * rv / rr
            assert np.allclose(g, want, rtol=2e-3, atol=1e-6), (kid, rr, g, want)


def test_dimension_free_kernels_keep_their_3d_constants():
    """The cubic, poly6 and spiky kernels (kernels.glsl:3-46) carry 3D
    constants in either build: the planar values equal the 3D ones."""
    for kid in (0, 2):
        for r in ([0.2, 0.1, 0.0], [0.5, -0.3, 0.0]):
            h2, g2 = O.kernel(kid, r, 1.5, dims=2)
            h3, g3 = O.kernel(kid, r, 1.5, dims=3)
            assert h2 == h3 and np.array_equal(g2, g3)
