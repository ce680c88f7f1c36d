"""CPU tests: the oracle against the reference's own self-test vectors, brute
force, closed forms and invariants, and the HIP library's exported interface.
Nothing here needs a GPU."""
import ctypes
import json
import math
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FIX = json.load(open(os.path.join(HERE, "golden", "reference_selftest.json")))


# ------------------------------------------------------------------ list algorithms
def test_prefix_sum_vectors():
    v = FIX["prefix_sum"]
    assert O.prefix_sum(v["values"]).tolist() == v["expected"]


def test_prefix_sum_long_running_sum():
    rng = np.random.default_rng(0)
    vals = (rng.integers(0, 2, 512 * 512 + 1000)).astype(np.uint32)
    assert np.array_equal(O.prefix_sum(vals), np.cumsum(vals).astype(np.uint32))


@pytest.mark.parametrize("name", ["sort", "sort_small_values"])
def test_sort_vectors(name):
    v = FIX[name]
    k, idx = O.stable_radix_sort(v["keys"], np.arange(len(v["keys"])))
    assert k.tolist() == v["expected_keys"] and idx.tolist() == v["expected_values"]


@pytest.mark.parametrize("bits", [4, 16, 32])
def test_sort_is_stable_sort(bits):
    rng = np.random.default_rng(bits)
    keys = rng.integers(0, 1 << bits, 512 * 512 + 123, dtype=np.uint64).astype(np.uint32)
    idx = np.arange(len(keys), dtype=np.uint32)
    k, v = O.stable_radix_sort(keys, idx)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order]) and np.array_equal(v, idx[order])


def test_sort_upper_bound_passes():
    """Digits are visited while upper_bound >> offset != 0 (algorithms.cpp:73):
    upper_bound 15 sorts on the low 4 bits only, 16 on 8 bits (the reference
    counts 2^k as k+1 bits)."""
    keys = np.array([0x35, 0x12, 0x21, 0x03], np.uint32)
    _, v = O.stable_radix_sort(keys, np.arange(4), 15)
    assert v.tolist() == [2, 1, 3, 0]
    _, v = O.stable_radix_sort(keys, np.arange(4), 16)
    assert v.tolist() == [3, 1, 2, 0]


@pytest.mark.parametrize("name", ["indexed_list_apply_hidden_edit_1", "indexed_list_apply_hidden_edit_2",
                                  "indexed_list_apply_hidden_edit_3"])
def test_apply_hidden_edit_vectors(name):
    v = FIX[name]
    a, _ = O.apply_hidden_edit(v["list_a"], v["hidden_length"], v["edit"], 100)
    b, _ = O.apply_hidden_edit(v["list_b"], v["hidden_length"], v["edit"], 100)
    assert a.tolist() == v["expected_a"] and b.tolist() == v["expected_b"]


def test_delete_these_vectors():
    v = FIX["delete_these_1"]
    e = O.delete_edit_list(v["list_b"], 13)
    assert np.array(v["hidden_values"])[e].tolist() == v["expected_hidden"]
    a, _ = O.apply_hidden_edit(v["list_a"], 13, e, 100)
    c, _ = O.apply_hidden_edit(v["list_c"], 13, e, 100)
    assert len(a) == v["expected_length_a"] and len(c) == v["expected_length_c"]
    assert c.tolist() == [0, 1]
    v3 = FIX["delete_these_3"]
    e3 = O.delete_edit_list(v3["list_b"], 8)
    assert e3.tolist() == v3["expected_hidden"]
    a3, _ = O.apply_hidden_edit(v3["list_a"], 8, e3, 100)
    assert a3.tolist() == v3["expected_a"]
    e2 = O.delete_edit_list(FIX["delete_these_2"]["list_b"], 13)
    assert len(e2) == 0


def test_duplicate_these_vectors():
    v = FIX["duplicate_these"]
    edit = list(range(10)) + v["list_a"]
    b, _ = O.apply_hidden_edit(v["list_b"], 10, edit, 100)
    c, _ = O.apply_hidden_edit(v["list_c"], 10, edit, 100)
    assert b.tolist() == v["expected_b"] and c.tolist() == v["expected_c"]
    assert np.array(v["hidden_values"])[edit].tolist() == v["expected_hidden"]


def test_apply_hidden_edit_matches_definition():
    """Rebuild = for each new slot h ascending, the positions pointing at its
    source slot, highest position first (sequential atomic_swap order)."""
    rng = np.random.default_rng(1)
    for _ in range(20):
        nh = int(rng.integers(1, 40))
        il = rng.integers(0, nh, int(rng.integers(0, 60)))
        ed = rng.integers(0, nh, int(rng.integers(1, 50)))
        want_i, want_e = [], []
        for h, src in enumerate(ed):
            for p in sorted(np.nonzero(il == src)[0], reverse=True):
                want_i.append(h)
                want_e.append(p)
        ni, ne = O.apply_hidden_edit(il, nh, ed, 10_000)
        assert ni.tolist() == want_i and ne.tolist() == want_e
        cap = max(1, len(want_i) // 2)
        ni2, _ = O.apply_hidden_edit(il, nh, ed, cap)
        assert ni2.tolist() == want_i[:cap]


# --------------------------------------------------------------------- hashing
def _naive_hash(p, mn, mx, res):
    out = []
    for q in p:
        g = []
        for d in range(3):
            x = np.float32(np.float32(q[d]) / np.float32(262144.0))
            v = np.float32((x - np.float32(mn[d])) / (np.float32(mx[d]) - np.float32(mn[d])) * np.float32(1 << res))
            g.append(0 if not v > 0 else (0xFFFFFFFF if v >= 4294967296.0 else int(v)))
        h = 0
        for i in range(res):
            for d in range(3):
                h += ((g[d] >> i) & 1) << (i * 3 + d)
        out.append(h & 0xFFFFFFFF)
    return np.array(out, np.uint32)


def test_position_hash_bit_interleave():
    rng = np.random.default_rng(2)
    p = np.zeros((300, 4), np.int32)
    p[:, :3] = (rng.uniform(-2, 18, (300, 3)) * 262144).astype(np.int32)
    mn, mx = (0.0, 0.0, 0.0), (16.0, 8.0, 4.0)
    for res in (1, 3, 5):
        assert np.array_equal(O.position_hash(p, mn, mx, res), _naive_hash(p, mn, mx, res))
    # known answers: cell (1,0,0) -> 1, (0,1,0) -> 2, (0,0,1) -> 4, (1,1,1) -> 7
    q = np.zeros((4, 4), np.int32)
    q[:, :3] = np.array([[1.5, 0.5, 0.5], [0.5, 1.5, 0.5], [0.5, 0.5, 1.5], [1.5, 1.5, 1.5]]) * 262144
    assert O.position_hash(q, (0, 0, 0), (16, 16, 16), 4).tolist() == [1, 2, 4, 7]


def test_position_code_known_answers():
    p = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [-1, 0, 0, 0]], np.int32)
    assert O.position_code(p, 0)[:3].tolist() == [1, 2, 4]
    # -1 sets every bit of x: bits 0,3,6,... of the 96-bit code
    full = [O.position_code(p[3:], s)[0] for s in range(3)]
    x_bits = sum(1 << b for b in range(0, 96, 3))
    assert sum(int(v) << (32 * s) for s, v in enumerate(full)) == x_bits


def test_find_value_ranges_equal_range():
    rng = np.random.default_rng(3)
    keys = np.sort(rng.integers(0, 100, 1000).astype(np.uint32))
    s, e = O.find_value_ranges(keys, 100)
    for c in range(100):
        lo, hi = np.searchsorted(keys, c, "left"), np.searchsorted(keys, c, "right")
        assert (s[c], e[c]) == ((lo, hi) if hi > lo else (0, 0))


def _cloud(n_side, seed, jitter=0.05):
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(*[np.arange(n_side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pw = 0.25 + 0.5 * g + rng.uniform(-jitter, jitter, g.shape)
    p = np.zeros((len(pw), 4), np.int32)
    p[:, :3] = np.trunc(pw * 262144.0).astype(np.int32)
    return p


@pytest.mark.parametrize("res,scale", [(2, 1.0), (3, 1.5), (4, 1.0)])
def test_green_equals_brute_force(res, scale):
    mn, mx = (0.0, 0.0, 0.0), (6.0, 6.0, 6.0)
    p = _cloud(10, res)
    h = O.position_hash(p, mn, mx, res)
    order = np.argsort(h, kind="stable")
    p, h = p[order], h[order]
    rng = np.random.default_rng(9)
    rad = rng.uniform(0.5, 1.2, len(p)).astype(np.float32)
    s, e = O.find_value_ranges(h, 1 << (3 * res))
    green = O.neighborhood_green(p, rad, s, e, scale, mn, mx, res, 1_000_000)
    brute = O.neighborhood_brute_force(p, rad, scale, 1_000_000)
    key = lambda a: a[np.lexsort((a[:, 1], a[:, 0]))]
    assert np.array_equal(key(green), key(brute))
    # directed: (i, j) in the list iff distance <= range_i (not range_j)
    assert len(green) > 0


# ---------------------------------------------------------------------- math
def test_exp_within_one_ulp_of_libm():
    xs = np.concatenate([np.linspace(-85.9, 0, 20001), np.linspace(0, 88, 4001)]).astype(np.float32)
    worst = 0
    for x in xs[::7]:
        got = np.float32(O.exp(float(x)))
        ref = np.float32(math.exp(float(x)))
        worst = max(worst, abs(int(got.view(np.int32)) - int(ref.view(np.int32))))
    assert worst <= 1
    assert O.exp(0.0) == 1.0 and O.exp(-90.0) == 0.0


def test_pow_correctly_rounded_mostly():
    rng = np.random.default_rng(4)
    xs = rng.uniform(1e-3, 50, 3000).astype(np.float32)
    for y in (3.0, 2.0 / 3.0, 1.5, 1.0 / 3.0, 2.0, 9.0):
        yy = float(np.float32(y))
        for x in xs[:500]:
            got = np.float32(O.pow(float(x), yy))
            ref = np.float32(float(x) ** yy)
            assert abs(int(got.view(np.int32)) - int(ref.view(np.int32))) <= 1
    assert O.pow(0.5, 3.0) == 0.125 and O.pow(2.0, 0.0) == 1.0 and O.pow(0.0, 3.0) == 0.0
    assert math.isnan(O.pow(-1.0, 2.0))


def test_gauss_kernel_closed_form():
    """W(0) = c for the Gaussian with c = 0.6/(kw/2)^3 (kernels.glsl:84-89,117),
    the gradient is -2 pi c^(2/3) r W and zero below 1e-4."""
    for kw in (0.5, 1.0, 4.0):
        c = 0.6 / (kw / 2) ** 3
        h0, g0 = O.kernel(1, [0, 0, 0], kw)
        assert abs(h0 - c) <= 2e-6 * c and np.all(g0 == 0)
        r = np.array([0.3, -0.2, 0.1], np.float32) * kw
        h, g = O.kernel(1, r, kw)
        ivd = c ** (2 / 3) * math.pi
        w = c * math.exp(-float(np.dot(r, r)) * ivd)
        assert abs(h - w) <= 1e-5 * w
        assert np.allclose(g, -2 * ivd * r * w, rtol=1e-5, atol=1e-9)


def test_kernels_vanish_outside_support_and_point_inward():
    r_out = [1.2, 0.0, 0.0]
    for kid in (0, 2, 3):
        h, _ = O.kernel(kid, r_out, 1.0)
        assert h == 0.0
    for kid in (0, 1, 2, 3, 4):
        _, g = O.kernel(kid, [0.3, 0.0, 0.0], 1.0)
        assert g[0] < 0 and g[1] == 0 and g[2] == 0  # gradient points back to the centre


# --------------------------------------------------------------- library interface
def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "apbf.h")).read()
    declared = set(re.findall(r"\b(apbf_[a-z0-9_]+)\s*\(", header))
    lib = ctypes.CDLL(os.path.join(ROOT, "apbf_amd", "libapbf_hip.so"))
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    from apbf_amd import _lib
    assert declared == set(_lib.exported_symbols())


def test_library_without_device_fails_loudly():
    from apbf_amd import _lib
    L = _lib.lib()
    if L.apbf_device_count() > 0:
        pytest.skip("a device is present")
    from apbf_amd import config as cfg
    from apbf_amd.scene import Scene
    p, _ = cfg.dam_break(n_side=4)
    with pytest.raises(_lib.ApbfError):
        Scene(p)


def test_power_of_two_division_is_exact_scaling():
    """The neighbour sums divide by inverse masses; for a power of two m whose
    reciprocal is a normal float, x / m equals x * (1 / m) bit for bit (the
    HIP kernels take that path when a whole wavefront qualifies)."""
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.standard_normal(4000).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, 4000),
        np.array([0.0, -0.0, np.inf, -np.inf, 1e-45, -1e-45, 1e-40, 3.4e38, -3.4e38, 1.17549435e-38],
                 np.float32),
        (rng.integers(1, 1 << 23, 500, dtype=np.uint32)).view(np.float32),  # subnormals
    ]).astype(np.float32)
    with np.errstate(over="ignore", under="ignore", invalid="ignore"):
        for e in range(1, 254):
            for sign in (0, 1):
                m = np.array([(sign << 31) | (e << 23)], np.uint32).view(np.float32)[0]
                rim = np.array([(sign << 31) | ((254 - e) << 23)], np.uint32).view(np.float32)[0]
                assert rim == np.float32(1.0) / m
                q = x / m
                s = x * rim
                assert np.array_equal(q.view(np.uint32), s.view(np.uint32)), (e, sign)
