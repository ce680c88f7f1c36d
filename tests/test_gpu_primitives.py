"""GPU parity of the list algorithms and single kernels against the reference's
self-test vectors (tests/golden/reference_selftest.json, from source/test.cpp)
and the oracle. Integer outputs must match bit-exactly."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_selftest.json")))


def T(a, device):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.uint64).astype(np.uint32).view(np.int32))).to(device)


def N(t):
    return t.cpu().numpy().view(np.uint32)


def test_prefix_sum_vectors(device):
    from apbf_amd import ops
    v = FIX["prefix_sum"]
    assert N(ops.prefix_sum(T(v["values"], device))).tolist() == v["expected"]
    # short list in a long buffer (test.cpp:269-282)
    buf = np.zeros(10000, np.uint32)
    buf[:7] = v["values"]
    out = N(ops.prefix_sum(T(buf, device), n=7))
    assert out[:7].tolist() == v["expected"]


@pytest.mark.parametrize("n", [1, 63, 64, 511, 512, 513, 2047, 2048, 2049, 1000, 512 * 512 + 1000])
def test_prefix_sum_running_sum(device, n):
    from apbf_amd import ops
    rng = np.random.default_rng(n)
    vals = rng.integers(0, 4, n).astype(np.uint32)
    out = N(ops.prefix_sum(T(vals, device)))
    assert np.array_equal(out, np.cumsum(vals, dtype=np.uint64).astype(np.uint32))


def test_prefix_sum_wraps(device):
    from apbf_amd import ops
    vals = np.full(5000, 0xF0000000, np.uint32)
    out = N(ops.prefix_sum(T(vals, device)))
    assert np.array_equal(out, np.cumsum(vals.astype(np.uint64)).astype(np.uint32))


@pytest.mark.parametrize("name", ["sort", "sort_small_values"])
def test_sort_vectors(device, name):
    from apbf_amd import ops
    v = FIX[name]
    n = len(v["keys"])
    k, vals = ops.radix_sort(T(v["keys"], device), T(np.arange(n), device))
    assert N(k).tolist() == v["expected_keys"]
    assert N(vals).tolist() == v["expected_values"]


@pytest.mark.parametrize("n,bits", [(512 * 512 + 123, 32), (512 * 512 + 1000, 4), (100000, 24), (3, 8),
                                    (2048, 12), (2049, 31)])
def test_sort_equals_stable_sort(device, n, bits):
    from apbf_amd import ops
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    idx = np.arange(n, dtype=np.uint32)
    k, v = ops.radix_sort(T(keys, device), T(idx, device))
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(N(k), keys[order])
    assert np.array_equal(N(v), idx[order])


def test_sort_upper_bound_matches_oracle(device):
    """Keys >= upper_bound are only sorted on the digits the reference visits."""
    from apbf_amd import ops
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 1 << 20, 30000, dtype=np.uint64).astype(np.uint32)
    idx = np.arange(len(keys), dtype=np.uint32)
    # upper bound 2^8: the oracle does 4-bit passes 0..8, the GPU one 8-bit pass
    # (while upper >> offset != 0); both give a stable sort on bits [0, 8) ... and
    # for keys < upper a full sort.
    small = keys & 0xFF
    k, v = ops.radix_sort(T(small, device), T(idx, device), upper_bound=1 << 8)
    ok, ov = O.stable_radix_sort(small, idx, 1 << 8)
    assert np.array_equal(N(k), ok) and np.array_equal(N(v), ov)


def _positions(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 4), np.int32)
    p[:, :3] = (rng.uniform(lo, hi, (n, 3)) * 262144.0).astype(np.int32)
    return p


def test_position_hash_matches_oracle(device):
    from apbf_amd import ops
    import torch
    p = _positions(50000, -3.0, 19.0, 1)  # includes positions outside [0, 16]
    mn, mx = (0.0, 0.0, 0.0), (16.0, 8.0, 12.0)
    for res in (1, 4, 6, 10):
        g = N(ops.position_hash(torch.from_numpy(p).to(device), mn, mx, res))
        assert np.array_equal(g, O.position_hash(p, mn, mx, res))


def test_position_code_matches_oracle(device):
    from apbf_amd import ops
    import torch
    p = _positions(20000, -100.0, 100.0, 2)
    for sec in range(3):
        g = N(ops.position_code(torch.from_numpy(p).to(device), sec))
        assert np.array_equal(g, O.position_code(p, sec))
    unit = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [-1, 0, 0, 0]], np.int32)
    g0 = N(ops.position_code(torch.from_numpy(unit).to(device), 0))
    assert g0[:3].tolist() == [1, 2, 4]


def test_find_value_ranges_matches_oracle(device):
    from apbf_amd import ops
    rng = np.random.default_rng(3)
    keys = np.sort(rng.integers(0, 4096, 20000).astype(np.uint32))
    cs, ce = ops.find_value_ranges(T(keys, device), 4096)
    rs, re_ = O.find_value_ranges(keys, 4096)
    assert np.array_equal(N(cs), rs) and np.array_equal(N(ce), re_)
    # every cell range is the equal_range of its key
    for c in rng.integers(0, 4096, 50):
        lo, hi = np.searchsorted(keys, c, "left"), np.searchsorted(keys, c, "right")
        if hi > lo:
            assert (N(cs)[c], N(ce)[c]) == (lo, hi)


def _sorted_cloud(n, seed, mn, mx, res, spacing=0.5):
    rng = np.random.default_rng(seed)
    side = int(round(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pw = 0.25 + spacing * g + rng.uniform(-0.05, 0.05, g.shape)
    p = np.zeros((len(pw), 4), np.int32)
    p[:, :3] = np.trunc(pw * 262144.0).astype(np.int32)
    h = O.position_hash(p, mn, mx, res)
    order = np.argsort(h, kind="stable")
    return p[order], h[order]


@pytest.mark.parametrize("res,scale", [(3, 1.0), (4, 1.5), (2, 1.0)])
def test_neighborhood_green_matches_oracle(device, res, scale):
    from apbf_amd import ops
    import torch
    mn, mx = (0.0, 0.0, 0.0), (8.0, 8.0, 8.0)
    p, h = _sorted_cloud(4096, res, mn, mx, res)
    rng = np.random.default_rng(5)
    rad = rng.uniform(0.6, 1.4, len(p)).astype(np.float32)
    rs, re_ = O.find_value_ranges(h, 1 << (3 * res))
    want = O.neighborhood_green(p, rad, rs, re_, scale, mn, mx, res, 10_000_000)
    row, col, npairs = ops.neighborhood_green(torch.from_numpy(p).to(device), torch.from_numpy(rad).to(device),
                                              T(rs, device), T(re_, device), scale, mn, mx, res, 10_000_000)
    npairs = int(N(npairs)[0])
    assert npairs == len(want)
    r = N(row)
    got = np.stack([np.repeat(np.arange(len(p)), np.diff(r)), N(col)[:npairs]], 1)
    assert np.array_equal(got, want)
    # and the set equals the brute-force definition (neighborhood_brute_force.comp:43-48)
    brute = O.neighborhood_brute_force(p, rad, scale, 10_000_000)
    a = got[np.lexsort((got[:, 1], got[:, 0]))]
    b = brute[np.lexsort((brute[:, 1], brute[:, 0]))]
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n,cap", [(1000, 10_000_000), (1331, 4000), (0, 100)])
def test_neighborhood_brute_force_matches_oracle(device, n, cap):
    """neighborhood_brute_force.comp:31-50 in index order, with the clamp."""
    from apbf_amd import ops
    import torch
    rng = np.random.default_rng(n)
    p = np.zeros((n, 4), np.int32)
    p[:, :3] = np.trunc(rng.uniform(0.0, 6.0, (n, 3)) * 262144.0).astype(np.int32)
    rad = rng.uniform(0.5, 1.5, n).astype(np.float32)
    want = O.neighborhood_brute_force(p, rad, 1.25, cap)
    row, col, npairs = ops.neighborhood_brute_force(torch.from_numpy(p).to(device), torch.from_numpy(rad).to(device),
                                                    1.25, cap)
    npairs = int(N(npairs)[0])
    assert npairs == len(want)
    r = np.minimum(N(row), cap)
    got = np.stack([np.repeat(np.arange(n), np.diff(r)), N(col)[:npairs]], 1) if n else np.zeros((0, 2))
    assert np.array_equal(got, want)


def test_neighborhood_green_capacity_clamp(device):
    from apbf_amd import ops
    import torch
    mn, mx, res = (0.0, 0.0, 0.0), (8.0, 8.0, 8.0), 3
    p, h = _sorted_cloud(1000, 9, mn, mx, res)
    rad = np.full(len(p), 1.0, np.float32)
    rs, re_ = O.find_value_ranges(h, 1 << (3 * res))
    cap = 5000
    want = O.neighborhood_green(p, rad, rs, re_, 1.0, mn, mx, res, cap)
    row, col, npairs = ops.neighborhood_green(torch.from_numpy(p).to(device), torch.from_numpy(rad).to(device),
                                              T(rs, device), T(re_, device), 1.0, mn, mx, res, cap)
    assert int(N(npairs)[0]) == cap == len(want)
    r = np.minimum(N(row), cap)
    got = np.stack([np.repeat(np.arange(len(p)), np.diff(r)), N(col)[:cap]], 1)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("name", ["indexed_list_apply_hidden_edit_1", "indexed_list_apply_hidden_edit_2",
                                  "indexed_list_apply_hidden_edit_3"])
def test_apply_hidden_edit_vectors(device, name):
    from apbf_amd import ops
    v = FIX[name]
    for lst, exp in (("list_a", "expected_a"), ("list_b", "expected_b")):
        ni, ne = ops.apply_hidden_edit(T(v[lst], device), v["hidden_length"], T(v["edit"], device), 100)
        assert N(ni).tolist() == v[exp]


def test_apply_hidden_edit_matches_oracle(device):
    from apbf_amd import ops
    rng = np.random.default_rng(11)
    for trial in range(5):
        nh = int(rng.integers(1, 3000))
        il = rng.integers(0, nh, int(rng.integers(0, 4000))).astype(np.uint32)
        ed = rng.integers(0, nh, int(rng.integers(1, 3000))).astype(np.uint32)
        mx = int(rng.integers(1, 6000))
        ni, ne = ops.apply_hidden_edit(T(il, device), nh, T(ed, device), mx)
        oi, oe = O.apply_hidden_edit(il, nh, ed, mx)
        assert np.array_equal(N(ni), oi) and np.array_equal(N(ne), oe)


def test_delete_edit_list(device):
    from apbf_amd import ops
    v = FIX["delete_these_1"]
    edit = N(ops.delete_edit_list(T(v["list_b"], device), len(v["hidden_values"])))
    hidden = np.array(v["hidden_values"])[edit]
    assert hidden.tolist() == v["expected_hidden"]
    rng = np.random.default_rng(4)
    d = rng.choice(5000, 1234, replace=False).astype(np.uint32)
    assert np.array_equal(N(ops.delete_edit_list(T(d, device), 5000)), O.delete_edit_list(d, 5000))


def test_copy_scattered_read(device):
    from apbf_amd import ops
    v = FIX["gpu_list_apply_edit"]
    out = N(ops.copy_scattered_read(T(v["list"], device), T(v["edit"], device), 4))
    assert out.tolist() == v["expected"]
    vec = np.array(FIX["gpu_list_concatenation_2"]["expected"], np.float32)
    out3 = ops.copy_scattered_read(T(vec.view(np.uint32).ravel(), device), T([4, 0, 2], device), 12)
    assert np.array_equal(N(out3).view(np.float32).reshape(3, 3), vec[[4, 0, 2]])


def _world_dist(pa, pb):
    """Float32 distance between fixed-point positions the way the walk forms it:
    world coordinates first, then difference, dot product and square root."""
    wa = pa[:, :3].astype(np.float32) / np.float32(262144.0)
    wb = pb[:, :3].astype(np.float32) / np.float32(262144.0)
    d = wa - wb
    d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    return np.sqrt(d2.astype(np.float32))


def test_neighborhood_green_boundary_ranges(device):
    """Ranges equal to a neighbour's rounded distance, and one float step either
    side: the walk's squared-distance bound must keep exactly the neighbours
    whose rounded distance is <= range."""
    from apbf_amd import ops
    import torch
    mn, mx, res = (0.0, 0.0, 0.0), (8.0, 8.0, 8.0), 3
    p, h = _sorted_cloud(3375, 21, mn, mx, res)
    rng = np.random.default_rng(8)
    partner = np.clip(np.arange(len(p)) + rng.integers(-40, 40, len(p)), 0, len(p) - 1)
    partner[partner == np.arange(len(p))] = (partner[partner == np.arange(len(p))] + 1) % len(p)
    rad = _world_dist(p, p[partner]).astype(np.float32)
    step = rng.integers(-1, 2, len(p))
    rad = np.where(step < 0, np.nextafter(rad, np.float32(0)), np.where(step > 0, np.nextafter(rad, np.float32(np.inf)), rad)).astype(np.float32)
    rad[:5] = np.array([0.0, -1.0, np.nan, np.inf, 1e-30], np.float32)
    rs, re_ = O.find_value_ranges(h, 1 << (3 * res))
    cap = 20_000_000
    want = O.neighborhood_green(p, rad, rs, re_, 1.0, mn, mx, res, cap)
    row, col, npairs = ops.neighborhood_green(torch.from_numpy(p).to(device), torch.from_numpy(rad).to(device),
                                              T(rs, device), T(re_, device), 1.0, mn, mx, res, cap)
    npairs = int(N(npairs)[0])
    assert npairs == len(want)
    got = np.stack([np.repeat(np.arange(len(p)), np.diff(N(row))), N(col)[:npairs]], 1)
    assert np.array_equal(got, want)
    # the chosen partner sits exactly on the boundary for step 0 (kept) and -1 (dropped)
    pairs = set(map(tuple, got.tolist()))
    for i in range(5, len(p)):
        if step[i] == 0:
            assert (i, partner[i]) in pairs
        elif step[i] < 0:
            assert (i, partner[i]) not in pairs
