"""The gpu_list / indexed_list operations of the reference's self-test
(source/test.cpp:47-115, 486-531; vectors in tests/golden/reference_selftest.json)
through the C ABI list operators: append with its capacity clamp
(gpu_list::operator+=, append_list.comp:21-25), increase_length
(write_increasing_sequence.comp), duplicate_these (indexed_list.h:141-152:
iota + append + write_increasing_sequence_from_to + the hidden gather and the
owners' index rebuild)."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_selftest.json")))


def T(a, device, width=None):
    import torch
    arr = np.asarray(a)
    if arr.dtype.kind == "f":
        arr = arr.astype(np.float32).view(np.uint32)
    return torch.from_numpy(np.ascontiguousarray(arr.astype(np.uint64).astype(np.uint32).view(np.int32).ravel())).to(device)


def N(t):
    return t.cpu().numpy().view(np.uint32)


def test_concatenation_1(device):
    from apbf_amd import ops
    import torch
    v = FIX["gpu_list_concatenation_1"]
    buf = torch.zeros(len(v["a"]) + len(v["b"]), dtype=torch.int32, device=device)
    buf[: len(v["a"])] = T(v["a"], device)
    out, n = ops.append_list(buf, len(v["a"]), T(v["b"], device), len(v["b"]), 4)
    assert n == len(v["expected"]) and N(out)[:n].tolist() == v["expected"]


def test_concatenation_2_vec3(device):
    """listB = listA + listB + listC with 12-byte elements (test.cpp:59-73)."""
    from apbf_amd import ops
    import torch
    v = FIX["gpu_list_concatenation_2"]
    f = lambda rows: np.asarray(rows, np.float32).view(np.uint32).ravel()
    cap = len(v["expected"])
    buf = torch.zeros(3 * cap, dtype=torch.int32, device=device)
    a = f(v["a"])
    buf[: len(a)] = T(a, device)
    out, n = ops.append_list(buf, len(v["a"]), T(f(v["b"]), device), len(v["b"]), 12)
    out, n = ops.append_list(out, n, T(f(v["c"]), device), len(v["c"]), 12)
    assert n == cap
    got = N(out).view(np.float32).reshape(cap, 3)
    assert np.array_equal(got, np.asarray(v["expected"], np.float32))


def test_append_empty(device):
    from apbf_amd import ops
    import torch
    v = FIX["gpu_list_append_empty"]
    buf = torch.zeros(len(v["a"]) + 5, dtype=torch.int32, device=device)
    buf[: len(v["a"])] = T(v["a"], device)
    out, n = ops.append_list(buf, len(v["a"]), torch.zeros(0, dtype=torch.int32, device=device), 0, 4)
    assert n == v["expected_length"] and N(out)[:n].tolist() == v["expected"]


def test_append_clamps_at_capacity(device):
    """append_list.comp:21-25: the new length is min(sum, capacity) and
    elements beyond the capacity are dropped; the words before stay."""
    from apbf_amd import ops
    import torch
    backing = torch.full((10,), 99, dtype=torch.int32, device=device)
    backing[:4] = T([3, 64, 12683, 7], device)
    buf = backing[:6]  # a list of capacity 6 inside a larger allocation
    out, n = ops.append_list(buf, 4, T([1, 2, 3, 4, 5], device), 5, 4)
    assert n == 6 and N(out).tolist() == [3, 64, 12683, 7, 1, 2]
    assert N(backing[6:]).tolist() == [99] * 4  # nothing written past the capacity
    out, n = ops.append_list(buf, 6, T([8], device), 1, 4)  # full list: nothing written
    assert n == 6 and N(backing).tolist() == [3, 64, 12683, 7, 1, 2, 99, 99, 99, 99]


def test_increase_length(device):
    """indexed_list::increase_length (test.cpp:107-115): indices [0, 1, 2] of a
    hidden list of capacity 5; a second call continues at 3 and stops at the
    hidden capacity."""
    from apbf_amd import ops
    v = FIX["indexed_list_write_increasing_sequence"]
    seq, new_min, n = ops.write_increasing_sequence(v["added"], 0, v["hidden_capacity"], v["added"], device)
    assert N(seq).tolist() == v["expected"] and new_min == 3 and n == 3
    seq, new_min, n = ops.write_increasing_sequence(4, new_min, v["hidden_capacity"], 4, device)
    assert N(seq).tolist() == [3, 4] and new_min == 5 and n == 2


@pytest.mark.parametrize("name", ["duplicate_these", "duplicate_these_empty"])
def test_duplicate_these(device, name):
    """indexed_list::duplicate_these (indexed_list.h:141-152) composed from the
    list operators: edit = [0, hidden) + list_a; the hidden list is gathered
    by the edit list; list_a's copies get [hidden, edit length); every other
    owner's index list is rebuilt by the hidden edit (indexed_list.h:289-308)."""
    from apbf_amd import ops
    import torch
    v = FIX[name]
    hidden = np.asarray(v["hidden_values"], np.uint32)
    nh = len(hidden)
    la = v["list_a"]
    cap = nh + len(la)
    edit = torch.zeros(cap, dtype=torch.int32, device=device)
    edit[:nh] = ops.write_sequence(nh, 0, 1, device)
    edit, ne = ops.append_list(edit, nh, T(la, device) if la else torch.zeros(0, dtype=torch.int32, device=device),
                               len(la), 4)
    new_ids, k = ops.write_increasing_sequence_from_to(len(la), nh, ne, device)
    assert k == len(la) and N(new_ids).tolist() == list(range(nh, ne))
    new_hidden = N(ops.copy_scattered_read(T(hidden, device), edit[:ne], 4))
    assert new_hidden.tolist() == v["expected_hidden"] and ne == v["expected_hidden_length"]
    ni, _ = ops.apply_hidden_edit(T(v["list_b"], device), nh, edit[:ne], 1000)
    assert N(ni).tolist() == v["expected_b"]
    if "list_c" in v:
        ni, _ = ops.apply_hidden_edit(T(v["list_c"], device), nh, edit[:ne], 1000)
        assert N(ni).tolist() == v["expected_c"]
