"""apbf_radix_sort_noalloc (the sort with its workspace in a fixed device
array of the library) against apbf_radix_sort (caller workspace), bit for bit,
through the C ABI: two calls in a row, the second on a smaller input, so the
second call runs on the array the first one left behind."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sort_with_workspace(L, torch, keys, vals, n, cap, ub, stream):
    ok = torch.empty(cap, dtype=torch.int32, device=keys.device)
    ov = torch.empty(cap, dtype=torch.int32, device=keys.device)
    ws = torch.empty(L.apbf_radix_sort_workspace_size(cap), dtype=torch.uint8, device=keys.device)
    d_len = torch.tensor([n], dtype=torch.int32, device=keys.device)
    st = L.apbf_radix_sort(C.c_void_p(keys.data_ptr()), C.c_void_p(vals.data_ptr()), C.c_void_p(d_len.data_ptr()),
                           cap, ub, C.c_void_p(ok.data_ptr()), C.c_void_p(ov.data_ptr()),
                           C.c_void_p(ws.data_ptr()), ws.numel(), stream)
    assert st == 0
    torch.cuda.synchronize()
    return ok[:n].cpu().numpy(), ov[:n].cpu().numpy()


def _sort_noalloc(L, torch, keys, vals, n, cap, ub, stream):
    ok = torch.empty(cap, dtype=torch.int32, device=keys.device)
    ov = torch.empty(cap, dtype=torch.int32, device=keys.device)
    d_len = torch.tensor([n], dtype=torch.int32, device=keys.device)
    st = L.apbf_radix_sort_noalloc(C.c_void_p(keys.data_ptr()), C.c_void_p(vals.data_ptr()),
                                   C.c_void_p(d_len.data_ptr()), cap, ub, C.c_void_p(ok.data_ptr()),
                                   C.c_void_p(ov.data_ptr()), stream)
    assert st == 0
    torch.cuda.synchronize()
    return ok[:n].cpu().numpy(), ov[:n].cpu().numpy()


def test_noalloc_sort_matches_workspace_sort(device):
    import torch
    from apbf_amd import _lib
    L = _lib.lib()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(7)
    # first call: the largest input the array takes, full 32-bit keys (4 passes);
    # second call: a smaller input with 24-bit keys (3 passes), length below capacity
    cases = [(1 << 20, 1 << 20, 0xFFFFFFFF), (300_001, 400_000, 1 << 24)]
    for n, cap, ub in cases:
        print(f"noalloc sort n={n} cap={cap} upper_bound={ub:#x}", flush=True)
        hk = rng.integers(0, min(ub, 0xFFFFFFFF), cap, dtype=np.uint64).astype(np.uint32)
        hk[: n // 3] = hk[n // 3: 2 * (n // 3)]  # many equal keys: stability shows in the values
        keys = torch.from_numpy(hk.view(np.int32)).to(device)
        vals = torch.arange(cap, dtype=torch.int32, device=device)
        ref_k, ref_v = _sort_with_workspace(L, torch, keys, vals, n, cap, ub, stream)
        got_k, got_v = _sort_noalloc(L, torch, keys, vals, n, cap, ub, stream)
        assert np.array_equal(got_k, ref_k)
        assert np.array_equal(got_v, ref_v)
        order = np.argsort(hk[:n], kind="stable")
        assert np.array_equal(ref_v, order.astype(np.int32))


def test_noalloc_sort_rejects_capacity_beyond_the_array(device):
    import torch
    from apbf_amd import _lib
    L = _lib.lib()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    big = (1 << 20) + 1
    keys = torch.zeros(big, dtype=torch.int32, device=device)
    vals = torch.zeros(big, dtype=torch.int32, device=device)
    ok = torch.empty_like(keys)
    ov = torch.empty_like(vals)
    d_len = torch.tensor([big], dtype=torch.int32, device=device)
    st = L.apbf_radix_sort_noalloc(C.c_void_p(keys.data_ptr()), C.c_void_p(vals.data_ptr()),
                                   C.c_void_p(d_len.data_ptr()), big, 0xFFFFFFFF, C.c_void_p(ok.data_ptr()),
                                   C.c_void_p(ov.data_ptr()), stream)
    assert st == -3  # APBF_ERR_CAPACITY, nothing launched
