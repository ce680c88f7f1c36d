"""List algorithms and single kernels on device tensors (torch is used only to
hold device memory and streams; the work is done by libapbf_hip.so).

Each function mirrors a reference interface: algorithms::sort / prefix_sum
(source/algorithms.h), shader_provider::calculate_position_hash,
calculate_position_code, find_value_ranges, neighborhood_green,
copy_scattered_read (source/shader_provider.h), and the index-list edits of
source/indexed_list.h (apply_hidden_edit, delete_these).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import check


def _p(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


_keep: list = []


def _len(n: int, device) -> torch.Tensor:
    """Device-resident u32 length. The tensor is kept alive until the next
    synchronisation (the library reads it asynchronously)."""
    t = torch.tensor([n], dtype=torch.int32, device=device)
    _keep.append(t)
    return t


def _sync():
    torch.cuda.current_stream().synchronize()
    _keep.clear()


def _ws(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 16), dtype=torch.uint8, device=device)


def _u32(t):  # torch has no uint32 arithmetic; int32 storage with the same bits
    assert t.dtype == torch.int32 and t.is_cuda and t.is_contiguous()
    return t


def prefix_sum(values: torch.Tensor, n: int | None = None) -> torch.Tensor:
    """Inclusive prefix sum (algorithms::prefix_sum) of the first n values, in place."""
    L = _lib.lib()
    _u32(values)
    n = values.numel() if n is None else n
    cap = values.numel()
    ws = _ws(L.apbf_prefix_sum_workspace_size(cap), values.device)
    check(L.apbf_prefix_sum(_p(values), _p(_len(n, values.device)), cap, _p(ws), ws.numel(), _stream()),
          "apbf_prefix_sum")
    _sync()
    return values


def radix_sort(keys: torch.Tensor, values: torch.Tensor, upper_bound: int = 0xFFFFFFFF,
               n: int | None = None):
    """Stable sort of (key, value) pairs by key (algorithms::sort)."""
    L = _lib.lib()
    _u32(keys)
    _u32(values)
    cap = keys.numel()
    n = cap if n is None else n
    ok = torch.empty_like(keys)
    ov = torch.empty_like(values)
    ws = _ws(L.apbf_radix_sort_workspace_size(cap), keys.device)
    check(L.apbf_radix_sort(_p(keys), _p(values), _p(_len(n, keys.device)), cap, upper_bound & 0xFFFFFFFF,
                            _p(ok), _p(ov), _p(ws), ws.numel(), _stream()), "apbf_radix_sort")
    _sync()
    return ok, ov


def position_hash(position: torch.Tensor, mn, mx, res: int, dims: int = 3) -> torch.Tensor:
    L = _lib.lib()
    n = position.shape[0]
    out = torch.empty(n, dtype=torch.int32, device=position.device)
    fmn = (C.c_float * 3)(*mn)
    fmx = (C.c_float * 3)(*mx)
    check(L.apbf_calculate_position_hash(_p(position), _p(_len(n, position.device)), n, fmn, fmx, res, dims,
                                         _p(out), _stream()), "apbf_calculate_position_hash")
    _sync()
    return out


def position_code(position: torch.Tensor, section: int) -> torch.Tensor:
    L = _lib.lib()
    n = position.shape[0]
    out = torch.empty(n, dtype=torch.int32, device=position.device)
    check(L.apbf_calculate_position_code(_p(position), _p(_len(n, position.device)), n, section, _p(out),
                                         _stream()), "apbf_calculate_position_code")
    _sync()
    return out


def find_value_ranges(sorted_keys: torch.Tensor, n_cells: int):
    L = _lib.lib()
    n = sorted_keys.numel()
    cs = torch.empty(n_cells, dtype=torch.int32, device=sorted_keys.device)
    ce = torch.empty(n_cells, dtype=torch.int32, device=sorted_keys.device)
    check(L.apbf_find_value_ranges(_p(sorted_keys), _p(_len(n, sorted_keys.device)), n, n_cells, _p(cs),
                                   _p(ce), _stream()), "apbf_find_value_ranges")
    _sync()
    return cs, ce


def neighborhood_green(position, rng, cell_start, cell_end, range_scale, mn, mx, res, pair_cap, dims=3):
    """CSR neighbour lists (row offsets, neighbour ids, pair count)."""
    L = _lib.lib()
    n = position.shape[0]
    dev = position.device
    row = torch.empty(n + 1, dtype=torch.int32, device=dev)
    col = torch.empty(max(pair_cap, 1), dtype=torch.int32, device=dev)
    npairs = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = _ws(L.apbf_neighborhood_green_workspace_size(n), dev)
    fmn = (C.c_float * 3)(*mn)
    fmx = (C.c_float * 3)(*mx)
    check(L.apbf_neighborhood_green(_p(position), _p(rng), _p(_len(n, dev)), n, float(range_scale), fmn, fmx,
                                    res, dims, _p(cell_start), _p(cell_end), _p(row), _p(col), pair_cap, _p(npairs),
                                    _p(ws), ws.numel(), _stream()), "apbf_neighborhood_green")
    _sync()
    return row, col, npairs


def neighborhood_brute_force(position, rng, range_scale, pair_cap):
    """CSR neighbour lists of the brute-force search (row offsets, neighbour ids, pair count)."""
    L = _lib.lib()
    n = position.shape[0]
    dev = position.device
    row = torch.empty(n + 1, dtype=torch.int32, device=dev)
    col = torch.empty(max(pair_cap, 1), dtype=torch.int32, device=dev)
    npairs = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = _ws(L.apbf_neighborhood_green_workspace_size(n), dev)
    check(L.apbf_neighborhood_brute_force(_p(position), _p(rng), _p(_len(n, dev)), n, float(range_scale), _p(row),
                                          _p(col), pair_cap, _p(npairs), _p(ws), ws.numel(), _stream()),
          "apbf_neighborhood_brute_force")
    _sync()
    return row, col, npairs


def copy_scattered_read(src: torch.Tensor, edit: torch.Tensor, stride: int, out_len: int | None = None):
    L = _lib.lib()
    n = edit.numel() if out_len is None else out_len
    dst = torch.empty(max(n * stride // 4, 1), dtype=torch.int32, device=src.device)
    check(L.apbf_copy_scattered_read(_p(src), _p(dst), _p(edit), _p(_len(n, src.device)), edit.numel(),
                                     stride, _stream()), "apbf_copy_scattered_read")
    _sync()
    return dst[: n * stride // 4]


def apply_hidden_edit(index_list: torch.Tensor, n_hidden: int, edit: torch.Tensor, max_new: int):
    L = _lib.lib()
    dev = edit.device
    m, e = index_list.numel(), edit.numel()
    cap_m, cap_e = max(m, 1), max(e, 1)
    il = index_list if m else torch.zeros(1, dtype=torch.int32, device=dev)
    ed = edit if e else torch.zeros(1, dtype=torch.int32, device=dev)
    ni = torch.zeros(max(max_new, 1), dtype=torch.int32, device=dev)
    ne = torch.zeros(max(max_new, 1), dtype=torch.int32, device=dev)
    nl = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = _ws(L.apbf_hidden_edit_workspace_size(cap_m, max(n_hidden, 1), cap_e), dev)
    check(L.apbf_apply_hidden_edit(_p(il), _p(_len(m, dev)), cap_m, max(n_hidden, 1), _p(ed), _p(_len(e, dev)),
                                   cap_e, max_new, _p(ni), _p(ne), _p(nl), _p(ws), ws.numel(), _stream()),
          "apbf_apply_hidden_edit")
    k = int(nl.item())
    _sync()
    return ni[:k], ne[:k]


def delete_edit_list(deleted: torch.Tensor, n_hidden: int):
    L = _lib.lib()
    dev = deleted.device
    d = deleted if deleted.numel() else torch.zeros(1, dtype=torch.int32, device=dev)
    cap = max(n_hidden, 1)
    edit = torch.zeros(cap, dtype=torch.int32, device=dev)
    el = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = _ws(L.apbf_delete_edit_list_workspace_size(cap), dev)
    check(L.apbf_delete_edit_list(_p(d), _p(_len(deleted.numel(), dev)), max(d.numel(), 1),
                                  _p(_len(n_hidden, dev)), cap, _p(edit), _p(el), _p(ws), ws.numel(), _stream()),
          "apbf_delete_edit_list")
    _sync()
    return edit[: int(el.item())]


def append_list(target: torch.Tensor, n_target: int, appending: torch.Tensor, n_append: int, stride: int):
    """gpu_list::operator+= (shader_provider::append_list): `target` is the
    whole buffer (its capacity is target.numel() * 4 / stride); returns the
    buffer and the clamped new length."""
    L = _lib.lib()
    dev = target.device
    cap = target.numel() * 4 // stride
    a = appending if appending.numel() else torch.zeros(1, dtype=torch.int32, device=dev)
    new_len = torch.zeros(1, dtype=torch.int32, device=dev)
    check(L.apbf_append_list(_p(target), cap, _p(_len(n_target, dev)), _p(a), _p(_len(n_append, dev)), stride,
                             _p(new_len), appending.numel() * 4 // stride, _stream()), "apbf_append_list")
    _sync()
    return target, int(new_len.item())


def write_increasing_sequence(target_capacity: int, sequence_min: int, value_upper_bound: int, sequence_length: int,
                              device):
    """shader_provider::write_increasing_sequence: (list, new sequence min, new length)."""
    L = _lib.lib()
    out = torch.zeros(max(target_capacity, 1), dtype=torch.int32, device=device)
    new_min = torch.zeros(1, dtype=torch.int32, device=device)
    new_len = torch.zeros(1, dtype=torch.int32, device=device)
    check(L.apbf_write_increasing_sequence(_p(out), target_capacity, _p(_len(sequence_min, device)),
                                           value_upper_bound & 0xFFFFFFFF, sequence_length, _p(new_min),
                                           _p(new_len), _stream()), "apbf_write_increasing_sequence")
    _sync()
    n = int(new_len.item())
    return out[:n], int(new_min.item()), n


def write_increasing_sequence_from_to(out_capacity: int, start: int, stop: int, device):
    """shader_provider::write_increasing_sequence_from_to: [start, stop) and its length."""
    L = _lib.lib()
    out = torch.zeros(max(out_capacity, 1), dtype=torch.int32, device=device)
    n = torch.zeros(1, dtype=torch.int32, device=device)
    check(L.apbf_write_increasing_sequence_from_to(_p(out), out_capacity, _p(n), _p(_len(start, device)),
                                                   _p(_len(stop, device)), _stream()),
          "apbf_write_increasing_sequence_from_to")
    _sync()
    k = int(n.item())
    return out[:k], k


def write_sequence(n: int, start: int, step: int, device, capacity: int | None = None):
    """shader_provider::write_sequence: start + i * step for i < n."""
    L = _lib.lib()
    cap = n if capacity is None else capacity
    out = torch.zeros(max(cap, 1), dtype=torch.int32, device=device)
    check(L.apbf_write_sequence(_p(out), _p(_len(n, device)), cap, start & 0xFFFFFFFF, step & 0xFFFFFFFF, _stream()),
          "apbf_write_sequence")
    _sync()
    return out[:n]
