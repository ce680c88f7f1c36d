// C ABI of the list algorithms and single kernels (include/apbf.h).
#include <algorithm>

#include "plain_kernels.h"

namespace apbf {


__global__ __launch_bounds__(BLOCK) void k_count_values(const uint32_t* __restrict__ vals,
                                                        const uint32_t* __restrict__ d_len, uint32_t limit,
                                                        uint32_t* __restrict__ cnt) {
    const uint32_t n = *d_len;
    GRID_STRIDE(p, n) {
        uint32_t b = vals[p];
        if (b < limit) atomicAdd(&cnt[b], 1u);
    }
}
__global__ void k_set_plus_one(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src + 1u;
}
__global__ __launch_bounds__(BLOCK) void k_emit_counts(const uint32_t* __restrict__ edit,
                                                       const uint32_t* __restrict__ d_edit_len,
                                                       uint32_t limit, const uint32_t* __restrict__ cnt,
                                                       uint32_t* __restrict__ k) {
    const uint32_t e = *d_edit_len;
    GRID_STRIDE(h, e) {
        uint32_t b = edit[h];
        k[h] = b < limit ? cnt[b] : 0u;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) k[e] = 0;
}
// generate_new_index_and_edit_list.comp semantics in the sequential schedule:
// new slot h (ascending) emits the positions of its source slot, the most
// recently linked (highest position) first.
__global__ __launch_bounds__(BLOCK) void k_emit_entries(const uint32_t* __restrict__ edit,
                                                        const uint32_t* __restrict__ d_edit_len,
                                                        uint32_t limit, const uint32_t* __restrict__ cnt,
                                                        const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ sorted_pos,
                                                        const uint32_t* __restrict__ out_off,
                                                        uint32_t max_new, uint32_t* __restrict__ new_index,
                                                        uint32_t* __restrict__ new_edit) {
    const uint32_t e = *d_edit_len;
    GRID_STRIDE(h, e) {
        uint32_t b = edit[h];
        if (b >= limit) continue;
        uint32_t c = cnt[b], base = off[b], o = out_off[h];
        for (uint32_t q = 0; q < c; ++q) {
            uint32_t w = o + q;
            if (w >= max_new) break;
            new_index[w] = h;
            new_edit[w] = sorted_pos[base + c - 1u - q];
        }
    }
}
__global__ void k_set_value(uint32_t* __restrict__ dst, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = v;
}
__global__ void k_clamp_len(const uint32_t* __restrict__ total, uint32_t max_new, uint32_t* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = min(*total, max_new);
}
__global__ __launch_bounds__(BLOCK) void k_flag_deleted(const uint32_t* __restrict__ deleted,
                                                        const uint32_t* __restrict__ d_len,
                                                        const uint32_t* __restrict__ d_hidden_len,
                                                        uint32_t* __restrict__ flags) {
    const uint32_t n = *d_len, h = *d_hidden_len;
    GRID_STRIDE(p, n) {
        uint32_t d = deleted[p];
        if (d < h) flags[d] = 0u;
    }
}
__global__ __launch_bounds__(BLOCK) void k_scatter_kept(const uint32_t* __restrict__ flags,
                                                        const uint32_t* __restrict__ scan,
                                                        const uint32_t* __restrict__ d_len,
                                                        uint32_t* __restrict__ edit) {
    const uint32_t n = *d_len;
    GRID_STRIDE(i, n) if (flags[i]) edit[scan[i]] = i;
}

// append_list.comp:21-25: the new length is clamped to the target's capacity,
// and words that would land beyond it are dropped.
__global__ __launch_bounds__(BLOCK) void k_append(uint32_t* __restrict__ target, const uint32_t* __restrict__ app,
                                                  const uint32_t* __restrict__ d_tlen,
                                                  const uint32_t* __restrict__ d_alen, uint32_t target_cap,
                                                  uint32_t words, uint32_t* __restrict__ d_new_len) {
    const uint32_t tl = *d_tlen, al = *d_alen;
    const uint64_t room = tl < target_cap ? (uint64_t)(target_cap - tl) * words : 0ull;
    const uint64_t n = min((uint64_t)al * words, room);
    const uint64_t base = (uint64_t)tl * words;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        target[base + i] = app[i];
}
__global__ void k_append_len(const uint32_t* __restrict__ d_tlen, const uint32_t* __restrict__ d_alen,
                             uint32_t target_cap, uint32_t* __restrict__ d_new_len) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint64_t t = (uint64_t)*d_tlen + (uint64_t)*d_alen;
        *d_new_len = (uint32_t)min(t, (uint64_t)target_cap);
    }
}
// write_increasing_sequence.comp:23-33
__global__ __launch_bounds__(BLOCK) void k_increasing_sequence(uint32_t* __restrict__ target, uint32_t target_cap,
                                                               const uint32_t* __restrict__ d_min, uint32_t upper,
                                                               uint32_t seq_len) {
    const uint32_t mn = *d_min;
    const uint32_t len = min(min(target_cap, seq_len), upper > mn ? upper - mn : 0u);
    GRID_STRIDE(i, len) target[i] = mn + i;
}
__global__ void k_increasing_sequence_len(uint32_t target_cap, const uint32_t* __restrict__ d_min, uint32_t upper,
                                          uint32_t seq_len, uint32_t* __restrict__ d_new_min,
                                          uint32_t* __restrict__ d_new_len) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint32_t mn = *d_min;
        const uint32_t len = min(min(target_cap, seq_len), upper > mn ? upper - mn : 0u);
        *d_new_min = mn + len;
        *d_new_len = len;
    }
}
// write_increasing_sequence_from_to.comp:19-27 (the output capacity bounds the writes)
__global__ __launch_bounds__(BLOCK) void k_sequence_from_to(uint32_t* __restrict__ out, uint32_t out_cap,
                                                            const uint32_t* __restrict__ d_from,
                                                            const uint32_t* __restrict__ d_to) {
    const uint32_t from = *d_from, to = *d_to;
    const uint32_t len = min(to > from ? to - from : 0u, out_cap);
    GRID_STRIDE(i, len) out[i] = from + i;
}
__global__ void k_sequence_from_to_len(const uint32_t* __restrict__ d_from, const uint32_t* __restrict__ d_to,
                                       uint32_t* __restrict__ d_out_len) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *d_out_len = *d_to - *d_from;
}
// write_sequence.comp:15-20: out[i] = start + i * step for i < *d_len
__global__ __launch_bounds__(BLOCK) void k_write_sequence(uint32_t* __restrict__ out, const uint32_t* __restrict__ d_len,
                                                          uint32_t start, uint32_t step) {
    const uint32_t n = *d_len;
    GRID_STRIDE(i, n) out[i] = start + i * step;
}

}  // namespace apbf

using namespace apbf;

static inline hipStream_t S(apbf_stream s) { return static_cast<hipStream_t>(s); }
static inline apbf_status last_status() {
    return hipGetLastError() == hipSuccess ? APBF_OK : APBF_ERR_HIP;
}

extern "C" {

size_t apbf_prefix_sum_workspace_size(uint32_t capacity) { return scan_workspace_bytes(capacity); }
size_t apbf_radix_sort_workspace_size(uint32_t capacity) { return sort_workspace_bytes(capacity); }

apbf_status apbf_prefix_sum(uint32_t* values, const uint32_t* d_len, uint32_t capacity, void* ws,
                            size_t ws_bytes, apbf_stream stream) {
    if (!values || !d_len || !ws) return APBF_ERR_INVALID_ARGUMENT;
    if (ws_bytes < scan_workspace_bytes(capacity)) return APBF_ERR_WORKSPACE;
    scan_u32(values, values, d_len, capacity, true, nullptr, ws, S(stream));
    return last_status();
}

apbf_status apbf_radix_sort(const uint32_t* keys, const uint32_t* values, const uint32_t* d_len,
                            uint32_t capacity, uint32_t upper_bound, uint32_t* out_keys,
                            uint32_t* out_values, void* ws, size_t ws_bytes, apbf_stream stream) {
    if (!keys || !values || !d_len || !out_keys || !out_values || !ws) return APBF_ERR_INVALID_ARGUMENT;
    if (keys == out_keys || values == out_values) return APBF_ERR_INVALID_ARGUMENT;
    if (ws_bytes < sort_workspace_bytes(capacity)) return APBF_ERR_WORKSPACE;
    radix_sort_pairs(keys, values, d_len, capacity, upper_bound, out_keys, out_values, ws, S(stream));
    return last_status();
}

apbf_status apbf_append_list(void* target, uint32_t target_capacity, const uint32_t* d_target_len,
                             const void* appending, const uint32_t* d_append_len, uint32_t stride,
                             uint32_t* d_new_len, uint32_t append_capacity, apbf_stream stream) {
    if (!target || !d_target_len || !d_append_len || !d_new_len || stride == 0 || (stride % 4) != 0 ||
        (!appending && append_capacity))
        return APBF_ERR_INVALID_ARGUMENT;
    hipStream_t s = S(stream);
    const uint32_t words = stride / 4;
    if (appending && append_capacity)
        hipLaunchKernelGGL(k_append, dim3(stride_grid((uint64_t)append_capacity * words)), dim3(BLOCK), 0, s,
                           static_cast<uint32_t*>(target), static_cast<const uint32_t*>(appending), d_target_len,
                           d_append_len, target_capacity, words, d_new_len);
    hipLaunchKernelGGL(k_append_len, dim3(1), dim3(64), 0, s, d_target_len, d_append_len, target_capacity,
                       d_new_len);
    return last_status();
}

apbf_status apbf_write_increasing_sequence(uint32_t* target, uint32_t target_capacity,
                                           const uint32_t* d_sequence_min, uint32_t value_upper_bound,
                                           uint32_t sequence_length, uint32_t* d_new_sequence_min,
                                           uint32_t* d_new_target_len, apbf_stream stream) {
    if ((!target && target_capacity) || !d_sequence_min || !d_new_sequence_min || !d_new_target_len ||
        d_new_sequence_min == d_sequence_min)
        return APBF_ERR_INVALID_ARGUMENT;
    hipStream_t s = S(stream);
    const uint32_t grid_n = target_capacity < sequence_length ? target_capacity : sequence_length;
    if (grid_n)
        hipLaunchKernelGGL(k_increasing_sequence, dim3(stride_grid(grid_n)), dim3(BLOCK), 0, s, target,
                           target_capacity, d_sequence_min, value_upper_bound, sequence_length);
    hipLaunchKernelGGL(k_increasing_sequence_len, dim3(1), dim3(64), 0, s, target_capacity, d_sequence_min,
                       value_upper_bound, sequence_length, d_new_sequence_min, d_new_target_len);
    return last_status();
}

apbf_status apbf_write_increasing_sequence_from_to(uint32_t* out, uint32_t out_capacity, uint32_t* d_out_len,
                                                   const uint32_t* d_from, const uint32_t* d_to, apbf_stream stream) {
    if ((!out && out_capacity) || !d_out_len || !d_from || !d_to) return APBF_ERR_INVALID_ARGUMENT;
    hipStream_t s = S(stream);
    if (out_capacity)
        hipLaunchKernelGGL(k_sequence_from_to, dim3(stride_grid(out_capacity)), dim3(BLOCK), 0, s, out, out_capacity,
                           d_from, d_to);
    hipLaunchKernelGGL(k_sequence_from_to_len, dim3(1), dim3(64), 0, s, d_from, d_to, d_out_len);
    return last_status();
}

apbf_status apbf_write_sequence(uint32_t* out, const uint32_t* d_len, uint32_t capacity, uint32_t start, uint32_t step,
                                apbf_stream stream) {
    if ((!out && capacity) || !d_len) return APBF_ERR_INVALID_ARGUMENT;
    if (capacity)
        hipLaunchKernelGGL(k_write_sequence, dim3(stride_grid(capacity)), dim3(BLOCK), 0, S(stream), out, d_len, start,
                           step);
    return last_status();
}

apbf_status apbf_device_alloc(size_t bytes, void** out) {
    if (!out) return APBF_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (hipMalloc(out, bytes ? bytes : 4) != hipSuccess) return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_device_free(void* p) {
    if (p && hipFree(p) != hipSuccess) return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_copy(void* dst, const void* src, size_t bytes, apbf_stream stream) {
    if ((!dst || !src) && bytes) return APBF_ERR_INVALID_ARGUMENT;
    if (bytes && hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, S(stream)) != hipSuccess) return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_stream_synchronize(apbf_stream stream) {
    return hipStreamSynchronize(S(stream)) == hipSuccess ? APBF_OK : APBF_ERR_HIP;
}

apbf_status apbf_calculate_position_hash(const int32_t* position, const uint32_t* d_len, uint32_t capacity,
                                         const float* mn, const float* mx, uint32_t res_log2, uint32_t dimensions,
                                         uint32_t* out_hash, apbf_stream stream) {
    if (!position || !d_len || !mn || !mx || !out_hash || (dimensions != 2 && dimensions != 3) ||
        res_log2 * dimensions > 30)
        return APBF_ERR_INVALID_ARGUMENT;
    GridDesc g;
    g.dims = dimensions;
    for (int d = 0; d < 3; ++d) {
        g.mn[d] = mn[d];
        g.mx[d] = mx[d];
    }
    g.res = res_log2;
    launch_hash_iota(reinterpret_cast<const int4*>(position), d_len, capacity, g, out_hash, nullptr,
                     S(stream));
    return last_status();
}

apbf_status apbf_calculate_position_code(const int32_t* position, const uint32_t* d_len, uint32_t capacity,
                                         uint32_t section, uint32_t* out_code, apbf_stream stream) {
    if (!position || !d_len || !out_code || section > 2) return APBF_ERR_INVALID_ARGUMENT;
    launch_position_code(reinterpret_cast<const int4*>(position), d_len, capacity, section, out_code,
                         S(stream));
    return last_status();
}

apbf_status apbf_find_value_ranges(const uint32_t* sorted_keys, const uint32_t* d_len, uint32_t capacity,
                                   uint32_t n_cells, uint32_t* cell_start, uint32_t* cell_end,
                                   apbf_stream stream) {
    if (!sorted_keys || !d_len || !cell_start || !cell_end) return APBF_ERR_INVALID_ARGUMENT;
    fill_u32(cell_start, 0u, n_cells, S(stream));
    fill_u32(cell_end, 0u, n_cells, S(stream));
    launch_ranges(sorted_keys, d_len, capacity, cell_start, cell_end, S(stream));
    return last_status();
}

size_t apbf_neighborhood_green_workspace_size(uint32_t capacity) {
    return (size_t)C_COUNT * 4 + (size_t)(capacity + 1) * 4 + scan_workspace_bytes(capacity + 1);
}

apbf_status apbf_neighborhood_green(const int32_t* position, const float* range, const uint32_t* d_len,
                                    uint32_t capacity, float range_scale, const float* mn, const float* mx,
                                    uint32_t res_log2, uint32_t dimensions, const uint32_t* cell_start,
                                    const uint32_t* cell_end, uint32_t* row_offset, uint32_t* neighbors,
                                    uint32_t pair_capacity, uint32_t* d_n_pairs, void* ws, size_t ws_bytes,
                                    apbf_stream stream) {
    if (!position || !range || !d_len || !mn || !mx || !cell_start || !cell_end || !row_offset ||
        !neighbors || !d_n_pairs || !ws || (dimensions != 2 && dimensions != 3) || res_log2 * dimensions > 30)
        return APBF_ERR_INVALID_ARGUMENT;
    // workspace: count block + counts per particle (capacity + 1) + scan workspace
    if (ws_bytes < apbf_neighborhood_green_workspace_size(capacity)) return APBF_ERR_WORKSPACE;
    hipStream_t s = S(stream);
    uint32_t* counts = static_cast<uint32_t*>(ws);
    uint32_t* cnt = counts + C_COUNT;
    void* scan_ws = cnt + capacity + 1;
    GridDesc g;
    g.dims = dimensions;
    for (int d = 0; d < 3; ++d) {
        g.mn[d] = mn[d];
        g.mx[d] = mx[d];
    }
    g.res = res_log2;
    fill_u32(counts, 0u, C_COUNT, s);
    hipMemcpyAsync(counts + C_N, d_len, 4, hipMemcpyDeviceToDevice, s);
    launch_set_n_plus_1(counts, s);
    const int4* p4 = reinterpret_cast<const int4*>(position);
    launch_walk(0, false, p4, range, nullptr, nullptr, 0, cell_start, cell_end, counts, capacity, range_scale, g,
                pair_capacity, cnt, nullptr, nullptr, nullptr, nullptr, nullptr, s);
    scan_u32(cnt, row_offset, counts + C_N_PLUS_1, capacity + 1, false, nullptr, scan_ws, s);
    launch_pairs_total(counts, row_offset, pair_capacity, s);
    launch_walk(1, false, p4, range, nullptr, nullptr, 0, cell_start, cell_end, counts, capacity, range_scale, g,
                pair_capacity, nullptr, nullptr, row_offset, nullptr, neighbors, nullptr, s);
    hipMemcpyAsync(d_n_pairs, counts + C_PAIRS, 4, hipMemcpyDeviceToDevice, s);
    return last_status();
}

apbf_status apbf_neighborhood_brute_force(const int32_t* position, const float* range, const uint32_t* d_len,
                                          uint32_t capacity, float range_scale, uint32_t* row_offset,
                                          uint32_t* neighbors, uint32_t pair_capacity, uint32_t* d_n_pairs, void* ws,
                                          size_t ws_bytes, apbf_stream stream) {
    if (((!position || !range) && capacity) || !d_len || !row_offset || !neighbors || !d_n_pairs || !ws)
        return APBF_ERR_INVALID_ARGUMENT;  // an empty list (capacity 0) may come without particle arrays
    if (ws_bytes < apbf_neighborhood_green_workspace_size(capacity)) return APBF_ERR_WORKSPACE;
    hipStream_t s = S(stream);
    uint32_t* counts = static_cast<uint32_t*>(ws);
    uint32_t* cnt = counts + C_COUNT;
    void* scan_ws = cnt + capacity + 1;
    fill_u32(counts, 0u, C_COUNT, s);
    hipMemcpyAsync(counts + C_N, d_len, 4, hipMemcpyDeviceToDevice, s);
    launch_set_n_plus_1(counts, s);
    const int4* p4 = reinterpret_cast<const int4*>(position);
    launch_brute(false, p4, range, counts, capacity, range_scale, pair_capacity, cnt, nullptr, nullptr, s);
    scan_u32(cnt, row_offset, counts + C_N_PLUS_1, capacity + 1, false, nullptr, scan_ws, s);
    launch_pairs_total(counts, row_offset, pair_capacity, s);
    launch_brute(true, p4, range, counts, capacity, range_scale, pair_capacity, nullptr, row_offset, neighbors, s);
    hipMemcpyAsync(d_n_pairs, counts + C_PAIRS, 4, hipMemcpyDeviceToDevice, s);
    return last_status();
}

apbf_status apbf_copy_scattered_read(const void* src, void* dst, const uint32_t* edit,
                                     const uint32_t* d_edit_len, uint32_t capacity, uint32_t stride,
                                     apbf_stream stream) {
    if (!src || !dst || !edit || !d_edit_len || stride == 0 || (stride % 4) != 0)
        return APBF_ERR_INVALID_ARGUMENT;
    gather_words(static_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst), edit, d_edit_len, capacity,
                 stride / 4, S(stream));
    return last_status();
}

size_t apbf_hidden_edit_workspace_size(uint32_t index_capacity, uint32_t hidden_capacity,
                                       uint32_t edit_capacity) {
    size_t w = 64;
    w += (size_t)(hidden_capacity + 1) * 4 * 2;   // cnt, off
    w += (size_t)(index_capacity + 1) * 4 * 3;    // sorted keys, sorted positions, iota
    w += (size_t)(edit_capacity + 1) * 4 * 2;     // per-slot counts, offsets
    size_t scan = std::max(scan_workspace_bytes(hidden_capacity + 1), scan_workspace_bytes(edit_capacity + 1));
    w += std::max(scan, sort_workspace_bytes(index_capacity));
    return w + 256;
}

apbf_status apbf_apply_hidden_edit(const uint32_t* index_list, const uint32_t* d_index_len,
                                   uint32_t index_capacity, uint32_t hidden_capacity, const uint32_t* edit,
                                   const uint32_t* d_edit_len, uint32_t edit_capacity, uint32_t max_new,
                                   uint32_t* new_index, uint32_t* new_edit, uint32_t* d_new_len, void* ws,
                                   size_t ws_bytes, apbf_stream stream) {
    if (!index_list || !d_index_len || !edit || !d_edit_len || !new_index || !new_edit || !d_new_len || !ws)
        return APBF_ERR_INVALID_ARGUMENT;
    if (ws_bytes < apbf_hidden_edit_workspace_size(index_capacity, hidden_capacity, edit_capacity))
        return APBF_ERR_WORKSPACE;
    hipStream_t s = S(stream);
    uint32_t* small = static_cast<uint32_t*>(ws);  // [0] H+1, [1] E+1, [2] total
    uint32_t* cnt = small + 16;
    uint32_t* off = cnt + hidden_capacity + 1;
    uint32_t* skeys = off + hidden_capacity + 1;
    uint32_t* spos = skeys + index_capacity + 1;
    uint32_t* iota = spos + index_capacity + 1;
    uint32_t* kk = iota + index_capacity + 1;
    uint32_t* oo = kk + edit_capacity + 1;
    void* inner = oo + edit_capacity + 1;
    hipLaunchKernelGGL(k_set_value, dim3(1), dim3(64), 0, s, small, hidden_capacity + 1);
    fill_u32(cnt, 0u, (uint64_t)hidden_capacity + 1, s);
    hipLaunchKernelGGL(k_count_values, dim3(stride_grid(index_capacity)), dim3(BLOCK), 0, s, index_list,
                       d_index_len, hidden_capacity, cnt);
    scan_u32(cnt, off, small, hidden_capacity + 1, false, nullptr, inner, s);
    // positions grouped by hidden id, ascending position within a group
    iota_u32(iota, d_index_len, index_capacity, s);
    radix_sort_pairs(index_list, iota, d_index_len, index_capacity, 0xFFFFFFFFu, skeys, spos, inner, s);
    hipLaunchKernelGGL(k_set_plus_one, dim3(1), dim3(64), 0, s, d_edit_len, small + 1);
    hipLaunchKernelGGL(k_emit_counts, dim3(stride_grid(edit_capacity)), dim3(BLOCK), 0, s, edit, d_edit_len,
                       hidden_capacity, cnt, kk);
    scan_u32(kk, oo, small + 1, edit_capacity + 1, false, small + 2, inner, s);
    hipLaunchKernelGGL(k_emit_entries, dim3(stride_grid(edit_capacity)), dim3(BLOCK), 0, s, edit, d_edit_len,
                       hidden_capacity, cnt, off, spos, oo, max_new, new_index, new_edit);
    hipLaunchKernelGGL(k_clamp_len, dim3(1), dim3(64), 0, s, small + 2, max_new, d_new_len);
    return last_status();
}

size_t apbf_delete_edit_list_workspace_size(uint32_t hidden_capacity) {
    return scan_workspace_bytes(hidden_capacity) + (size_t)hidden_capacity * 8 + 64;  // flags, scan, scan tiles
}

apbf_status apbf_delete_edit_list(const uint32_t* deleted, const uint32_t* d_deleted_len,
                                  uint32_t deleted_capacity, const uint32_t* d_hidden_len,
                                  uint32_t hidden_capacity, uint32_t* edit, uint32_t* d_edit_len, void* ws,
                                  size_t ws_bytes, apbf_stream stream) {
    if (!deleted || !d_deleted_len || !d_hidden_len || !edit || !d_edit_len || !ws)
        return APBF_ERR_INVALID_ARGUMENT;
    if (ws_bytes < apbf_delete_edit_list_workspace_size(hidden_capacity)) return APBF_ERR_WORKSPACE;
    hipStream_t s = S(stream);
    uint32_t* flags = static_cast<uint32_t*>(ws);
    uint32_t* scan = flags + hidden_capacity;
    void* inner = scan + hidden_capacity;
    fill_u32_dev(flags, 1u, d_hidden_len, hidden_capacity, s);
    hipLaunchKernelGGL(k_flag_deleted, dim3(stride_grid(deleted_capacity)), dim3(BLOCK), 0, s, deleted,
                       d_deleted_len, d_hidden_len, flags);
    scan_u32(flags, scan, d_hidden_len, hidden_capacity, false, d_edit_len, inner, s);
    hipLaunchKernelGGL(k_scatter_kept, dim3(stride_grid(hidden_capacity)), dim3(BLOCK), 0, s, flags, scan,
                       d_hidden_len, edit);
    return last_status();
}

}  // extern "C"
