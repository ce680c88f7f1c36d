// Device-wide scan and radix sort for gfx950 (64-lane wavefronts).
//
// Replaces algorithms::prefix_sum / algorithms::sort (algorithms.cpp:59-118)
// and their shaders (prefix_sum_apply_on_block_level.comp,
// prefix_sum_spread_from_block_level.comp, radix_sort_apply_on_block_level.comp,
// radix_sort_scattered_write.comp). Lengths are read from device memory so the
// host never synchronises; grids are sized for the capacity and tiles past the
// length exit or contribute zeros.
//
// Scan: reduce-then-scan over tiles of 2048 values (256 threads x 8 values),
// two launches: the tile sums, then per tile the sum of the tile sums before
// it plus the scan of its own values. Wave-level inclusive scans use
// __shfl_up over the 64 lanes; the four waves of a block combine through LDS.
//
// Radix sort: LSD on 4-, 8- (default: half the passes of the reference's
// 4-bit digits) or 11-bit digits, tiles of 4, 8 or 16 values per thread.
// Per pass: (1) per-tile digit histograms written bucket-major [digit][tile],
// (2) one exclusive scan of that table gives every (digit, tile) its global
// output offset, (3) a stable scatter: each tile walks its values in 256-wide
// rounds in index order; inside a round a value's rank among equal digits
// comes from a 64-lane match built from one __ballot mask per digit bit, and
// per-wave digit counts in LDS give the cross-wave and cross-round offsets.
#include <algorithm>

#include "apbf_internal.h"

namespace apbf {

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive block scan of one value per thread; returns the exclusive prefix
// and the block total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds4, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) lds4[wave] = inc;
    __syncthreads();
    uint32_t w0 = lds4[0], w1 = lds4[1], w2 = lds4[2], w3 = lds4[3];
    uint32_t woff = (wave > 0 ? w0 : 0u) + (wave > 1 ? w1 : 0u) + (wave > 2 ? w2 : 0u);
    total = w0 + w1 + w2 + w3;
    __syncthreads();
    return woff + inc - v;
}

__global__ __launch_bounds__(BLOCK) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                       const uint32_t* __restrict__ d_len,
                                                       uint32_t n_host, uint32_t* __restrict__ tile_sums) {
    __shared__ uint32_t lds[4];
    const uint32_t n = d_len ? *d_len : n_host;
    const uint32_t tile = blockIdx.x;
    const uint64_t base = (uint64_t)tile * SCAN_TILE;
    uint32_t sum = 0;
    if (base < n) {
        uint64_t i0 = base + (uint64_t)threadIdx.x * SCAN_ITEMS;
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; ++k)
            if (i0 + k < n) sum += in[i0 + k];
    }
    uint32_t total;
    block_excl_scan(sum, lds, total);
    if (threadIdx.x == 0) tile_sums[tile] = (base < n) ? total : 0u;
}

// Second pass: every tile adds the sum of the tile sums before it (at most a
// few thousand values, read by the block from L2) to the scan of its own
// values; block 0 also writes the grand total. Two launches per scan.
__global__ __launch_bounds__(BLOCK) void k_scan_apply(const uint32_t* __restrict__ in,
                                                      uint32_t* __restrict__ out,
                                                      const uint32_t* __restrict__ d_len, uint32_t n_host,
                                                      const uint32_t* __restrict__ tile_sums, uint32_t n_tiles,
                                                      int inclusive, uint32_t* __restrict__ total_out) {
    __shared__ uint32_t lds[4];
    const uint32_t n = d_len ? *d_len : n_host;
    const uint32_t tile = blockIdx.x;
    const uint64_t base = (uint64_t)tile * SCAN_TILE;
    if (tile == 0 && total_out) {  // uniform in the block
        uint32_t acc = 0;
        for (uint32_t k = threadIdx.x; k < n_tiles; k += BLOCK) acc += tile_sums[k];
        uint32_t all;
        block_excl_scan(acc, lds, all);
        if (threadIdx.x == 0) *total_out = all;
    }
    if (base >= n) return;
    uint32_t acc = 0;
    for (uint32_t k = threadIdx.x; k < tile; k += BLOCK) acc += tile_sums[k];
    uint32_t before;
    block_excl_scan(acc, lds, before);
    uint64_t i0 = base + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (i0 + k < n) ? in[i0 + k] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t run = block_excl_scan(sum, lds, total) + before;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        uint32_t x = v[k];
        if (i0 + k < n) out[i0 + k] = inclusive ? run + x : run;
        run += x;
    }
}

size_t scan_workspace_bytes(uint32_t capacity) {
    return (size_t)(div_up(capacity ? capacity : 1, SCAN_TILE) + 1) * 4 + 256;
}

static void scan_impl(const uint32_t* in, uint32_t* out, const uint32_t* d_len, uint32_t n_host,
                      uint32_t capacity, bool inclusive, uint32_t* total, void* ws, hipStream_t s) {
    uint32_t tiles = div_up(capacity ? capacity : 1, SCAN_TILE);
    uint32_t* tile_sums = static_cast<uint32_t*>(ws);
    hipLaunchKernelGGL(k_scan_reduce, dim3(tiles), dim3(BLOCK), 0, s, in, d_len, n_host, tile_sums);
    hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(BLOCK), 0, s, in, out, d_len, n_host, tile_sums, tiles,
                       inclusive ? 1 : 0, total);
}

void scan_u32(const uint32_t* in, uint32_t* out, const uint32_t* d_len, uint32_t capacity,
              bool inclusive, uint32_t* total, void* ws, hipStream_t s) {
    scan_impl(in, out, d_len, 0, capacity, inclusive, total, ws, s);
}

void scan_u32_n(const uint32_t* in, uint32_t* out, uint32_t n, bool inclusive, uint32_t* total,
                void* ws, hipStream_t s) {
    scan_impl(in, out, nullptr, n, n, inclusive, total, ws, s);
}

// ------------------------------------------------------------------ radix sort
template <int BITS, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_sort_hist(const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ d_len,
                                                     uint32_t shift, uint32_t tiles_cap,
                                                     uint32_t* __restrict__ hist) {
    constexpr uint32_t R = 1u << BITS;
    __shared__ uint32_t h[R];
    const uint32_t n = *d_len;
    const uint32_t tile = blockIdx.x;
    for (uint32_t b = threadIdx.x; b < R; b += BLOCK) h[b] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)tile * (ITEMS * BLOCK);
    if (base < n) {
#pragma unroll
        for (int r = 0; r < ITEMS; ++r) {
            uint64_t i = base + (uint64_t)r * BLOCK + threadIdx.x;
            if (i < n) atomicAdd(&h[(keys[i] >> shift) & (R - 1)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < R; b += BLOCK) hist[(uint64_t)b * tiles_cap + tile] = h[b];
}

// A tile's values in ITEMS rounds of 256 in index order. In a round, a
// value's rank among the equal digits of its wave is a 64-lane match built
// from BITS ballots; the first lane of each digit group publishes the
// group's size per wave in LDS, every value adds the groups of the waves
// before it and the running offset of its digit, and the group leaders then
// advance the running offsets and clear their entries (no pass over all
// 2^BITS digits per round).
template <int BITS, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_sort_scatter(const uint32_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ vals,
                                                        const uint32_t* __restrict__ d_len,
                                                        uint32_t shift, uint32_t tiles_cap,
                                                        const uint32_t* __restrict__ offsets,
                                                        uint32_t* __restrict__ out_keys,
                                                        uint32_t* __restrict__ out_vals) {
    constexpr uint32_t R = 1u << BITS;
    __shared__ uint32_t running[R];
    __shared__ uint32_t wcnt[4][R];
    const uint32_t n = *d_len;
    const uint32_t tile = blockIdx.x;
    const uint64_t base = (uint64_t)tile * (ITEMS * BLOCK);
    if (base >= n) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t b = threadIdx.x; b < R; b += BLOCK) {
        running[b] = offsets[(uint64_t)b * tiles_cap + tile];
        wcnt[0][b] = wcnt[1][b] = wcnt[2][b] = wcnt[3][b] = 0;
    }
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    __syncthreads();
    for (int r = 0; r < ITEMS; ++r) {
        uint64_t i = base + (uint64_t)r * BLOCK + threadIdx.x;
        bool valid = i < n;
        uint32_t key = valid ? keys[i] : 0u;
        uint32_t val = valid ? vals[i] : 0u;
        uint32_t d = (key >> shift) & (R - 1);
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            uint64_t bb = __ballot((d >> b) & 1u);
            match &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(match & lt_mask);
        const uint32_t cnt = (uint32_t)__popcll(match);
        const bool leader = valid && rank == 0;
        if (leader) wcnt[wave][d] = cnt;
        __syncthreads();
        if (valid) {
            uint32_t off = running[d] + rank;
            for (int w = 0; w < wave; ++w) off += wcnt[w][d];
            out_keys[off] = key;
            out_vals[off] = val;
        }
        __syncthreads();
        if (leader) {
            atomicAdd(&running[d], cnt);
            wcnt[wave][d] = 0;
        }
        __syncthreads();
    }
}

__global__ void k_copy_pairs(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                             uint32_t* __restrict__ oa, uint32_t* __restrict__ ob,
                             const uint32_t* __restrict__ d_len) {
    const uint32_t n = *d_len;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        oa[i] = a[i];
        ob[i] = b[i];
    }
}

bool sort_shape_valid(SortShape sh) {
    return (sh.bits == 4u || sh.bits == 8u || sh.bits == 11u) && (sh.items == 4u || sh.items == 8u || sh.items == 16u);
}

size_t sort_workspace_bytes(uint32_t capacity, SortShape sh) {
    const uint32_t tiles = div_up(capacity ? capacity : 1, (uint64_t)BLOCK * sh.items);
    const size_t digits = (size_t)tiles << sh.bits;
    const size_t tmp = (size_t)(capacity ? capacity : 1) * 8;
    return digits * 4 + scan_workspace_bytes((uint32_t)digits) + tmp + 1024;
}

size_t sort_workspace_bytes_any(uint32_t capacity) {
    size_t m = 0;
    for (uint32_t bits : {4u, 8u, 11u})
        for (uint32_t items : {4u, 8u, 16u}) m = std::max(m, sort_workspace_bytes(capacity, SortShape{bits, items}));
    return m;
}

namespace {
template <int BITS, int ITEMS>
void sort_pass(const uint32_t* src_k, const uint32_t* src_v, const uint32_t* d_len, uint32_t shift, uint32_t tiles,
               uint32_t* hist, void* scan_ws, uint32_t* dk, uint32_t* dv, hipStream_t s) {
    hipLaunchKernelGGL((k_sort_hist<BITS, ITEMS>), dim3(tiles), dim3(BLOCK), 0, s, src_k, d_len, shift, tiles, hist);
    scan_u32_n(hist, hist, tiles << BITS, false, nullptr, scan_ws, s);
    hipLaunchKernelGGL((k_sort_scatter<BITS, ITEMS>), dim3(tiles), dim3(BLOCK), 0, s, src_k, src_v, d_len, shift,
                       tiles, hist, dk, dv);
}
template <int BITS>
void sort_pass_items(uint32_t items, const uint32_t* src_k, const uint32_t* src_v, const uint32_t* d_len,
                     uint32_t shift, uint32_t tiles, uint32_t* hist, void* scan_ws, uint32_t* dk, uint32_t* dv,
                     hipStream_t s) {
    if (items == 4u)
        sort_pass<BITS, 4>(src_k, src_v, d_len, shift, tiles, hist, scan_ws, dk, dv, s);
    else if (items == 16u)
        sort_pass<BITS, 16>(src_k, src_v, d_len, shift, tiles, hist, scan_ws, dk, dv, s);
    else
        sort_pass<BITS, 8>(src_k, src_v, d_len, shift, tiles, hist, scan_ws, dk, dv, s);
}
}  // namespace

void radix_sort_pairs(const uint32_t* keys, const uint32_t* vals, const uint32_t* d_len,
                      uint32_t capacity, uint32_t upper_bound, uint32_t* out_keys,
                      uint32_t* out_vals, void* ws, hipStream_t s, SortShape sh) {
    if (!sort_shape_valid(sh)) sh = SORT_DEFAULT;
    const uint32_t cap = capacity ? capacity : 1;
    const uint32_t tiles = div_up(cap, (uint64_t)BLOCK * sh.items);
    char* w = static_cast<char*>(ws);
    uint32_t* hist = reinterpret_cast<uint32_t*>(w);
    w += ((size_t)tiles << sh.bits) * 4;
    void* scan_ws = w;
    w += scan_workspace_bytes(tiles << sh.bits);
    uint32_t* tk = reinterpret_cast<uint32_t*>(w);
    uint32_t* tv = tk + cap;
    int passes = 0;
    for (uint32_t off = 0; off < 32u && (upper_bound >> off) != 0u; off += sh.bits) ++passes;
    if (passes == 0) {
        hipLaunchKernelGGL(k_copy_pairs, dim3(stride_grid(cap)), dim3(BLOCK), 0, s, keys, vals,
                           out_keys, out_vals, d_len);
        return;
    }
    // Ping-pong so that the last pass lands in out_*.
    const uint32_t* src_k = keys;
    const uint32_t* src_v = vals;
    for (int p = 0; p < passes; ++p) {
        // pass p writes to out_* when (passes - 1 - p) is even
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t* dk = to_out ? out_keys : tk;
        uint32_t* dv = to_out ? out_vals : tv;
        const uint32_t shift = (uint32_t)p * sh.bits;
        if (sh.bits == 4u)
            sort_pass_items<4>(sh.items, src_k, src_v, d_len, shift, tiles, hist, scan_ws, dk, dv, s);
        else if (sh.bits == 11u)
            sort_pass_items<11>(sh.items, src_k, src_v, d_len, shift, tiles, hist, scan_ws, dk, dv, s);
        else
            sort_pass_items<8>(sh.items, src_k, src_v, d_len, shift, tiles, hist, scan_ws, dk, dv, s);
        src_k = dk;
        src_v = dv;
    }
}

// ------------------------------------------------------------------ helpers
__global__ void k_fill(uint32_t* __restrict__ p, uint32_t v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_dev(uint32_t* __restrict__ p, uint32_t v, const uint32_t* __restrict__ d_len) {
    const uint32_t n = *d_len;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_copy_dev(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                           const uint32_t* __restrict__ d_len) {
    const uint32_t n = *d_len;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
__global__ void k_iota(uint32_t* __restrict__ p, const uint32_t* __restrict__ d_len) {
    const uint32_t n = *d_len;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}
__global__ void k_gather_words(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                               const uint32_t* __restrict__ edit, const uint32_t* __restrict__ d_len,
                               uint32_t words) {
    const uint64_t n = (uint64_t)(*d_len) * words;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t e = i / words, o = i - e * words;
        dst[i] = src[(uint64_t)edit[e] * words + o];
    }
}
__global__ void k_copy1(uint32_t* dst, const uint32_t* src) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

void fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill, dim3(stride_grid(n)), dim3(BLOCK), 0, s, p, value, n);
}
void fill_u32_dev(uint32_t* p, uint32_t value, const uint32_t* d_len, uint32_t capacity, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_dev, dim3(stride_grid(capacity)), dim3(BLOCK), 0, s, p, value, d_len);
}
void copy_u32_dev(uint32_t* dst, const uint32_t* src, const uint32_t* d_len, uint32_t capacity, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_dev, dim3(stride_grid(capacity)), dim3(BLOCK), 0, s, dst, src, d_len);
}
void iota_u32(uint32_t* p, const uint32_t* d_len, uint32_t capacity, hipStream_t s) {
    hipLaunchKernelGGL(k_iota, dim3(stride_grid(capacity)), dim3(BLOCK), 0, s, p, d_len);
}
void gather_words(const uint32_t* src, uint32_t* dst, const uint32_t* edit, const uint32_t* d_len,
                  uint32_t capacity, uint32_t words, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_words, dim3(stride_grid((uint64_t)capacity * words)), dim3(BLOCK), 0, s,
                       src, dst, edit, d_len, words);
}
void copy_u32(uint32_t* dst, const uint32_t* src, hipStream_t s) {
    hipLaunchKernelGGL(k_copy1, dim3(1), dim3(64), 0, s, dst, src);
}

// Every float bit pattern through the hardware conversions (to_i32 / to_u32)
// and through their definitions (to_i32_ref / to_u32_ref): the number of
// patterns where they differ, the smallest such pattern, and the hardware
// results of every `stride`-th pattern.
__global__ __launch_bounds__(BLOCK) void k_check_conversions(uint32_t stride, unsigned long long* mismatches,
                                                             uint32_t* first_bad, int32_t* sample_i,
                                                             uint32_t* sample_u) {
    unsigned long long bad = 0;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < (1ull << 32);
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = (uint32_t)k;
        const float x = bits_f(b);
        const int32_t hi = to_i32(x);
        const uint32_t hu = to_u32(x);
        if (hi != to_i32_ref(x) || hu != to_u32_ref(x)) {
            ++bad;
            atomicMin(first_bad, b);
        }
        if (b % stride == 0u) {
            sample_i[b / stride] = hi;
            sample_u[b / stride] = hu;
        }
    }
    if (bad) atomicAdd(mismatches, bad);
}

// Every float bit pattern in the stated ranges through sqrt_in_range /
// recip_in_range and through the IEEE square root and 1/x, and 2^32 quotients
// through quotient_by_recip and the IEEE division (512 divisors, each with
// every dividend significand): the number of cases where the bits differ and
// the smallest such case (bit pattern, or divisor index << 23 | significand),
// per function (slot 0 quotient, 1 square root, 2 reciprocal).
__global__ __launch_bounds__(BLOCK) void k_check_math(unsigned long long* mismatches, uint32_t* first_bad) {
    unsigned long long bad[3] = {0, 0, 0};
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < (1ull << 32);
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = (uint32_t)k;
        const float x = bits_f(b);
        const float ax = __builtin_fabsf(x);
        bool d[3] = {false, false, false};
        {
            // divisor d: a hashed significand (d = 0: all ones, d = 1: 1.5) with exponents
            // of -40 to 40, the dividend 1.m with exponents of -80 to 80 and either sign
            const uint32_t di = b >> 23, m = b & 0x7FFFFFu;
            uint32_t h = di * 0x9E3779B1u;
            h ^= h >> 15;
            h *= 0x85EBCA77u;
            h ^= h >> 13;
            const uint32_t bm = di == 0u ? 0x7FFFFFu : (di == 1u ? 0x400000u : (h & 0x7FFFFFu));
            const uint32_t be = 127u + (di % 81u) - 40u;
            const uint32_t xe = 127u + ((di * 7u + (m >> 11)) % 161u) - 80u;
            const float dv = bits_f((be << 23) | bm);
            const float xv = bits_f(((m & 1u) << 31) | (xe << 23) | m);
            d[0] = __builtin_bit_cast(uint32_t, quotient_by_recip(xv, dv, recip_in_range(dv))) !=
                   __builtin_bit_cast(uint32_t, xv / dv);
        }
        if (x == 0.0f || (x >= 0x1p-96f && x <= 0x1p120f))
            d[1] = __builtin_bit_cast(uint32_t, sqrt_in_range(x)) != __builtin_bit_cast(uint32_t, __builtin_sqrtf(x));
        if (ax >= 0x1p-120f && ax <= 0x1p120f)
            d[2] = __builtin_bit_cast(uint32_t, recip_in_range(x)) != __builtin_bit_cast(uint32_t, 1.0f / x);
        for (int f = 0; f < 3; ++f)
            if (d[f]) {
                ++bad[f];
                atomicMin(&first_bad[f], b);
            }
    }
    for (int f = 0; f < 3; ++f)
        if (bad[f]) atomicAdd(&mismatches[f], bad[f]);
}

}  // namespace apbf

extern "C" apbf_status apbf_check_math(uint64_t* mismatches3, uint32_t* first_bad3) {
    using namespace apbf;
    if (!mismatches3 || !first_bad3) return APBF_ERR_INVALID_ARGUMENT;
    void* mem = nullptr;
    if (hipMalloc(&mem, 64) != hipSuccess) return APBF_ERR_HIP;
    unsigned long long* d_bad = static_cast<unsigned long long*>(mem);
    uint32_t* d_first = reinterpret_cast<uint32_t*>(static_cast<char*>(mem) + 32);
    const uint64_t zero[4] = {0, 0, 0, 0};
    const uint32_t ones[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    apbf_status st = APBF_OK;
    if (hipMemcpy(d_bad, zero, 32, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_first, ones, 16, hipMemcpyHostToDevice) != hipSuccess)
        st = APBF_ERR_HIP;
    if (st == APBF_OK) {
        hipLaunchKernelGGL(k_check_math, dim3(4096), dim3(BLOCK), 0, nullptr, d_bad, d_first);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) st = APBF_ERR_HIP;
    }
    unsigned long long bad[3] = {0, 0, 0};
    if (st == APBF_OK && (hipMemcpy(bad, d_bad, 24, hipMemcpyDeviceToHost) != hipSuccess ||
                          hipMemcpy(first_bad3, d_first, 12, hipMemcpyDeviceToHost) != hipSuccess))
        st = APBF_ERR_HIP;
    for (int f = 0; f < 3; ++f) mismatches3[f] = bad[f];
    (void)hipFree(mem);
    return st;
}

extern "C" apbf_status apbf_check_conversions(uint32_t stride, uint64_t* mismatches, uint32_t* first_bad,
                                              int32_t* sample_i32, uint32_t* sample_u32) {
    using namespace apbf;
    if (!mismatches || !first_bad || !sample_i32 || !sample_u32 || stride < 1024u) return APBF_ERR_INVALID_ARGUMENT;
    const uint64_t samples = ((1ull << 32) - 1) / stride + 1;
    void* mem = nullptr;
    const size_t bytes = 16 + samples * 8;
    if (hipMalloc(&mem, bytes) != hipSuccess) return APBF_ERR_HIP;
    unsigned long long* d_bad = static_cast<unsigned long long*>(mem);
    uint32_t* d_first = reinterpret_cast<uint32_t*>(static_cast<char*>(mem) + 8);
    int32_t* d_si = reinterpret_cast<int32_t*>(static_cast<char*>(mem) + 16);
    uint32_t* d_su = reinterpret_cast<uint32_t*>(d_si + samples);
    const uint32_t init[4] = {0u, 0u, 0xFFFFFFFFu, 0u};
    apbf_status st = APBF_OK;
    if (hipMemcpy(mem, init, 16, hipMemcpyHostToDevice) != hipSuccess) st = APBF_ERR_HIP;
    if (st == APBF_OK) {
        hipLaunchKernelGGL(k_check_conversions, dim3(4096), dim3(BLOCK), 0, nullptr, stride, d_bad, d_first, d_si,
                           d_su);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) st = APBF_ERR_HIP;
    }
    unsigned long long bad = 0;
    if (st == APBF_OK && (hipMemcpy(&bad, d_bad, 8, hipMemcpyDeviceToHost) != hipSuccess ||
                          hipMemcpy(first_bad, d_first, 4, hipMemcpyDeviceToHost) != hipSuccess ||
                          hipMemcpy(sample_i32, d_si, samples * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                          hipMemcpy(sample_u32, d_su, samples * 4, hipMemcpyDeviceToHost) != hipSuccess))
        st = APBF_ERR_HIP;
    *mismatches = bad;
    (void)hipFree(mem);
    return st;
}
