// Internal declarations shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/apbf.h"
#include "apbf_math.h"

namespace apbf {

// shaders/cpu_gpu_shared_config.h:7-17 (DIMENSIONS is a scene parameter)
// A hash grid resolution the cell tables can index: 2^(D res) cells, at
// most 2^30.
inline bool hash_res_ok(uint32_t res, uint32_t dims) { return res >= 1u && res * dims <= 30u; }
constexpr float POS_RES = 262144.0f;
constexpr float KW_RES = 262144.0f;
constexpr float INCOMP_RES = 262144.0f;
constexpr float KERNEL_SCALE = 4.0f;
constexpr float KW_PROPAGATION = 0.5f;
constexpr float PI_F = 3.14159265f;
constexpr float INCOMP_EPSILON = 0.01f;
// initialize_split_particles.comp:26 stores the GLSL integer 1/0, which glslang
// folds to INT_MAX; as a float that is 2^31.
constexpr float SPLIT_INV_MASS = 2147483648.0f;

constexpr int BLOCK = 256;       // threads per block for streaming kernels
constexpr int SCAN_ITEMS = 8;    // values per thread in the scan tiles
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;
constexpr int SORT_ITEMS = 8;
constexpr int SORT_TILE = BLOCK * SORT_ITEMS;
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;

// Grid-stride loop over [0, n).
#define GRID_STRIDE(i, n)                                                              \
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Grid for a grid-stride kernel over up to `capacity` items.
inline uint32_t stride_grid(uint64_t capacity, uint32_t max_blocks = 4096) {
    uint32_t g = div_up(capacity ? capacity : 1, BLOCK);
    return g < max_blocks ? g : max_blocks;
}

// ---------------------------------------------------------------- primitives
size_t scan_workspace_bytes(uint32_t capacity);
// Exclusive (inclusive = false) or inclusive scan of d_len[0] u32 values from
// `in` into `out` (may alias). If `total` is non-null it receives the sum.
void scan_u32(const uint32_t* in, uint32_t* out, const uint32_t* d_len, uint32_t capacity,
              bool inclusive, uint32_t* total, void* ws, hipStream_t s);
// Same, over a length known on the host.
void scan_u32_n(const uint32_t* in, uint32_t* out, uint32_t n, bool inclusive, uint32_t* total,
                void* ws, hipStream_t s);

// Digit width (4, 8 or 11 bits) and values per thread of a tile (4, 8 or
// 16; 256 threads) of the radix sort. Every shape gives the same stable
// order; they differ in passes and launch sizes.
struct SortShape {
    uint32_t bits;
    uint32_t items;
};
constexpr SortShape SORT_DEFAULT{8u, 8u};
bool sort_shape_valid(SortShape shape);
size_t sort_workspace_bytes(uint32_t capacity, SortShape shape = SORT_DEFAULT);
size_t sort_workspace_bytes_any(uint32_t capacity);  // enough for every shape
void radix_sort_pairs(const uint32_t* keys, const uint32_t* vals, const uint32_t* d_len,
                      uint32_t capacity, uint32_t upper_bound, uint32_t* out_keys,
                      uint32_t* out_vals, void* ws, hipStream_t s, SortShape shape = SORT_DEFAULT);

void fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);
void fill_u32_dev(uint32_t* p, uint32_t value, const uint32_t* d_len, uint32_t capacity,
                  hipStream_t s);
void copy_u32_dev(uint32_t* dst, const uint32_t* src, const uint32_t* d_len, uint32_t capacity, hipStream_t s);
void iota_u32(uint32_t* p, const uint32_t* d_len, uint32_t capacity, hipStream_t s);
void gather_words(const uint32_t* src, uint32_t* dst, const uint32_t* edit,
                  const uint32_t* d_len, uint32_t capacity, uint32_t words, hipStream_t s);
void copy_u32(uint32_t* dst, const uint32_t* src, hipStream_t s);

}  // namespace apbf
