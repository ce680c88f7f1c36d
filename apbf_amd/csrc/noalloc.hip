// apbf_radix_sort_noalloc: the stable radix sort of apbf_radix_sort with its
// workspace (digit histograms, their scan's tile sums, the ping-pong key and
// value buffers) in a fixed device array of this file instead of a caller
// buffer, for callers that cannot allocate device memory (code under graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// Every GPU the library runs on holds its own copy of `scratch` until the
// process ends. Host code never passes `scratch` itself to a kernel: on the
// host the name is a stand-in, not a device address. Its device address is
// looked up once per device with hipGetSymbolAddress (with that device
// current) and kept in `g_scratch_ptr`; kernels receive that pointer.
//
// Nothing in the workspace has to start at zero: every histogram entry, tile
// sum and ping-pong element is written before it is read in each call. Calls
// on one GPU share the array, so calls that overlap in time on one GPU (on
// different streams, or from different threads) race; calls on one stream are
// ordered and safe.
#include <mutex>

#include "apbf_internal.h"

namespace apbf {
namespace {

constexpr uint32_t NOALLOC_SORT_MAX = APBF_RADIX_SORT_NOALLOC_MAX;  // elements per call
constexpr size_t NOALLOC_SCRATCH_BYTES = (NOALLOC_SORT_MAX / SORT_TILE) * (size_t)RADIX * 4 +  // histograms
                                         ((size_t)NOALLOC_SORT_MAX / SORT_TILE * RADIX / SCAN_TILE + 1) * 4 +
                                         256 +                               // scan tile sums
                                         (size_t)NOALLOC_SORT_MAX * 8 + 1024;  // ping-pong keys + values

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of this device's scratch array (looked up on the first call
// on each device); nullptr on failure.
void* scratch_for_current_device(apbf_status* st) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        *st = APBF_ERR_NO_DEVICE;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
            *st = APBF_ERR_HIP;
            return nullptr;
        }
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace apbf

extern "C" {

apbf_status apbf_radix_sort_noalloc(const uint32_t* keys, const uint32_t* values, const uint32_t* d_len,
                                    uint32_t capacity, uint32_t upper_bound, uint32_t* out_keys,
                                    uint32_t* out_values, apbf_stream stream) {
    using namespace apbf;
    if (!keys || !values || !d_len || !out_keys || !out_values) return APBF_ERR_INVALID_ARGUMENT;
    if (keys == out_keys || values == out_values) return APBF_ERR_INVALID_ARGUMENT;
    if (capacity > NOALLOC_SORT_MAX) return APBF_ERR_CAPACITY;
    if (sort_workspace_bytes(capacity) > NOALLOC_SCRATCH_BYTES) return APBF_ERR_CAPACITY;
    apbf_status st = APBF_OK;
    void* ws = scratch_for_current_device(&st);
    if (!ws) return st;
    radix_sort_pairs(keys, values, d_len, capacity, upper_bound, out_keys, out_values, ws,
                     static_cast<hipStream_t>(stream));
    return hipGetLastError() == hipSuccess ? APBF_OK : APBF_ERR_HIP;
}

}  // extern "C"
