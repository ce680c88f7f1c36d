// neighborhood_binary_search (neighborhood_binary_search.cpp:22-75,
// neighborhood_binary_search.comp:109-276): particles in the order of the
// 96-bit z-order code of their raw fixed-point positions, and per particle
// the 27 cells of side 2^(digits/3) around it found by lower-bound searches
// over the sorted codes.
//
// The search reproduces the shader's: the 31-step bisection clamped to the
// list end, started inside the searching invocation's 256-particle workgroup
// whenever the workgroup's first and last codes bracket the target (the
// shader's shared-memory shortcut). Workgroup entries past the list end are
// never written by the shader; they count as the largest code here. The walk
// visits the cells in the shader's order (x fastest, then y, then z) and keeps
// d <= range (a NaN distance or range keeps nothing), so every row comes out
// in the order a sequential run appends it.
#include "scene.h"

namespace apbf {

struct C96 {
    uint32_t v[3];
};

__device__ __forceinline__ uint32_t code_section(int4 p, uint32_t section) {
    const uint32_t u[3] = {(uint32_t)p.x, (uint32_t)p.y, (uint32_t)p.z};
    uint32_t r = 0;
#pragma unroll
    for (uint32_t k = 0; k < 32u; ++k) {
        const uint32_t b = section * 32u + k;
        r |= ((u[b % 3u] >> (b / 3u)) & 1u) << k;
    }
    return r;
}

__device__ __forceinline__ bool gt96(const C96& a, const C96& b) {
    if (a.v[2] != b.v[2]) return a.v[2] > b.v[2];
    if (a.v[1] != b.v[1]) return a.v[1] > b.v[1];
    return a.v[0] > b.v[0];
}
__device__ __forceinline__ C96 plus96(const C96& a, const C96& b) {  // :98-103
    C96 r{{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]}};
    const bool y = r.v[0] < a.v[0];
    const bool z = r.v[1] < a.v[1] || (y && r.v[1] == 0xFFFFFFFFu);
    r.v[1] += y ? 1u : 0u;
    r.v[2] += z ? 1u : 0u;
    return r;
}
__device__ __forceinline__ C96 minus96(const C96& a, const C96& b) {  // :105-109
    const bool y = a.v[0] < b.v[0];
    const bool z = a.v[1] < b.v[1] || (y && a.v[1] == b.v[1]);
    return C96{{a.v[0] - b.v[0], a.v[1] - b.v[1] - (y ? 1u : 0u), a.v[2] - b.v[2] - (z ? 1u : 0u)}};
}
__device__ __forceinline__ C96 shl96(const C96& a, uint32_t d) {  // leftShift96 for 0 < d < 32
    return C96{{a.v[0] << d, (a.v[1] << d) | (a.v[0] >> (32u - d)), (a.v[2] << d) | (a.v[1] >> (32u - d))}};
}
__device__ __forceinline__ C96 and96(const C96& a, const C96& b) {
    return C96{{a.v[0] & b.v[0], a.v[1] & b.v[1], a.v[2] & b.v[2]}};
}
__device__ __forceinline__ C96 or96(const C96& a, const C96& b) {
    return C96{{a.v[0] | b.v[0], a.v[1] | b.v[1], a.v[2] | b.v[2]}};
}
__device__ __forceinline__ C96 not96(const C96& a) { return C96{{~a.v[0], ~a.v[1], ~a.v[2]}}; }

// uint(ceil(log2(x))) with the exact logarithm: 0 for x <= 1 or NaN, UINT_MAX
// for +inf (saturating conversion).
__device__ __forceinline__ uint32_t ceil_log2_u(float x) {
    if (!(x > 1.0f)) return 0u;
    const uint32_t b = __builtin_bit_cast(uint32_t, x);
    const uint32_t e = b >> 23;  // x > 1: normal or inf
    if (e == 0xFFu) return 0xFFFFFFFFu;
    return (e - 127u) + ((b & 0x7FFFFFu) ? 1u : 0u);
}

struct Codes {
    const uint32_t *c0, *c1, *c2;
    __device__ C96 at(uint32_t k) const { return C96{{c0[k], c1[k], c2[k]}}; }
};

__device__ uint32_t lower_bound96(const Codes& cs, uint32_t n, const C96& code, uint32_t id) {
    const C96 ones{{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}};
    const uint32_t b0 = (id / 256u) * 256u;
    auto block_code = [&](uint32_t k) { return k < n ? cs.at(k) : ones; };
    uint32_t lo = 0u, hi = n;
    if (!gt96(block_code(b0), code) && !gt96(code, block_code(b0 + 255u))) {
        lo = b0;
        hi = b0 + 256u;
    }
    uint32_t mid = (lo + hi) / 2u;
    for (int it = 0; it < 31; ++it) {
        const C96 c = (mid - b0 < 256u) ? block_code(mid) : cs.at(mid);
        if (gt96(code, c))
            lo = mid + 1u;
        else
            hi = mid;
        mid = min((lo + hi) / 2u, n - 1u);
    }
    return lo;
}

__global__ __launch_bounds__(BLOCK) void k_code_gather(const int4* __restrict__ pos, const uint32_t* __restrict__ idx,
                                                       const uint32_t* __restrict__ counts, uint32_t section,
                                                       uint32_t* __restrict__ out) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(j, n) out[j] = code_section(pos[idx ? idx[j] : j], section);
}

template <bool FILL>
__global__ __launch_bounds__(BLOCK) void k_bsearch_walk(const int4* __restrict__ pos, const float* __restrict__ kw,
                                                        Codes cs, const uint32_t* __restrict__ counts,
                                                        float range_scale, uint32_t* __restrict__ cnt,
                                                        const uint32_t* __restrict__ row, uint32_t* __restrict__ col,
                                                        uint32_t pcap) {
    const uint32_t n = counts[C_N];
    const C96 xm{{011111111111u, 022222222222u, 04444444444u}};
    const C96 ym{{022222222222u, 04444444444u, 011111111111u}};
    const C96 zm{{04444444444u, 011111111111u, 022222222222u}};
    GRID_STRIDE(id, n) {
        const float range = kw[id] * range_scale;
        const C96 code = cs.at(id);
        const uint32_t digits = ceil_log2_u(range * POS_RES) * 3u;
        C96 mask;
        mask.v[0] = (digits < 32u ? 1u << digits : 0u) - 1u;
        mask.v[1] = (digits < 64u ? 1u << (max(digits, 32u) - 32u) : 0u) - 1u;
        mask.v[2] = (digits < 96u ? 1u << (max(digits, 64u) - 64u) : 0u) - 1u;
        const C96 center{{code.v[0] - (code.v[0] & mask.v[0]), code.v[1] - (code.v[1] & mask.v[1]),
                          code.v[2] - (code.v[2] & mask.v[2])}};
        const C96 step = plus96(mask, C96{{1u, 0u, 0u}});
        C96 X[3], Y[3], Z[3];
        X[0] = and96(minus96(and96(center, xm), step), xm);
        X[2] = and96(plus96(or96(center, not96(xm)), step), xm);
        Y[0] = and96(minus96(and96(center, ym), shl96(step, 1u)), ym);
        Y[2] = and96(plus96(or96(center, not96(ym)), shl96(step, 1u)), ym);
        Z[0] = and96(minus96(and96(center, zm), shl96(step, 2u)), zm);
        Z[2] = and96(plus96(or96(center, not96(zm)), shl96(step, 2u)), zm);
        X[1] = and96(center, xm);
        Y[1] = and96(center, ym);
        Z[1] = and96(center, zm);
        const v3f p = pos_world(pos[id]);
        const float t = sqrt_le_bound(range);  // distance <= range  <=>  d2 <= t
        uint32_t c = 0;
        const uint32_t base = FILL ? row[id] : 0u;
        for (uint32_t cell = 0; cell < 27u; ++cell) {
            const C96 cc = or96(or96(X[cell % 3u], Y[(cell / 3u) % 3u]), Z[cell / 9u]);
            const C96 last = or96(cc, mask);
            for (uint32_t j = lower_bound96(cs, n, cc, id); j < n && !gt96(cs.at(j), last); ++j) {
                const v3f d = p - pos_world(pos[j]);
                if (j != id && dot(d, d) <= t) {
                    if (FILL && base + c < pcap) col[base + c] = j;
                    ++c;
                }
            }
        }
        if (!FILL) cnt[id] = c;
    }
    if (!FILL && blockIdx.x == 0 && threadIdx.x == 0) cnt[n] = 0;
}

void launch_code_gather(const int4* pos, const uint32_t* idx, const uint32_t* counts, uint32_t cap, uint32_t section,
                        uint32_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_code_gather, dim3(stride_grid(cap)), dim3(BLOCK), 0, s, pos, idx, counts, section, out);
}

void launch_bsearch_walk(bool fill, const int4* pos, const float* kw, const uint32_t* c0, const uint32_t* c1,
                         const uint32_t* c2, const uint32_t* counts, uint32_t cap, float range_scale, uint32_t* cnt,
                         const uint32_t* row, uint32_t* col, uint32_t pcap, hipStream_t s) {
    const Codes cs{c0, c1, c2};
    if (fill)
        hipLaunchKernelGGL(k_bsearch_walk<true>, dim3(stride_grid(cap, 8192)), dim3(BLOCK), 0, s, pos, kw, cs, counts,
                           range_scale, cnt, row, col, pcap);
    else
        hipLaunchKernelGGL(k_bsearch_walk<false>, dim3(stride_grid(cap, 8192)), dim3(BLOCK), 0, s, pos, kw, cs,
                           counts, range_scale, cnt, row, col, pcap);
}

// neighborhood_binary_search.cpp:22-75 on the scene.
void Scene::neighborhood_bs(float range_scale) {
    hipStream_t s = stream;
    // chained stable sorts by code sections 0, 1, 2 (:44-48)
    iota_u32(vals, counts + C_N, cap, s);
    launch_code_gather(pos[cur], nullptr, counts, cap, 0u, keys, s);
    radix_sort_pairs(keys, vals, counts + C_N, cap, 0xFFFFFFFFu, skeys, perm, ws_mem, s, sort_shape);
    launch_code_gather(pos[cur], perm, counts, cap, 1u, keys, s);
    radix_sort_pairs(keys, perm, counts + C_N, cap, 0xFFFFFFFFu, skeys, vals, ws_mem, s, sort_shape);
    launch_code_gather(pos[cur], vals, counts, cap, 2u, keys, s);
    radix_sort_pairs(keys, vals, counts + C_N, cap, 0xFFFFFFFFu, skeys, perm, ws_mem, s, sort_shape);
    // particles into code order (:50-51)
    ParticleArrays src = arrays(cur), dst = arrays(1 - cur);
    bool tk = tbegin(K_PERMUTE);
    launch_permute(perm, counts, cap, src, dst, s);
    tend(tk);
    cur = 1 - cur;
    remap_transfers_after_permute();
    // codes of the sorted particles (:55-57), then the pair build
    uint32_t *c0 = keys, *c1 = vals, *c2 = skeys;
    launch_code_gather(pos[cur], nullptr, counts, cap, 0u, c0, s);
    launch_code_gather(pos[cur], nullptr, counts, cap, 1u, c1, s);
    launch_code_gather(pos[cur], nullptr, counts, cap, 2u, c2, s);
    launch_set_n_plus_1(counts, s);
    tk = tbegin(K_WALK_COUNT);
    launch_bsearch_walk(false, pos[cur], kernel_width[cur], c0, c1, c2, counts, cap, range_scale, cnt, nullptr,
                        nullptr, pcap, s);
    tend(tk);
    scan_u32(cnt, row, counts + C_N_PLUS_1, cap + 1, false, nullptr, ws_mem, s);
    launch_pairs_total(counts, row, pcap, s);
    tk = tbegin(K_WALK_FILL);
    launch_bsearch_walk(true, pos[cur], kernel_width[cur], c0, c1, c2, counts, cap, range_scale, nullptr, row, col,
                        pcap, s);
    tend(tk);
    pairs_valid = true;
}

}  // namespace apbf
