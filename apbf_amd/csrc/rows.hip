// Per-particle kernels over CSR neighbour rows, with a group of L lanes of a
// 64-lane wavefront cooperating on one row (L = 16 by default): the lanes read
// consecutive entries of the row (coalesced) and gather the neighbours'
// positions, and the row totals come from xor-shuffle reductions inside the
// group. Every per-pair quantity the reference accumulates is a truncated
// integer (incompressibility_1.comp:55-59, incompressibility_3.comp:66,
// find_split_and_merge_1.comp:30-31), so the reduction order does not change
// any bit of the result.
//
// The incompressibility iteration (incompressibility.cpp:12-45) is one kernel
// per iteration plus a light apply pass:
//   k_solve_rows: incompressibility_0/1 sums over row i, incompressibility_2
//     (lambda, boundariness), then the position shifts of
//     incompressibility_2.comp:106-109 (particle i) and _3.comp:62-66 (every
//     neighbour j of row i, with the gradient of the first loop), summed per
//     block in LDS and added with int32 atomics into a per-particle
//     accumulator, and the boundariness filter of _3.comp:41-46;
//   k_solve_apply: position += accumulator (and the box collision that opens
//     the next solver iteration).
// Positions are read-only during k_solve_rows, as in the reference, where
// _1 and _3 use gradients of the positions _1 saw.
#include "scene.h"

namespace apbf {

constexpr int ROW_LANES = 16;

// Group reductions (every lane of the wave active). Inside a 16-lane row of
// the wave they use DPP lane moves, which read a neighbour lane's register in
// the same VALU instruction: swap of adjacent lanes, swap of adjacent lane
// pairs, the mirror inside 8 lanes, the mirror inside 16 lanes. After each
// step every lane holds the reduction over its 2, 4, 8, 16 lanes. Wider
// groups continue with xor shuffles. Sums of integers, minima and maxima do
// not depend on the order of the operands, so the results equal those of
// any other reduction order. The moves are written as DPP updates of a zero
// with bound_ctrl set (every source lane of these patterns lies in the same
// row, so neither takes effect), the form the compiler folds into the
// integer add or minimum that consumes the move: one instruction per step.
template <int L, typename Op>
__device__ __forceinline__ uint32_t greduce(uint32_t v, Op op) {
    static_assert(L >= 1 && L <= 64 && (L & (L - 1)) == 0, "group of a power of two lanes");
    constexpr int QUAD_XOR1 = 0xB1, QUAD_XOR2 = 0x4E, ROW_HALF_MIRROR = 0x141, ROW_MIRROR = 0x140;
    auto dpp = [](uint32_t x, auto ctrl) {
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, decltype(ctrl)::value, 0xF, 0xF, true);
    };
    if (L >= 2) v = op(v, dpp(v, std::integral_constant<int, QUAD_XOR1>{}));
    if (L >= 4) v = op(v, dpp(v, std::integral_constant<int, QUAD_XOR2>{}));
    if (L >= 8) v = op(v, dpp(v, std::integral_constant<int, ROW_HALF_MIRROR>{}));
    if (L >= 16) v = op(v, dpp(v, std::integral_constant<int, ROW_MIRROR>{}));
    if (L >= 32) v = op(v, (uint32_t)__shfl_xor((int)v, 16, 64));
    if (L >= 64) v = op(v, (uint32_t)__shfl_xor((int)v, 32, 64));
    return v;
}
template <int L>
__device__ __forceinline__ uint32_t gsum(uint32_t v) {
    return greduce<L>(v, [](uint32_t a, uint32_t b) { return a + b; });
}
template <int L>
__device__ __forceinline__ uint32_t gmin(uint32_t v) {
    return greduce<L>(v, [](uint32_t a, uint32_t b) { return b < a ? b : a; });
}
template <int L>
__device__ __forceinline__ float gmaxf(float v) {
    return __builtin_bit_cast(float, greduce<L>(__builtin_bit_cast(uint32_t, v), [](uint32_t a, uint32_t b) {
                                  return __builtin_bit_cast(uint32_t, fmax_g(__builtin_bit_cast(float, a),
                                                                             __builtin_bit_cast(float, b)));
                              }));
}
// Mask of this lane's group within the wave's 64-bit ballot.
template <int L>
__device__ __forceinline__ uint64_t group_mask() {
    const int lane = threadIdx.x & 63;
    const uint64_t m = (L == 64) ? ~0ull : ((1ull << L) - 1ull);
    return m << (lane & ~(L - 1));
}

#define ROW_LOOP_BEGIN(L)                                                                          \
    const uint32_t n = counts[C_N];                                                                \
    const int gl = threadIdx.x & (L - 1);                                                          \
    const uint32_t groups = BLOCK / L;                                                             \
    for (uint32_t base = blockIdx.x * groups; base < n; base += gridDim.x * groups) {              \
        const uint32_t i = base + threadIdx.x / L;                                                 \
        const bool active = i < n;

#define ROW_LOOP_END }

// Kernel evaluation with the height / gradient kernel ids fixed at compile
// time (KH, KG >= 0) or read at run time (-1).
template <int KH, int KG>
__device__ __forceinline__ void kernel_eval_k(int hid, int gid, float dim, v3f r, float kw, GaussConst g,
                                              bool need_w, bool need_g, float* w, v3f* grad) {
    kernel_eval(KH >= 0 ? KH : hid, KG >= 0 ? KG : gid, dim, r, kw, g, need_w, need_g, w, grad);
}

// Gradients of the first C entries of a lane's share of the row stay in
// registers between the two neighbour loops of k_solve_rows; longer rows
// recompute the rest. 6 keeps the kernel at 79 VGPRs, 6 waves per SIMD,
// which hide the latency of the row's dependent loads better than a larger
// cache saves recomputation (8: 93 VGPRs, 5 waves, 1-4% slower).
constexpr int SOLVE_CACHE = 6;

// The position shifts of k_solve_rows are summed per block in a table in LDS
// before they reach the per-particle accumulator. A block takes 2 or 4
// groups of consecutive rows, and consecutive particles are neighbours (the
// particles are sorted by cell), so the rows share most neighbours and one
// global atomic per neighbour and component replaces one per pair; where
// many rows push the same particles (a fluid at rest) this removes the
// contention of the global atomics. Integer sums wrap modulo 2^32 in any
// order, so the accumulator ends with the same bits. The table maps a
// particle id to slot (id mod AGG_SLOTS) and probes the next AGG_PROBES - 1
// slots; an id that finds no free slot goes straight to the accumulator.
constexpr int AGG_SLOTS = 1024;
constexpr int AGG_PROBES = 4;
constexpr uint32_t AGG_STEPS = 4;  // row groups per table (64 rows at 16 lanes per row); long rows: 2

// Gradients of a lane's row entries after the SOLVE_CACHE in registers: the
// next SOLVE_SPILL go to LDS (one column per thread, so no thread reads
// another's), the rest are recomputed in the shift loop.
constexpr int SOLVE_SPILL = 3;
constexpr uint32_t AGG_EMPTY = 0xFFFFFFFFu;

struct ShiftTable {
    uint32_t* key;  // AGG_SLOTS particle ids (AGG_EMPTY: free)
    int32_t* sum;   // 3 x AGG_SLOTS: x sums, y sums, z sums
    int* acc;       // per-particle accumulator (x, y, z, unused)

    __device__ __forceinline__ void clear() {
        for (int t = threadIdx.x; t < AGG_SLOTS; t += BLOCK) {
            key[t] = AGG_EMPTY;
            sum[t] = 0;
            sum[AGG_SLOTS + t] = 0;
            sum[2 * AGG_SLOTS + t] = 0;
        }
    }
    __device__ __forceinline__ void add(uint32_t j, int32_t dx, int32_t dy, int32_t dz) {
        if ((dx | dy | dz) == 0) return;
#pragma unroll
        for (int p = 0; p < AGG_PROBES; ++p) {
            const uint32_t t = (j + (uint32_t)p) & (AGG_SLOTS - 1u);
            const uint32_t k = atomicCAS(&key[t], AGG_EMPTY, j);
            if (k == AGG_EMPTY || k == j) {
                atomicAdd(&sum[t], dx);
                atomicAdd(&sum[AGG_SLOTS + t], dy);
                atomicAdd(&sum[2 * AGG_SLOTS + t], dz);
                return;
            }
        }
        if (dx) atomicAdd(&acc[4 * j + 0], dx);
        if (dy) atomicAdd(&acc[4 * j + 1], dy);
        if (dz) atomicAdd(&acc[4 * j + 2], dz);
    }
    // Adds the table to the accumulator and empties it; the block's threads
    // must have finished their add() calls (a barrier before) and must not
    // start new ones until every thread is done (a barrier after).
    __device__ __forceinline__ void flush() {
        for (int t = threadIdx.x; t < AGG_SLOTS; t += BLOCK) {
            const uint32_t j = key[t];
            if (j == AGG_EMPTY) continue;
            const int32_t x = sum[t], y = sum[AGG_SLOTS + t], z = sum[2 * AGG_SLOTS + t];
            if (x) atomicAdd(&acc[4 * j + 0], x);
            if (y) atomicAdd(&acc[4 * j + 1], y);
            if (z) atomicAdd(&acc[4 * j + 2], z);
            key[t] = AGG_EMPTY;
            sum[t] = 0;
            sum[AGG_SLOTS + t] = 0;
            sum[2 * AGG_SLOTS + t] = 0;
        }
    }
};

// STEPS: row groups a block takes at a time and sums in its table before a
// flush. Long rows bring more distinct neighbours per row group, and a table
// flushed every 32 rows overflows less when a pass takes rows of every length
// (water drop, 100-160 entries per row: 0.75 against 0.80 ms per launch);
// short rows gain from fewer flushes (large pool, 33 entries per row: 4.04
// against 4.31 ms), and so does the water drop once its rows above 144
// entries go to the whole-wave pass (0.621 against 0.631 ms per iteration).
// The passes over the long-row lists flush after every row group (8 rows
// at 32 lanes per row, 4 at a wavefront per row).
// Six waves per SIMD (80 VGPRs, 25.6 KB of LDS per block) hide the latency
// of the row's dependent loads.
template <int L, int KH, int KG, int C, uint32_t STEPS>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void k_solve_rows(
    const int4* __restrict__ pos, const float* __restrict__ inv_mass, const float* __restrict__ radius,
    const float* __restrict__ kw, const float4* __restrict__ gc, const uint32_t* __restrict__ counts,
    const uint32_t* __restrict__ row, const uint32_t* __restrict__ col, uint32_t pcap, SolverArgs a,
    float* __restrict__ boundariness, float* __restrict__ lambda, int* __restrict__ acc) {
    if (a.skip && *a.skip) return;  // the same word for every block
    __shared__ double s_err;
    __shared__ uint32_t s_key[AGG_SLOTS];
    __shared__ int32_t s_sum[3 * AGG_SLOTS];
    __shared__ float s_g[3][SOLVE_SPILL][BLOCK];
    ShiftTable tab{s_key, s_sum, acc};
    if (threadIdx.x == 0) s_err = 0.0;
    tab.clear();
    __syncthreads();
    // A block takes STEPS consecutive groups of BLOCK / L rows, then
    // flushes the table; the trip counts are uniform per block, so every
    // thread meets the barriers.
    const uint32_t n = a.list ? counts[a.list_slot] : counts[C_N];
    const int gl = threadIdx.x & (L - 1);
    constexpr uint32_t groups = BLOCK / L;
    constexpr uint32_t span = groups * STEPS;
    for (uint32_t base = blockIdx.x * span; base < n; base += gridDim.x * span) {
        for (uint32_t step = 0; step < STEPS; ++step) {
            const uint32_t slot = base + step * groups + threadIdx.x / L;
            bool active = slot < n;
            const uint32_t i = (a.list && active) ? a.list[slot] : slot;
            int4 pi = make_int4(0, 0, 0, 0);
            float im = 1.0f, kwi = 1.0f, ird = 1.0f;
            uint32_t self = 0;
            GaussConst g{1.0f, 1.0f};
            uint32_t rb = 0, re = 0;
            if (active) {
                rb = row[i];
                re = min(row[i + 1], pcap);
                // a row longer than split_len is left to the pass over the long-row list
                if (a.split_len && re > rb && re - rb > a.split_len) {
                    active = false;
                    rb = re = 0;
                }
            }
            if (active) {
                pi = pos[i];
                im = inv_mass[i];
                kwi = kw[i];
                const float4 g4 = gc[i];
                g = GaussConst{g4.x, g4.y};
                ird = g4.z;
                self = __builtin_bit_cast(uint32_t, g4.w);
            }
            // incompressibility_1.comp:43-68 over this lane's share of row i
            uint32_t dens = 0, sq = 0, s0 = 0, s1 = 0, s2 = 0, m0 = 0, m1 = 0, m2 = 0, m3 = 0;
            uint32_t cj[C];
            v3f cg[C];
            auto pair1 = [&](const int4 pj, const float imN, v3f& g_out) {
                const v3f diff = vec_i(isub(pj.x, pi.x), isub(pj.y, pi.y), isub(pj.z, pi.z)) / POS_RES;
                float w;
                v3f grad;
                kernel_eval_k<KH, KG>(a.hid, a.gid, a.dim, diff, kwi, g, true, true, &w, &grad);
                const float gg = dot(grad, grad);
                float wq, ggq;
                v3f wg;
                // Inverse masses are powers of two unless a merge mixed two masses.
                // Dividing by a power of two whose reciprocal is a normal float is the
                // exact scaling by that reciprocal (one rounding of the same real
                // value). Other divisors of 2^-40 to 2^40 take quotient_by_recip, the
                // IEEE quotient for dividends of at least 2^-80; smaller dividends
                // give quotients below 2^-39 either way, whose products with
                // INCOMP_RES truncate to 0. So the integers below are those of the
                // IEEE quotients in every case (finite w and gg; otherwise the
                // full division runs).
                const uint32_t ib = __builtin_bit_cast(uint32_t, imN);
                const uint32_t ie = (ib >> 23) & 0xFFu;
                if (__all((ib & 0x7FFFFFu) == 0u && ie >= 1u && ie <= 253u)) {
                    const float rim = __builtin_bit_cast(float, (ib & 0x80000000u) | ((254u - ie) << 23));
                    wq = w * rim;
                    wg = grad * rim;
                    ggq = gg * rim;
                } else if (__all(ie >= 87u && ie <= 167u && w < __builtin_inff() && gg < __builtin_inff())) {
                    const float y = recip_in_range(imN);
                    wq = quotient_by_recip(w, imN, y);
                    wg = mk(quotient_by_recip(grad.x, imN, y), quotient_by_recip(grad.y, imN, y),
                            quotient_by_recip(grad.z, imN, y));
                    ggq = quotient_by_recip(gg, imN, y);
                } else {
                    wq = w / imN;
                    wg = grad / imN;
                    ggq = gg / imN;
                }
                dens += to_u32(wq * INCOMP_RES);
                s0 += (uint32_t)to_i32(wg.x * INCOMP_RES);
                s1 += (uint32_t)to_i32(wg.y * INCOMP_RES);
                s2 += (uint32_t)to_i32(wg.z * INCOMP_RES);
                sq += to_u32(ggq * INCOMP_RES);
                if (a.method == 1) {
                    const v3f wd = vec_i(isub(pj.x, pi.x), isub(pj.y, pi.y), isub(pj.z, pi.z)) / (imN * kwi);
                    m0 += (uint32_t)to_i32(wd.x);
                    m1 += (uint32_t)to_i32(wd.y);
                    m2 += (uint32_t)to_i32(wd.z);
                    m3 += (uint32_t)to_i32(INCOMP_RES / imN);
                }
                g_out = grad;
            };
            // The cached part of the row goes in batches of B entries per lane: the
            // batch's list entries, then its neighbours' positions and masses, are
            // loaded before any arithmetic, so a lane waits for memory twice per
            // batch instead of twice per entry. Lanes past the row end load the row's
            // last entry and discard it.
            constexpr int B = (C % 8 == 0) ? 8 : ((C % 4 == 0) ? 4 : C);
            static_assert(C % B == 0, "the cached entries split into whole batches");
        #pragma unroll
            for (int t = 0; t < C; ++t) {
                cj[t] = 0;
                cg[t] = mk(0, 0, 0);
            }
            if (rb < re) {
                const uint32_t klast = re - 1u;
        #pragma unroll
                for (int t0 = 0; t0 < C; t0 += B) {
                    int4 bp[B];
                    float bm[B];
        #pragma unroll
                    for (int t = 0; t < B; ++t) cj[t0 + t] = col[min(rb + gl + (uint32_t)(t0 + t) * L, klast)];
        #pragma unroll
                    for (int t = 0; t < B; ++t) {
                        bp[t] = pos[cj[t0 + t]];
                        bm[t] = inv_mass[cj[t0 + t]];
                    }
        #pragma unroll
                    for (int t = 0; t < B; ++t)
                        if (rb + gl + (uint32_t)(t0 + t) * L < re) pair1(bp[t], bm[t], cg[t0 + t]);
                }
            }
            {
                uint32_t idx = 0;
                for (uint32_t k = rb + gl + (uint32_t)C * L; k < re; k += L, ++idx) {
                    const uint32_t j = col[k];
                    v3f gr;
                    pair1(pos[j], inv_mass[j], gr);
                    if (idx < (uint32_t)SOLVE_SPILL) {
                        s_g[0][idx][threadIdx.x] = gr.x;
                        s_g[1][idx][threadIdx.x] = gr.y;
                        s_g[2][idx][threadIdx.x] = gr.z;
                    }
                }
            }
            dens = gsum<L>(dens);
            sq = gsum<L>(sq);
            s0 = gsum<L>(s0);
            s1 = gsum<L>(s1);
            s2 = gsum<L>(s2);
            if (a.method == 1) {
                m0 = gsum<L>(m0);
                m1 = gsum<L>(m1);
                m2 = gsum<L>(m2);
                m3 = gsum<L>(m3);
            }
            dens += self;  // incompressibility_0.comp:43: the particle's own contribution
            // incompressibility_2.comp:77-109 (every lane of the group computes the same values)
            const int32_t gs0 = (int32_t)s0, gs1 = (int32_t)s1, gs2 = (int32_t)s2;
            const float densf = (float)dens / INCOMP_RES;
            const v3f wgs = vec_i(gs0, gs1, gs2) / INCOMP_RES;
            float sgs = (float)sq / INCOMP_RES;
            const v3f selfGrad = (-wgs) * ird;
            const float wgss = dot(wgs, wgs);
            const float sgl = __builtin_sqrtf(wgss) * ird;
            sgs += wgss * im;
            const float under = 1.0f - densf * ird;
            float b = active ? boundariness[i] : 0.0f;
            if (a.update_b && !a.lam_override) {
                float bb;
                if (a.method == 1) {
                    const float totalMass = (float)(int32_t)m3 / INCOMP_RES + 1.0f / im;
                    const float dev = length(vec_i((int32_t)m0, (int32_t)m1, (int32_t)m2)) / (POS_RES * totalMass);
                    bb = dev * a.b_sgl_factor;
                } else {
                    bb = sgl * kwi * a.b_sgl_factor + fmax_g(0.0f, under) * a.b_under_factor;
                }
                bb = bb >= 1.0f ? 1.0f : 0.0f;
                b = fmin_g(1.0f, move_towards_add(b, bb, a.b_speed));
            }
            float lam = under / (ird * ird * (sgs + INCOMP_EPSILON));
            lam /= a.lambda_scale_base / ird * im;
            if (a.lam_override) lam = active ? a.lam_override[i] : 0.0f;
            if (active && gl == 0) {
                if (a.it_err || a.tol_sum) atomicAdd(&s_err, (double)fmax_g(0.0f, densf * ird - 1.0f));
                if (a.lam_acc) a.lam_acc[i] += fminf(lam, 0.0f);
                lambda[i] = lam;
                if (!(lam >= 0.0f)) {  // own shift; a NaN lambda shifts by ivec3(NaN) = 0
                    const v3f d = selfGrad * (lam * im * POS_RES);
                    tab.add(i, to_i32(d.x), to_i32(d.y), to_i32(d.z));
                }
            }
            // neighbour shifts (incompressibility_3.comp:62-66) and the boundariness
            // filter (:41-46), over this lane's share of row i, with the gradients of
            // the first loop (positions are unchanged in between)
            const bool do_shift = active && lam < 0.0f;
            const bool do_filter = active && a.method == 2 && !a.lam_override && __builtin_bit_cast(uint32_t, b) != 0u;
            bool hit = false;
            if (do_shift || do_filter) {
                const float sc = lam * POS_RES;
                const v3f empty = -normalize(vec_i(gs0, gs1, gs2));
                auto pair3 = [&](uint32_t j, v3f grad) {
                    if (do_filter && !hit) {
                        // cos(empty, grad) > 0.6 needs dot(empty, grad) > 0 and
                        // dot^2 > 0.36 |grad|^2; with a 3% margin over the rounding
                        // of either form, entries that fail the squared test cannot
                        // pass the normalised one, which then runs only for the rest
                        // (and for gradients too short for the squares to be exact).
                        const float dg = dot(empty, grad), gg = dot(grad, grad);
                        const bool maybe = !(gg >= 1e-30f) || (dg > 0.0f && dg * dg > 0.35f * gg);
                        if (maybe && dot(empty, normalize(grad)) > 0.6f) hit = true;
                    }
                    if (do_shift) {
                        const v3f d = grad * sc;
                        tab.add(j, to_i32(d.x), to_i32(d.y), to_i32(d.z));
                    }
                };
        #pragma unroll
                for (int t = 0; t < C; ++t) {
                    const uint32_t k = rb + gl + (uint32_t)t * L;
                    if (k < re) pair3(cj[t], cg[t]);
                }
                uint32_t idx = 0;
                for (uint32_t k = rb + gl + (uint32_t)C * L; k < re; k += L, ++idx) {
                    const uint32_t j = col[k];
                    v3f grad;
                    if (idx < (uint32_t)SOLVE_SPILL) {
                        grad = mk(s_g[0][idx][threadIdx.x], s_g[1][idx][threadIdx.x], s_g[2][idx][threadIdx.x]);
                    } else {
                        const int4 pj = pos[j];
                        const v3f diff = vec_i(isub(pj.x, pi.x), isub(pj.y, pi.y), isub(pj.z, pi.z)) / POS_RES;
                        float w;
                        kernel_eval_k<KH, KG>(a.hid, a.gid, a.dim, diff, kwi, g, false, true, &w, &grad);
                    }
                    pair3(j, grad);
                }
            }
            const uint64_t ballot = __ballot(hit);
            if (ballot & group_mask<L>()) b = 0.0f;
            if (active && gl == 0) boundariness[i] = b;
        }  // step
        __syncthreads();
        tab.flush();
        __syncthreads();
    }  // base
    if (a.it_err || a.tol_sum) {
        __syncthreads();
        if (threadIdx.x == 0 && s_err != 0.0) {
            if (a.it_err) atomicAdd(a.it_err, s_err / (double)counts[C_N]);
            if (a.tol_sum) atomicAdd(a.tol_sum, s_err);
        }
    }
}

// Warm-start features of particle i (16 floats, learned lambda warm start):
// the sums of incompressibility_1 over row i and the quantities
// incompressibility_2 derives from them (density ratio, self-gradient length,
// lambda of a cold start), plus neighbour count, previous total lambda, radius
// ratio to the neighbours, kernel width / radius, boundariness, rest-density
// factor, log radius, mean neighbour distance / kernel width and the distance
// moved per substep / radius.
template <int L, int KH, int KG>
__global__ __launch_bounds__(BLOCK) void k_features_rows(const int4* __restrict__ pos,
                                                         const float4* __restrict__ vel,
                                                         const float* __restrict__ inv_mass,
                                                         const float* __restrict__ radius,
                                                         const float* __restrict__ kw,
                                                         const float4* __restrict__ gc,
                                                         const float* __restrict__ boundariness,
                                                         const float* __restrict__ lam_prev,
                                                         const uint32_t* __restrict__ counts,
                                                         const uint32_t* __restrict__ row,
                                                         const uint32_t* __restrict__ col, uint32_t pcap,
                                                         SolverArgs a, float dt, float* __restrict__ feat) {
    ROW_LOOP_BEGIN(L)
    int4 pi = make_int4(0, 0, 0, 0);
    float im = 1.0f, kwi = 1.0f, ird = 1.0f;
    uint32_t self = 0;
    GaussConst g{1.0f, 1.0f};
    uint32_t rb = 0, re = 0;
    if (active) {
        pi = pos[i];
        im = inv_mass[i];
        kwi = kw[i];
        const float4 g4 = gc[i];
        g = GaussConst{g4.x, g4.y};
        ird = g4.z;
        self = __builtin_bit_cast(uint32_t, g4.w);
        rb = row[i];
        re = min(row[i + 1], pcap);
    }
    uint32_t dens = 0, sq = 0, s0 = 0, s1 = 0, s2 = 0;
    float rsum = 0.0f, dsum = 0.0f;
    auto entry = [&](const int4 pj, const float imN, const float rj) {
        const v3f diff = vec_i(isub(pj.x, pi.x), isub(pj.y, pi.y), isub(pj.z, pi.z)) / POS_RES;
        float w;
        v3f grad;
        kernel_eval_k<KH, KG>(a.hid, a.gid, a.dim, diff, kwi, g, true, true, &w, &grad);
        const float gg = dot(grad, grad);
        float wq, ggq;
        v3f wg;
        // the quotients by the neighbour's inverse mass, as in k_solve_rows
        const uint32_t ib = __builtin_bit_cast(uint32_t, imN);
        const uint32_t ie = (ib >> 23) & 0xFFu;
        if (__all((ib & 0x7FFFFFu) == 0u && ie >= 1u && ie <= 253u)) {
            const float rim = __builtin_bit_cast(float, (ib & 0x80000000u) | ((254u - ie) << 23));
            wq = w * rim;
            wg = grad * rim;
            ggq = gg * rim;
        } else {
            wq = w / imN;
            wg = grad / imN;
            ggq = gg / imN;
        }
        dens += to_u32(wq * INCOMP_RES);
        s0 += (uint32_t)to_i32(wg.x * INCOMP_RES);
        s1 += (uint32_t)to_i32(wg.y * INCOMP_RES);
        s2 += (uint32_t)to_i32(wg.z * INCOMP_RES);
        sq += to_u32(ggq * INCOMP_RES);
        rsum += rj;
        dsum += sqrt_in_range(dot(diff, diff));  // length(diff): 0 or in [2^-36, 2^34]
    };
    // batches of 4 entries per lane, loaded before the arithmetic; each lane
    // still adds its entries in row order
    uint32_t k = rb + gl;
    for (; k + 3u * L < re; k += 4u * L) {
        uint32_t jb[4];
        int4 pb[4];
        float mb[4], rbj[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) jb[t] = col[k + (uint32_t)t * L];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            pb[t] = pos[jb[t]];
            mb[t] = inv_mass[jb[t]];
            rbj[t] = radius[jb[t]];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) entry(pb[t], mb[t], rbj[t]);
    }
    for (; k < re; k += L) {
        const uint32_t j = col[k];
        entry(pos[j], inv_mass[j], radius[j]);
    }
    dens = gsum<L>(dens) + self;
    sq = gsum<L>(sq);
    s0 = gsum<L>(s0);
    s1 = gsum<L>(s1);
    s2 = gsum<L>(s2);
#pragma unroll
    for (int m = L / 2; m > 0; m >>= 1) {
        rsum += __shfl_xor(rsum, m, 64);
        dsum += __shfl_xor(dsum, m, 64);
    }
    if (active && gl == 0) {
        const float c = (float)(re > rb ? re - rb : 0u);
        const float densf = (float)dens / INCOMP_RES;
        const v3f wgs = vec_i((int32_t)s0, (int32_t)s1, (int32_t)s2) / INCOMP_RES;
        const float wgss = dot(wgs, wgs);
        const float sgs = (float)sq / INCOMP_RES + wgss * im;
        const float under = 1.0f - densf * ird;
        float lam = under / (ird * ird * (sgs + INCOMP_EPSILON));
        lam /= a.lambda_scale_base / ird * im;
        const float r = radius[i];
        const float4 v = vel[i];
        float* f = feat + (size_t)i * 16;
        f[0] = c * (1.0f / 64.0f);
        f[1] = densf * ird;
        f[2] = under;
        f[3] = sqrtf(wgss) * ird;
        f[4] = (float)sq / INCOMP_RES * ird * ird;
        f[5] = ird * ird * (sgs + INCOMP_EPSILON);
        f[6] = lam;
        f[7] = fminf(lam, 0.0f);
        f[8] = lam_prev[i];
        f[9] = c > 0.0f ? r * c / rsum : 1.0f;
        f[10] = kwi / r;
        f[11] = boundariness[i];
        f[12] = ird;
        f[13] = log2f(r);
        f[14] = c > 0.0f ? dsum / (c * kwi) : 0.0f;
        f[15] = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z) * dt / r;
    }
    ROW_LOOP_END
}

// Applies the accumulated shifts, clears the accumulator, and (when `boxes`
// holds colliders) runs the box and sphere collision of the next solver
// iteration.
// With `action` set, the iteration's decision (SolveAction) selects between
// adding the shifts, leaving everything as it is, and putting back the
// positions `saved` before the warm-start shift (then the accumulator is
// cleared and the next iteration's colliders run as after a shift).
__global__ __launch_bounds__(BLOCK) void k_solve_apply(int4* __restrict__ pos, int4* __restrict__ acc,
                                                       const float* __restrict__ radius,
                                                       const uint32_t* __restrict__ counts, BoxSet boxes,
                                                       const uint32_t* __restrict__ action,
                                                       const int4* __restrict__ saved, SegmentTable seg) {
    const uint32_t act = action ? *action : (uint32_t)SOLVE_APPLY;
    if (act == SOLVE_SKIP) return;
    const uint32_t n = counts[C_N];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        int4 p = pos[j];
        const int4 d = acc[j];
        if (act == SOLVE_RESTORE) {
            p = saved[j];
        } else {
            p.x = iadd(p.x, d.x);
            p.y = iadd(p.y, d.y);
            p.z = iadd(p.z, d.z);
        }
        acc[j] = make_int4(0, 0, 0, 0);
        if (boxes.n) {
            if (seg.sid) {  // a packed batch: the particle's own scene's boxes and list position
                const uint32_t sc = seg.sid[j];
                const BoxSet& b = seg.boxes[sc];
                if (b.n) p = box_collide(j - seg.seg_start[sc], p, radius[j], b);
            } else {
                p = box_collide(j, p, radius[j], boxes);
            }
        }
        if (boxes.sphere) p = sphere_collide(j, p, radius[j], boxes);
        pos[j] = p;
    }
}

// Density error per particle (the sums of incompressibility_0/1), reduced to
// per-block partials: the sum of max(0, rho/rho0 - 1) in part_sum[block], the
// sum of |rho/rho0 - 1| in part_sum[gridDim.x + block], the max in part_max.
// Rows exist for particles below counts[C_ROWS].
template <int L, int KH>
__global__ __launch_bounds__(BLOCK) void k_density_rows(const int4* __restrict__ pos,
                                                        const float* __restrict__ inv_mass,
                                                        const float* __restrict__ radius,
                                                        const float* __restrict__ kw,
                                                        const float4* __restrict__ gc,
                                                        const uint32_t* __restrict__ counts,
                                                        const uint32_t* __restrict__ row,
                                                        const uint32_t* __restrict__ col, uint32_t pcap,
                                                        int hid, int gid, float dim, double* __restrict__ part_sum,
                                                        float* __restrict__ part_max) {
    __shared__ double ssum[BLOCK];
    __shared__ double sabs[BLOCK];
    __shared__ float smax[BLOCK];
    double acc_sum = 0.0, acc_abs = 0.0;
    float acc_max = 0.0f;
    const uint32_t rows = counts[C_ROWS];
    ROW_LOOP_BEGIN(L)
    int4 pi = make_int4(0, 0, 0, 0);
    float kwi = 1.0f, ird = 1.0f;
    uint32_t self = 0;
    GaussConst g{1.0f, 1.0f};
    uint32_t rb = 0, re = 0;
    if (active) {
        pi = pos[i];
        kwi = kw[i];
        const float4 g4 = gc[i];
        g = GaussConst{g4.x, g4.y};
        ird = g4.z;
        self = __builtin_bit_cast(uint32_t, g4.w);
        if (i < rows) {
            rb = row[i];
            re = min(row[i + 1], pcap);
        }
    }
    uint32_t dens = 0;
    v3f unused;
    auto pair = [&](const int4 pj, const float imN) {
        const v3f diff = vec_i(isub(pj.x, pi.x), isub(pj.y, pi.y), isub(pj.z, pi.z)) / POS_RES;
        float w;
        kernel_eval_k<KH, 0>(hid, gid, dim, diff, kwi, g, true, false, &w, &unused);
        const uint32_t ib = __builtin_bit_cast(uint32_t, imN);
        const uint32_t ie = (ib >> 23) & 0xFFu;
        float wq;
        if (__all((ib & 0x7FFFFFu) == 0u && ie >= 1u && ie <= 253u))  // exact scaling, as in k_solve_rows
            wq = w * __builtin_bit_cast(float, (ib & 0x80000000u) | ((254u - ie) << 23));
        else
            wq = w / imN;
        dens += to_u32(wq * INCOMP_RES);
    };
    // batches of four entries per lane, loads first (as in k_solve_rows)
    uint32_t k = rb + gl;
    for (; k + 3u * L < re; k += 4u * L) {
        uint32_t jb[4];
        int4 pb[4];
        float mb[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) jb[t] = col[k + (uint32_t)t * L];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            pb[t] = pos[jb[t]];
            mb[t] = inv_mass[jb[t]];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) pair(pb[t], mb[t]);
    }
    for (; k < re; k += L) {
        const uint32_t j = col[k];
        pair(pos[j], inv_mass[j]);
    }
    dens = gsum<L>(dens);
    if (active && gl == 0) {
        dens += self;
        const float dev = (float)dens / INCOMP_RES * ird - 1.0f;
        const float e = fmax_g(0.0f, dev);
        acc_sum += (double)e;
        acc_abs += (double)__builtin_fabsf(dev);
        acc_max = fmax_g(acc_max, e);
    }
    ROW_LOOP_END
    ssum[threadIdx.x] = acc_sum;
    sabs[threadIdx.x] = acc_abs;
    smax[threadIdx.x] = acc_max;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            ssum[threadIdx.x] += ssum[threadIdx.x + s];
            sabs[threadIdx.x] += sabs[threadIdx.x + s];
            smax[threadIdx.x] = fmax_g(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_sum[blockIdx.x] = ssum[0];
        part_sum[gridDim.x + blockIdx.x] = sabs[0];
        part_max[blockIdx.x] = smax[0];
    }
}

// find_split_and_merge_1.comp:20-31 and _2.comp:20-32 over row i: flood-fill
// boundary distance, shortest neighbour distance, and the last neighbour in
// row order at that distance (the last writer of the sequential schedule). A
// particle without neighbours keeps UINT_MAX and names itself.
template <int L>
__global__ __launch_bounds__(BLOCK) void k_fsm12_rows(const int4* __restrict__ pos,
                                                      const uint32_t* __restrict__ counts,
                                                      const uint32_t* __restrict__ row,
                                                      const uint32_t* __restrict__ col, uint32_t pcap,
                                                      const uint32_t* __restrict__ old_bd,
                                                      uint32_t* __restrict__ bd, uint32_t* __restrict__ nearest) {
    ROW_LOOP_BEGIN(L)
    int4 pi = make_int4(0, 0, 0, 0);
    uint32_t rb = 0, re = 0;
    if (active) {
        pi = pos[i];
        rb = row[i];
        re = min(row[i + 1], pcap);
    }
    uint32_t bmin = 0xFFFFFFFFu, best = 0xFFFFFFFFu, best_k = 0u;
    bool found = false;
    auto entry = [&](uint32_t k, const int4 pj, uint32_t obd) {
        // squared distances of fixed-point differences: 0 or in [1, 2^64], where
        // sqrt_in_range is the IEEE square root of length()
        const v3f dv = vec_i(isub(pj.x, pi.x), isub(pj.y, pi.y), isub(pj.z, pi.z));
        const uint32_t dist = to_u32(sqrt_in_range(dot(dv, dv)));
        const uint32_t cand = obd + dist;
        bmin = cand < bmin ? cand : bmin;
        if (!found || dist <= best) {  // later entries of the lane win ties
            best = dist;
            best_k = k;
            found = true;
        }
    };
    // batches of four entries per lane, loads first (as in k_solve_rows)
    uint32_t k = rb + gl;
    for (; k + 3u * L < re; k += 4u * L) {
        uint32_t jb[4], ob[4];
        int4 pb[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) jb[t] = col[k + (uint32_t)t * L];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            pb[t] = pos[jb[t]];
            ob[t] = old_bd[jb[t]];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) entry(k + (uint32_t)t * L, pb[t], ob[t]);
    }
    for (; k < re; k += L) {
        const uint32_t j = col[k];
        entry(k, pos[j], old_bd[j]);
    }
    bmin = gmin<L>(bmin);
    // group arg-min of (distance, -position): smallest distance, then the last position
#pragma unroll
    for (int m = L / 2; m > 0; m >>= 1) {
        const uint32_t ob = __shfl_xor(best, m, 64);
        const uint32_t ok = __shfl_xor(best_k, m, 64);
        const bool of = __shfl_xor((int)found, m, 64) != 0;
        if (of && (!found || ob < best || (ob == best && ok > best_k))) {
            best = ob;
            best_k = ok;
            found = true;
        }
    }
    if (active && gl == 0) {
        bd[i] = bmin;
        nearest[i] = found ? col[best_k] : i;
    }
    ROW_LOOP_END
}

// k_fsm12_rows and k_density_rows in one pass: after the solver loop the
// substep's error measurement and the split/merge find read the same rows at
// the same positions, so one walk over the rows loads them once. The grid is
// that of k_density_rows (its per-block partial sums).
template <int L, int KH>
__global__ __launch_bounds__(BLOCK) void k_fsm12_density_rows(
    const int4* __restrict__ pos, const float* __restrict__ inv_mass, const float* __restrict__ kw,
    const float4* __restrict__ gc, const uint32_t* __restrict__ counts, const uint32_t* __restrict__ row,
    const uint32_t* __restrict__ col, uint32_t pcap, const uint32_t* __restrict__ old_bd, uint32_t* __restrict__ bd,
    uint32_t* __restrict__ nearest, int hid, int gid, float dim, double* __restrict__ part_sum, float* __restrict__ part_max) {
    __shared__ double ssum[BLOCK];
    __shared__ double sabs[BLOCK];
    __shared__ float smax[BLOCK];
    double acc_sum = 0.0, acc_abs = 0.0;
    float acc_max = 0.0f;
    const uint32_t rows = counts[C_ROWS];
    ROW_LOOP_BEGIN(L)
    int4 pi = make_int4(0, 0, 0, 0);
    float kwi = 1.0f, ird = 1.0f;
    uint32_t self = 0;
    GaussConst g{1.0f, 1.0f};
    uint32_t rb = 0, re = 0, rbd = 0, red = 0;
    if (active) {
        pi = pos[i];
        kwi = kw[i];
        const float4 g4 = gc[i];
        g = GaussConst{g4.x, g4.y};
        ird = g4.z;
        self = __builtin_bit_cast(uint32_t, g4.w);
        rb = row[i];
        re = min(row[i + 1], pcap);
        if (i < rows) {  // the density pass covers the rows of the current list
            rbd = rb;
            red = re;
        }
    }
    uint32_t bmin = 0xFFFFFFFFu, best = 0xFFFFFFFFu, best_k = 0u, dens = 0;
    bool found = false;
    v3f unused;
    auto entry = [&](uint32_t k, const int4 pj, uint32_t obd, float imN) {
        const int32_t dx = isub(pj.x, pi.x), dy = isub(pj.y, pi.y), dz = isub(pj.z, pi.z);
        // find_split_and_merge_1/2 (as k_fsm12_rows)
        const v3f dv = vec_i(dx, dy, dz);
        const uint32_t dist = to_u32(sqrt_in_range(dot(dv, dv)));  // as in k_fsm12_rows
        const uint32_t cand = obd + dist;
        bmin = cand < bmin ? cand : bmin;
        if (!found || dist <= best) {
            best = dist;
            best_k = k;
            found = true;
        }
        // the density term (as k_density_rows)
        if (k >= rbd && k < red) {
            const v3f diff = vec_i(dx, dy, dz) / POS_RES;
            float w;
            kernel_eval_k<KH, 0>(hid, gid, dim, diff, kwi, g, true, false, &w, &unused);
            const uint32_t ib = __builtin_bit_cast(uint32_t, imN);
            const uint32_t ie = (ib >> 23) & 0xFFu;
            float wq;
            if (__all((ib & 0x7FFFFFu) == 0u && ie >= 1u && ie <= 253u))  // exact scaling, as in k_solve_rows
                wq = w * __builtin_bit_cast(float, (ib & 0x80000000u) | ((254u - ie) << 23));
            else
                wq = w / imN;
            dens += to_u32(wq * INCOMP_RES);
        }
    };
    uint32_t k = rb + gl;
    for (; k + 3u * L < re; k += 4u * L) {
        uint32_t jb[4], ob[4];
        int4 pb[4];
        float mb[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) jb[t] = col[k + (uint32_t)t * L];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            pb[t] = pos[jb[t]];
            ob[t] = old_bd[jb[t]];
            mb[t] = inv_mass[jb[t]];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) entry(k + (uint32_t)t * L, pb[t], ob[t], mb[t]);
    }
    for (; k < re; k += L) {
        const uint32_t j = col[k];
        entry(k, pos[j], old_bd[j], inv_mass[j]);
    }
    bmin = gmin<L>(bmin);
    dens = gsum<L>(dens);
#pragma unroll
    for (int m = L / 2; m > 0; m >>= 1) {
        const uint32_t ob = __shfl_xor(best, m, 64);
        const uint32_t ok = __shfl_xor(best_k, m, 64);
        const bool of = __shfl_xor((int)found, m, 64) != 0;
        if (of && (!found || ob < best || (ob == best && ok > best_k))) {
            best = ob;
            best_k = ok;
            found = true;
        }
    }
    if (active && gl == 0) {
        bd[i] = bmin;
        nearest[i] = found ? col[best_k] : i;
        dens += self;
        const float dev = (float)dens / INCOMP_RES * ird - 1.0f;
        const float e = fmax_g(0.0f, dev);
        acc_sum += (double)e;
        acc_abs += (double)__builtin_fabsf(dev);
        acc_max = fmax_g(acc_max, e);
    }
    ROW_LOOP_END
    ssum[threadIdx.x] = acc_sum;
    sabs[threadIdx.x] = acc_abs;
    smax[threadIdx.x] = acc_max;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            ssum[threadIdx.x] += ssum[threadIdx.x + s];
            sabs[threadIdx.x] += sabs[threadIdx.x + s];
            smax[threadIdx.x] = fmax_g(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part_sum[blockIdx.x] = ssum[0];
        part_sum[gridDim.x + blockIdx.x] = sabs[0];
        part_max[blockIdx.x] = smax[0];
    }
}

// kernel_width.comp:22-56 over row i. Pass FILL = false: spread the original
// kernel width to every neighbour (atomicMax, order-free) and count the pairs
// within the cutoff; FILL = true: write those pairs from row2[i] in row order.
template <int L, bool FILL>
__global__ __launch_bounds__(BLOCK) void k_spread_rows(const int4* __restrict__ pos,
                                                       const float* __restrict__ radius,
                                                       const float* __restrict__ trad,
                                                       const float* __restrict__ kw_old,
                                                       const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ row,
                                                       const uint32_t* __restrict__ col, uint32_t pcap,
                                                       int on_target, uint32_t* __restrict__ ukw,
                                                       uint32_t* __restrict__ cnt2,
                                                       const uint32_t* __restrict__ row2,
                                                       uint32_t* __restrict__ col2) {
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    ROW_LOOP_BEGIN(L)
    int4 pi = make_int4(0, 0, 0, 0);
    float okw = 0.0f, cutoff = 0.0f;
    uint32_t rb = 0, re = 0, out = 0;
    if (active) {
        pi = pos[i];
        const float tr = on_target ? trad[i] : 0.0f;
        okw = fmax_g(radius[i], tr) * KERNEL_SCALE;
        cutoff = fmax_g(okw, kw_old[i]);
        rb = row[i];
        re = min(row[i + 1], pcap);
        if (FILL) out = row2[i];
    }
    uint32_t c = 0;
    // the group walks the row in chunks of L entries so kept pairs keep row order
    const uint32_t len = re - rb;
    const uint32_t maxlen = gmin<L>(0xFFFFFFFFu - len);  // max over the group, as a min of complements
    const uint32_t chunks = ((0xFFFFFFFFu - maxlen) + L - 1) / L;
    for (uint32_t ch = 0; ch < chunks; ++ch) {
        const uint32_t k = rb + ch * L + gl;
        bool keep = false;
        uint32_t j = 0;
        if (k < re) {
            j = col[k];
            const int4 pj = pos[j];
            const v3f diff = vec_i(isub(pj.x, pi.x), isub(pj.y, pi.y), isub(pj.z, pi.z)) / POS_RES;
            const float dist = length(diff);
            if (!FILL) {
                const float dfk = dist - okw;
                const float infl = fmax_g(0.0f, 1.0f - fmax_g(0.0f, dfk / (okw * KW_PROPAGATION)));
                atomicMax(&ukw[j], to_u32(okw * infl * KW_RES));
            }
            keep = dist <= cutoff;
        }
        const uint64_t bal = __ballot(keep) & group_mask<L>();
        if (FILL && keep) col2[out + c + (uint32_t)__popcll(bal & below)] = j;
        c += (uint32_t)__popcll(bal);
    }
    if (!FILL && active && gl == 0) cnt2[i] = c;
    ROW_LOOP_END
    if (!FILL && blockIdx.x == 0 && threadIdx.x == 0) cnt2[counts[C_N]] = 0;
}

// ------------------------------------------------------------ launch wrappers
static uint32_t row_grid(uint32_t cap, uint32_t lanes = ROW_LANES) {
    const uint32_t groups = BLOCK / lanes;
    uint32_t g = div_up(cap ? cap : 1, groups);
    return g < 8192u ? g : 8192u;
}

// Rows longer than a row group's gradient cache (SOLVE_CACHE entries per lane
// in registers, SOLVE_SPILL in LDS) go to a pass with a whole wavefront per
// row. In the settled water drop 6 % of the rows (the coarse interior
// particles, 145 to 1 000 entries) hold 37 % of the pairs; a 16-lane group
// takes them through its uncached tail loop one dependent load chain at a
// time, with 3 idle row groups beside it in the wave.
// Rows up to the 32-lane cache (288 entries) take 32 lanes per row, two rows
// per wavefront, which halves the per-row work (lambda, boundariness: about
// 280 instructions per row group) against a wavefront per row; the batch of
// 64 reference scenes has many rows of 145 to 256 entries (28.8 ms per batch
// substep with a wavefront for every row above 144 entries, 27.4 without
// the split).
static uint32_t solve_split_len(int lanes) { return (uint32_t)((SOLVE_CACHE + SOLVE_SPILL) * lanes); }

// Lists the rows of more than split_len entries (after the pair-capacity
// clamp) for the passes over long rows: those up to split2 entries in list,
// the longer ones in list2. One atomic per list and wavefront of 64 rows, the
// rows of a wavefront in order. counts[C_LONG] and counts[C_LONG2] are zero
// before.
__global__ __launch_bounds__(BLOCK) void k_long_rows(uint32_t* __restrict__ counts, const uint32_t* __restrict__ row,
                                                     uint32_t pcap, uint32_t split_len, uint32_t split2,
                                                     uint32_t* __restrict__ list, uint32_t* __restrict__ list2) {
    const uint32_t n = counts[C_N];
    const uint32_t lane = threadIdx.x & 63u;
    for (uint32_t base = (blockIdx.x * BLOCK + threadIdx.x) & ~63u; base < n; base += gridDim.x * BLOCK) {
        const uint32_t i = base + lane;
        uint32_t len = 0;
        if (i < n) {
            const uint32_t rb = row[i], re = min(row[i + 1], pcap);
            len = re > rb ? re - rb : 0u;
        }
        const bool in1 = len > split_len && len <= split2, in2 = len > split2;
        const uint64_t b1 = __ballot(in1), b2 = __ballot(in2);
        const uint64_t below = (1ull << lane) - 1ull;
        if (b1) {  // the same for every lane of the wavefront
            uint32_t at = 0;
            if (lane == 0) at = atomicAdd(&counts[C_LONG], (uint32_t)__popcll(b1));
            at = (uint32_t)__shfl((int)at, 0, 64);
            if (in1) list[at + (uint32_t)__popcll(b1 & below)] = i;
        }
        if (b2) {
            uint32_t at = 0;
            if (lane == 0) at = atomicAdd(&counts[C_LONG2], (uint32_t)__popcll(b2));
            at = (uint32_t)__shfl((int)at, 0, 64);
            if (in2) list2[at + (uint32_t)__popcll(b2 & below)] = i;
        }
    }
}

void launch_long_rows(uint32_t* counts, uint32_t cap, const uint32_t* row, uint32_t pcap, int lanes, uint32_t* list,
                      uint32_t* list2, hipStream_t s) {
    static_assert(C_LONG2 == C_LONG + 1, "adjacent counters");
    fill_u32(counts + C_LONG, 0u, 2, s);
    const uint32_t s1 = solve_split_len(lanes), s2 = solve_split_len(32);
    hipLaunchKernelGGL(k_long_rows, dim3(stride_grid(cap)), dim3(BLOCK), 0, s, counts, row, pcap, s1,
                       s2 > s1 ? s2 : s1, list, list2);
}

void launch_solve_rows(const int4* pos, const float* inv_mass, const float* radius, const float* kw,
                       const float4* gc, const uint32_t* counts, uint32_t cap, const uint32_t* row,
                       const uint32_t* col, uint32_t pcap, const SolverArgs& a_in, float* boundariness,
                       float* lambda, int4* acc, hipStream_t s, int lanes, bool long_rows,
                       const uint32_t* long_list, const uint32_t* long_list2) {
    int* ac = reinterpret_cast<int*>(acc);
    SolverArgs a = a_in;
    const bool split = long_list && long_list2 && a.hid == 1 && a.gid == 1 &&
                       (lanes == 8 || lanes == 16 || lanes == 32);
    if (split) a.split_len = solve_split_len(lanes);
#define SOLVE(L, KH, KG, C, F)                                                                                       \
    hipLaunchKernelGGL((k_solve_rows<L, KH, KG, C, F>), dim3(row_grid(cap, L)), dim3(BLOCK), 0, s, pos, inv_mass,    \
                       radius, kw, gc, counts, row, col, pcap, a, boundariness, lambda, ac)
    if (split) {  // the long rows (listed by launch_long_rows): 32 lanes per row, then a wavefront per row
        SolverArgs al = a;
        al.split_len = 0;
        al.list = long_list2;
        al.list_slot = C_LONG2;
        hipLaunchKernelGGL((k_solve_rows<64, 1, 1, SOLVE_CACHE, 1>), dim3(row_grid(cap, 64)), dim3(BLOCK), 0, s,
                           pos, inv_mass, radius, kw, gc, counts, row, col, pcap, al, boundariness, lambda, ac);
        if (lanes != 32) {  // with 32-lane groups the first list is empty
            al.list = long_list;
            al.list_slot = C_LONG;
            hipLaunchKernelGGL((k_solve_rows<32, 1, 1, SOLVE_CACHE, 1>), dim3(row_grid(cap, 32)), dim3(BLOCK), 0, s,
                               pos, inv_mass, radius, kw, gc, counts, row, col, pcap, al, boundariness, lambda, ac);
        }
    }
    if (a.hid == 1 && a.gid == 1 && lanes == 8)  // Gaussian height and gradient, 8 lanes per row
        SOLVE(8, 1, 1, SOLVE_CACHE, AGG_STEPS);
    else if (a.hid == 1 && a.gid == 1 && lanes == 32)  // 32 lanes: rows of up to 192 entries stay in registers
        SOLVE(32, 1, 1, SOLVE_CACHE, AGG_STEPS);
    else if (a.hid == 1 && a.gid == 1 && long_rows && !split)  // the reference's defaults (Gaussian), long rows
        SOLVE(ROW_LANES, 1, 1, SOLVE_CACHE, AGG_STEPS / 2);
    else if (a.hid == 1 && a.gid == 1)
        SOLVE(ROW_LANES, 1, 1, SOLVE_CACHE, AGG_STEPS);
    else
        SOLVE(ROW_LANES, -1, -1, SOLVE_CACHE, AGG_STEPS);
#undef SOLVE
}
void launch_features_rows(const int4* pos, const float4* vel, const float* inv_mass, const float* radius,
                          const float* kw, const float4* gc, const float* boundariness, const float* lam_prev,
                          const uint32_t* counts, uint32_t cap, const uint32_t* row, const uint32_t* col,
                          uint32_t pcap, const SolverArgs& a, float dt, float* feat, hipStream_t s) {
    if (a.hid == 1 && a.gid == 1)
        hipLaunchKernelGGL((k_features_rows<ROW_LANES, 1, 1>), dim3(row_grid(cap)), dim3(BLOCK), 0, s, pos, vel,
                           inv_mass, radius, kw, gc, boundariness, lam_prev, counts, row, col, pcap, a, dt, feat);
    else
        hipLaunchKernelGGL((k_features_rows<ROW_LANES, -1, -1>), dim3(row_grid(cap)), dim3(BLOCK), 0, s, pos, vel,
                           inv_mass, radius, kw, gc, boundariness, lam_prev, counts, row, col, pcap, a, dt, feat);
}

void launch_solve_apply(int4* pos, int4* acc, const float* radius, const uint32_t* counts, uint32_t cap,
                        const BoxSet& boxes, hipStream_t s, const uint32_t* action, const int4* saved,
                        const SegmentTable* seg) {
    const SegmentTable none{nullptr, nullptr, nullptr, nullptr, 0u};
    hipLaunchKernelGGL(k_solve_apply, dim3(stride_grid(cap)), dim3(BLOCK), 0, s, pos, acc, radius, counts, boxes,
                       action, saved, seg ? *seg : none);
}

// Start of a solve with a tolerance or a warm start: nothing decided yet, the
// error sums of this substep's passes cleared (slot 0: the warm-start pass,
// slot 1 + k: plain iteration k).
__global__ void k_solve_begin(uint32_t* __restrict__ counts, double* __restrict__ tol_err, uint32_t slots) {
    for (uint32_t k = threadIdx.x; k < slots; k += blockDim.x) tol_err[k] = 0.0;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        counts[C_SOLVE_DONE] = 0u;
        counts[C_SOLVE_ACTION] = SOLVE_APPLY;
        counts[C_SOLVE_ITERS] = 0u;
        counts[C_WARM_FALLBACK] = 0u;
    }
}
// After the row pass of plain iteration k: its mean density error e_k decides
// the apply pass. With a warm start, iteration 0 first judges the warm shift:
// when the error after it exceeds the error before it (e_0 of the warm pass),
// the positions from before the warm shift are put back. Then, with a
// tolerance, e_k <= tol ends the solve (this and later iterations apply
// nothing).
__global__ void k_solve_decide(uint32_t* __restrict__ counts, const double* __restrict__ tol_err, uint32_t k,
                               float tol, int warm) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (counts[C_SOLVE_DONE]) {
        counts[C_SOLVE_ACTION] = SOLVE_SKIP;
        return;
    }
    const uint32_t n = counts[C_N];
    const double e = n ? tol_err[1 + k] / (double)n : 0.0;
    if (warm && k == 0) {
        const double e0 = n ? tol_err[0] / (double)n : 0.0;
        if (e > e0) {
            counts[C_SOLVE_ACTION] = SOLVE_RESTORE;
            counts[C_WARM_FALLBACK] = 1u;
            return;
        }
    }
    if (tol > 0.0f && e <= (double)tol) {
        counts[C_SOLVE_DONE] = 1u;
        counts[C_SOLVE_ACTION] = SOLVE_SKIP;
        return;
    }
    counts[C_SOLVE_ACTION] = SOLVE_APPLY;
    counts[C_SOLVE_ITERS] += 1u;
}
// Per substep: stats[0] += iterations applied, stats[1] += 1, stats[2] += fallbacks.
__global__ void k_solve_stats(const uint32_t* __restrict__ counts, unsigned long long* __restrict__ stats,
                              uint32_t fixed_iterations) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    stats[0] += fixed_iterations ? fixed_iterations : counts[C_SOLVE_ITERS];
    stats[1] += 1ull;
    stats[2] += counts[C_WARM_FALLBACK];
}
void launch_solve_begin(uint32_t* counts, double* tol_err, uint32_t slots, hipStream_t s) {
    hipLaunchKernelGGL(k_solve_begin, dim3(1), dim3(64), 0, s, counts, tol_err, slots);
}
void launch_solve_decide(uint32_t* counts, const double* tol_err, uint32_t k, float tol, bool warm, hipStream_t s) {
    hipLaunchKernelGGL(k_solve_decide, dim3(1), dim3(64), 0, s, counts, tol_err, k, tol, warm ? 1 : 0);
}
void launch_solve_stats(const uint32_t* counts, unsigned long long* stats, uint32_t fixed_iterations, hipStream_t s) {
    hipLaunchKernelGGL(k_solve_stats, dim3(1), dim3(64), 0, s, counts, stats, fixed_iterations);
}
uint32_t density_rows_blocks(uint32_t cap) { return row_grid(cap) < 2048u ? row_grid(cap) : 2048u; }
void launch_density_rows(const int4* pos, const float* inv_mass, const float* radius, const float* kw,
                         const float4* gc, const uint32_t* counts, uint32_t cap, const uint32_t* row,
                         const uint32_t* col, uint32_t pcap, int hid, int gid, float dim, double* part_sum,
                         float* part_max, hipStream_t s) {
    if (hid == 1)
        hipLaunchKernelGGL((k_density_rows<ROW_LANES, 1>), dim3(density_rows_blocks(cap)), dim3(BLOCK), 0, s, pos,
                           inv_mass, radius, kw, gc, counts, row, col, pcap, hid, gid, dim, part_sum, part_max);
    else
        hipLaunchKernelGGL((k_density_rows<ROW_LANES, -1>), dim3(density_rows_blocks(cap)), dim3(BLOCK), 0, s, pos,
                           inv_mass, radius, kw, gc, counts, row, col, pcap, hid, gid, dim, part_sum, part_max);
}
void launch_fsm12_density_rows(const int4* pos, const float* inv_mass, const float* kw, const float4* gc,
                               const uint32_t* counts, uint32_t cap, const uint32_t* row, const uint32_t* col,
                               uint32_t pcap, const uint32_t* old_bd, uint32_t* bd, uint32_t* nearest, int hid,
                               int gid, float dim, double* part_sum, float* part_max, hipStream_t s) {
    if (hid == 1)
        hipLaunchKernelGGL((k_fsm12_density_rows<ROW_LANES, 1>), dim3(density_rows_blocks(cap)), dim3(BLOCK), 0, s,
                           pos, inv_mass, kw, gc, counts, row, col, pcap, old_bd, bd, nearest, hid, gid, dim, part_sum,
                           part_max);
    else
        hipLaunchKernelGGL((k_fsm12_density_rows<ROW_LANES, -1>), dim3(density_rows_blocks(cap)), dim3(BLOCK), 0, s,
                           pos, inv_mass, kw, gc, counts, row, col, pcap, old_bd, bd, nearest, hid, gid, dim, part_sum,
                           part_max);
}
void launch_fsm12_rows(const int4* pos, const uint32_t* counts, uint32_t cap, const uint32_t* row,
                       const uint32_t* col, uint32_t pcap, const uint32_t* old_bd, uint32_t* bd, uint32_t* nearest,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_fsm12_rows<ROW_LANES>, dim3(row_grid(cap)), dim3(BLOCK), 0, s, pos, counts, row, col, pcap,
                       old_bd, bd, nearest);
}
void launch_spread_rows(bool fill, const int4* pos, const float* radius, const float* trad, const float* kw_old,
                        const uint32_t* counts, uint32_t cap, const uint32_t* row, const uint32_t* col, uint32_t pcap,
                        int on_target, uint32_t* ukw, uint32_t* cnt2, const uint32_t* row2, uint32_t* col2,
                        hipStream_t s) {
    if (fill)
        hipLaunchKernelGGL((k_spread_rows<ROW_LANES, true>), dim3(row_grid(cap)), dim3(BLOCK), 0, s, pos, radius, trad,
                           kw_old, counts, row, col, pcap, on_target, ukw, cnt2, row2, col2);
    else
        hipLaunchKernelGGL((k_spread_rows<ROW_LANES, false>), dim3(row_grid(cap)), dim3(BLOCK), 0, s, pos, radius,
                           trad, kw_old, counts, row, col, pcap, on_target, ukw, cnt2, row2, col2);
}

}  // namespace apbf
