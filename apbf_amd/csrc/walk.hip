// Neighbour search over the cell tables (neighborhood_green.comp:57-86), one
// group of L lanes of a wavefront per particle. The group visits the cells of
// the particle's search box in the reference order (z, then y, then x, then
// the particles of a cell in sorted order) and tests L candidates of a cell
// per step; a wave ballot turns the accepted candidates of a step into
// consecutive output slots, so every row keeps the sequential order and the
// writes of a step are contiguous.
//
// Modes: COUNT writes each row's length, FILL writes the rows at the offsets
// of the scanned counts, clamped to the pair capacity the way the sequential
// schedule fills a bounded list: entries whose slot in the full list is at or
// beyond the capacity are dropped (neighbor_add.glsl:22-23). PAD does both in
// one walk: it writes every row into a fixed-stride slot of its own (stride =
// pair capacity / current particles) and counts it; when no row outgrew its
// slot and the list fits the capacity (the normal case), the rows are then
// packed into the list, and otherwise the exact COUNT-less FILL walk runs
// (the counts of PAD are the counts of COUNT). The choice is made on the device, so the substep
// needs no host synchronisation.
//
// With SPREAD the walk also carries kernel_width.comp:22-56: every entry of
// the (clamped) list spreads the particle's original kernel width to the
// neighbour (atomicMax, order-free; skipped when a plain load already shows a
// value at least as large, which is exact because the stored values only grow
// during the walk), and only the entries within the
// particle's cutoff are written, which is the list the spreading pass of the
// reference leaves behind (kernel_width.comp:54-60). The full list is then
// never materialised: COUNT counts both lists, FILL writes the short one.
#include "scene.h"

namespace apbf {

constexpr int WALK_LANES = 8;

struct WalkArgs {
    float range_scale;
    GridDesc g;
    int on_target;  // spreading: kernel width from max(radius, target radius)
    uint32_t pcap;
    uint32_t stride;        // PAD, and FILL after PAD: nonzero (the slot stride is slot_stride())
    const uint32_t* gate;   // FILL after PAD: the count block (C_FALLBACK, C_ROWFIX); null otherwise
    unsigned long long* cand;  // COUNT: when set, += the candidates the walk tests (a diagnostic)
    SegmentTable seg;       // a packed batch (SEG walks): each particle's grid and cell block
};

enum WalkMode { WALK_COUNT = 0, WALK_FILL = 1, WALK_PAD = 2 };

// Entries per particle slot in the PAD layout: the pair capacity shared by the
// current particles (not by the particle capacity), so slots are as wide as
// the list allows. Every kernel of one pair build reads the same particle
// count, so they agree on it.
__device__ __forceinline__ uint32_t slot_stride(uint32_t pcap, uint32_t n) { return pcap / (n ? n : 1u); }


// Eight waves per SIMD (64 VGPRs, a 28-byte spill) hide the latency of the
// cell-table and position loads better than seven without the cap
// (71 VGPRs): 8.04 against 8.24 ms per water-drop substep.
template <int L, int MODE, bool SPREAD, bool SEG>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8))) void k_walk_rows(
    const int4* __restrict__ pos, const float* __restrict__ kw, const float* __restrict__ radius,
    const float* __restrict__ trad, const uint32_t* __restrict__ cs, const uint32_t* __restrict__ ce,
    uint32_t* __restrict__ counts, WalkArgs wa, uint32_t* __restrict__ cnt_raw, uint32_t* __restrict__ cnt_f,
    const uint32_t* __restrict__ row_raw, const uint32_t* __restrict__ row_f, uint32_t* __restrict__ col,
    uint32_t* __restrict__ ukw) {
    constexpr bool FILL = MODE == WALK_FILL;
    constexpr bool PAD = MODE == WALK_PAD;
    // After a PAD walk the exact fill runs for every row when the list exceeds
    // the capacity, only for the rows that outgrew their slot otherwise, and
    // not at all in the normal case.
    bool only_long_rows = false;
    if (FILL && wa.gate) {
        if (wa.gate[C_FALLBACK] == 0u) {
            if (wa.gate[C_ROWFIX] == 0u) return;
            only_long_rows = true;
        }
    }
    const int lane = threadIdx.x & 63;
    // A ballot's bits of this lane's group, shifted down: their count is the
    // group's total, and those below the lane's bit its rank.
    const uint32_t gshift = (uint32_t)lane & ~(uint32_t)(L - 1);
    constexpr uint32_t GBITS = (L >= 32) ? 0xFFFFFFFFu : ((1u << L) - 1u);
    const uint32_t gbelow = (1u << (threadIdx.x & (L - 1))) - 1u;
    auto group_bits = [&](uint64_t bal) { return (uint32_t)(bal >> gshift) & GBITS; };
    const uint32_t n = counts[C_N];
    const uint32_t stride = wa.stride ? slot_stride(wa.pcap, n) : 0u;
    const int gl = threadIdx.x & (L - 1);
    const uint32_t groups = BLOCK / L;
    for (uint32_t base = blockIdx.x * groups; base < n; base += gridDim.x * groups) {
        const uint32_t i = base + threadIdx.x / L;
        if (i >= n) continue;  // uniform within the group
        if (only_long_rows && row_f[i + 1] - row_f[i] <= stride) continue;
        // a packed batch: the particle's own scene grid and block of cells
        const uint32_t sc = SEG ? wa.seg.sid[i] : 0u;
        const GridDesc& g = SEG ? wa.seg.grids[sc] : wa.g;
        const uint32_t cbase = SEG ? (sc << wa.seg.shift) : 0u;
        const int4 pi = pos[i];
        const v3f p = pos_world(pi);
        const float kwi = kw[i];
        const float range = kwi * wa.range_scale;
        uint32_t lo[3], hi[3];
        lo[0] = grid_cell(p.x - range, g.mn[0], g.mx[0], g.res);
        lo[1] = grid_cell(p.y - range, g.mn[1], g.mx[1], g.res);
        lo[2] = grid_cell(p.z - range, g.mn[2], g.mx[2], g.res);
        hi[0] = grid_cell(p.x + range, g.mn[0], g.mx[0], g.res);
        hi[1] = grid_cell(p.y + range, g.mn[1], g.mx[1], g.res);
        hi[2] = grid_cell(p.z + range, g.mn[2], g.mx[2], g.res);
        // neighborhood_green.comp:36: in 2D the walk stays in z cell 0
        if (g.dims == 2u) lo[2] = hi[2] = 0u;
        // Bound the walk to 4 * 2^res cells per axis: wider walks only revisit
        // aliased cells, and an unbounded span (infinite range) would not end.
        const uint32_t span = 4u << g.res;
        for (int d = 0; d < 3; ++d)
            if (hi[d] - lo[d] >= span && hi[d] >= lo[d]) hi[d] = lo[d] + span - 1u;
        // Distance tests compare squared distances with exact bounds (the
        // reference compares rounded square roots; sqrt_le_bound makes the two
        // agree for every input). A NaN range keeps every candidate.
        //
        // Both tests run on unscaled coordinates. The world distance is
        // p - pos_world(pj) = (float(xi) - float(xj)) * 2^-18 component-wise,
        // and the differences of integer-valued floats are 0 or at least 1 in
        // magnitude, so every rounding of the difference and of its squared
        // length commutes with the scaling by 2^-18 (no result is subnormal):
        // the squared world distance is exactly dot(u, u) * 2^-36 with
        // u = float(xi) - float(xj), and likewise for the fixed-point
        // difference. Comparing dot(u, u) with the bound times 2^36 (an exact
        // scaling; an overflow to infinity keeps the comparison's outcome, as
        // dot(u, u) * 2^-36 < 2^30) is then the same test.
        const float t_range = ((range != range) ? __builtin_inff() : sqrt_le_bound(range)) * 0x1p36f;
        const v3f pr = vec_i(pi.x, pi.y, pi.z);
        float okw = 0.0f, t_cut = 0.0f, t_okw = 0.0f;
        uint32_t umax = 0;
        if (SPREAD) {
            const float tr = wa.on_target ? trad[i] : 0.0f;
            okw = fmax_g(radius[i], tr) * KERNEL_SCALE;
            t_cut = sqrt_le_bound(fmax_g(okw, kwi)) * 0x1p36f;
            t_okw = sqrt_le_bound(okw) * 0x1p36f;
            umax = to_u32(okw * KW_RES);
        }
        // kernel_width.comp:41-53 for an entry at squared distance d2: the
        // spread width is to_u32(okw * influence * 2^18) with the influence in
        // [0, 1], so it never exceeds umax (rounding and to_u32 are monotone).
        // The neighbour's stored value starts at its own width and only grows,
        // so where it already reaches umax (the common case of neighbours no
        // smaller than the particle) the entry changes nothing and the
        // influence is not evaluated.
        // `cur` is the neighbour's stored value, loaded with its position.
        auto spread_to = [&](uint32_t j, float d2r, uint32_t cur) {  // d2r: squared distance * 2^36
            if (umax <= cur) return;
            uint32_t uval;
            if (d2r <= t_okw) {  // distance <= okw: influence exactly 1
                uval = umax;
            } else {
                const float dfk = __builtin_sqrtf(d2r * 0x1p-36f) - okw;
                const float infl = fmax_g(0.0f, 1.0f - fmax_g(0.0f, dfk / (okw * KW_PROPAGATION)));
                uval = to_u32(okw * infl * KW_RES);
            }
            if (uval > cur) atomicMax(&ukw[j], uval);
        };
        // Cell culling: a cell whose box lies farther than the range holds no
        // accepted candidate. The cell boxes are widened by margins far above
        // the rounding of the hash (grid_cell) and of the distance test, and
        // the cells at the grid border (which also hold clamped or aliased
        // particles) are never culled, so culling drops no accepted pair. A
        // search box reaching past the grid end can visit an aliased bucket
        // twice; such walks are not culled.
        const uint32_t last_cell = (1u << g.res) - 1u;
        const bool cull = t_range < __builtin_inff() && range >= 0.0f && hi[0] <= last_cell && hi[1] <= last_cell &&
                          hi[2] <= last_cell;
        const float thr = cull ? range * range * 1.001f + 1e-30f : __builtin_inff();  // NaN range: no cull
        auto axis_d2 = [&](int d, uint32_t c, float pc) -> float {
            if (!cull || c == 0u || c >= last_cell) return 0.0f;
            const float w = (g.mx[d] - g.mn[d]) / (float)(1u << g.res);
            const float m = 1e-4f * w + 1e-6f * (__builtin_fabsf(g.mn[d]) + __builtin_fabsf(g.mx[d]));
            const float a = g.mn[d] + (float)c * w - m;
            const float b = g.mn[d] + (float)(c + 1u) * w + m;
            const float t = pc < a ? a - pc : (pc > b ? pc - b : 0.0f);
            return t * t;
        };
        uint32_t craw = 0, cf = 0, braw = 0, bf = 0;
        uint64_t tested = 0;  // COUNT: candidates tested for this particle
        bool more = true;
        if (FILL) {
            braw = row_raw[i];
            bf = SPREAD ? row_f[i] : braw;
            more = braw < wa.pcap;
        }
        const size_t slot = PAD ? (size_t)i * stride : 0;
        // The cells of one x-line of the search box are taken CW at a time:
        // lane c < CW of the group looks up cell lo.x + c0 + c, a scan over
        // the group gives each cell's offset in the line's candidates taken
        // in cell order, and the offsets and index deltas of the CW cells are
        // broadcast to every lane of the group. Every step then tests the
        // next L of those candidates, so steps run full across cell ends; a
        // lane picks its candidate's cell by comparing with the offsets held
        // in its registers. A candidate's position and (spreading) stored
        // width are loaded together.
        constexpr int CW = L < 4 ? L : 4;
        const uint32_t nline = hi[0] >= lo[0] ? hi[0] - lo[0] + 1u : 0u;
        for (uint32_t z = lo[2]; more && z <= hi[2]; ++z) {
            const float dz2 = axis_d2(2, z, p.z);
            for (uint32_t y = lo[1]; more && y <= hi[1] && dz2 <= thr; ++y) {
                const float dyz2 = dz2 + axis_d2(1, y, p.y);
                const uint32_t hyz = zbits(y, 1u, g.res, g.dims) | (g.dims == 2u ? 0u : zbits(z, 2u, g.res, g.dims));
                for (uint32_t c0 = 0; more && c0 < nline && dyz2 <= thr; c0 += CW) {
                    const uint32_t nc = min(nline - c0, (uint32_t)CW);
                    uint32_t cstart = 0, clen = 0;
                    if ((uint32_t)gl < nc) {
                        const uint32_t x = lo[0] + c0 + (uint32_t)gl;
                        if (!(dyz2 + axis_d2(0, x, p.x) > thr)) {
                            const uint32_t h = cbase + (zbits(x, 0u, g.res, g.dims) | hyz);
                            cstart = cs[h];
                            const uint32_t cend = ce[h];
                            clen = cend > cstart ? cend - cstart : 0u;
                        }
                    }
                    uint32_t incl = clen;
#pragma unroll
                    for (int m = 1; m < CW; m <<= 1) {
                        const uint32_t v = (uint32_t)__shfl_up((int)incl, m, L);
                        if (gl >= m) incl += v;
                    }
                    // lanes c < CW past the last cell hold the total as offset
                    const uint32_t off = incl - clen;
                    const uint32_t delta = cstart - off;
                    const uint32_t total = (uint32_t)__shfl((int)incl, CW - 1, L);
                    if (MODE == WALK_COUNT) tested += total;
                    uint32_t o[CW], dl[CW];
#pragma unroll
                    for (int c = 0; c < CW; ++c) {
                        o[c] = (uint32_t)__shfl((int)off, c, L);
                        dl[c] = (uint32_t)__shfl((int)delta, c, L);
                    }
                    for (uint32_t t0 = 0; more && t0 < total; t0 += L) {
                        const uint32_t t = t0 + (uint32_t)gl;
                        uint32_t d = dl[0];
#pragma unroll
                        for (int c = 1; c < CW; ++c)
                            if (t >= o[c]) d = dl[c];
                        const uint32_t j = t + d;
                        bool green = false, keep = false;
                        float d2 = 0.0f;
                        uint32_t uj = 0u;
                        if (t < total && j != i) {
                            const int4 pj = pos[j];
                            if (SPREAD && MODE != WALK_COUNT) uj = ukw[j];
                            const v3f u = pr - vec_i(pj.x, pj.y, pj.z);
                            green = !(dot(u, u) > t_range);
                            if (SPREAD && green) {
                                const v3f v = vec_i(isub(pj.x, pi.x), isub(pj.y, pi.y), isub(pj.z, pi.z));
                                d2 = dot(v, v);
                                keep = d2 <= t_cut;
                            }
                        }
                        const uint32_t rbal = group_bits(__ballot(green));
                        if (PAD) {
                            if (SPREAD && green) spread_to(j, d2, uj);
                            const bool kept = SPREAD ? (green && keep) : green;
                            const uint32_t kbal = SPREAD ? group_bits(__ballot(kept)) : rbal;
                            const uint32_t k = cf + (uint32_t)__popc(kbal & gbelow);
                            if (kept && k < stride) col[slot + k] = j;
                            cf += (uint32_t)__popc(kbal);
                            craw += (uint32_t)__popc(rbal);
                        } else if (!FILL) {
                            craw += (uint32_t)__popc(rbal);
                            if (SPREAD) cf += (uint32_t)__popc(group_bits(__ballot(keep)));
                        } else {
                            const uint32_t ridx = braw + craw + (uint32_t)__popc(rbal & gbelow);
                            const bool incl_r = green && ridx < wa.pcap;
                            if (SPREAD) {
                                if (incl_r) spread_to(j, d2, uj);
                                const bool kept = incl_r && keep;
                                const uint32_t fbal = group_bits(__ballot(kept));
                                if (kept) col[bf + cf + (uint32_t)__popc(fbal & gbelow)] = j;
                                cf += (uint32_t)__popc(fbal);
                            } else if (incl_r) {
                                col[ridx] = j;
                            }
                            craw += (uint32_t)__popc(rbal);
                            more = braw + craw < wa.pcap;  // later entries fall beyond the capacity
                        }
                    }
                }
                if (y == 0xFFFFFFFFu) break;
            }
            if (z == 0xFFFFFFFFu) break;
        }
        if (gl == 0) {
            if (PAD) {
                cnt_raw[i] = craw;
                if (SPREAD) cnt_f[i] = cf;
                if (cf > stride) atomicMax(&counts[C_MAXROW], cf);
            } else if (!FILL) {
                cnt_raw[i] = craw;
                if (SPREAD) cnt_f[i] = cf;
                if (wa.cand) atomicAdd(wa.cand, (unsigned long long)tested);
            } else if (SPREAD && braw <= wa.pcap && row_raw[i + 1] > wa.pcap) {
                // the row the capacity cuts: its kept short-list length
                counts[C_CLAMP_ROW] = i;
                counts[C_CLAMP_KEPT] = cf;
            }
        }
    }
    if (!FILL && blockIdx.x == 0 && threadIdx.x == 0) {
        cnt_raw[n] = 0;
        if (SPREAD) cnt_f[n] = 0;
    }
}

// Inclusive scan over the 64 lanes of a wavefront (every lane active) of an
// associative, commutative op with identity 0 on unsigned values (sum, max):
// DPP row shifts inside each 16-lane row, then the row totals by readlane.
template <typename Op>
__device__ __forceinline__ uint32_t wave_scan(uint32_t v, Op op) {
    constexpr int ROW_SHR1 = 0x111, ROW_SHR2 = 0x112, ROW_SHR4 = 0x114, ROW_SHR8 = 0x118;
    auto shr = [](uint32_t x, auto ctrl) {  // lanes whose source leaves the row read 0
        return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, decltype(ctrl)::value, 0xF, 0xF, false);
    };
    v = op(v, shr(v, std::integral_constant<int, ROW_SHR1>{}));
    v = op(v, shr(v, std::integral_constant<int, ROW_SHR2>{}));
    v = op(v, shr(v, std::integral_constant<int, ROW_SHR4>{}));
    v = op(v, shr(v, std::integral_constant<int, ROW_SHR8>{}));
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)v, 15);
    const uint32_t r1 = op(r0, (uint32_t)__builtin_amdgcn_readlane((int)v, 31));
    const uint32_t r2 = op(r1, (uint32_t)__builtin_amdgcn_readlane((int)v, 47));
    const uint32_t row = (threadIdx.x & 63u) >> 4;
    const uint32_t carry = row == 0u ? 0u : (row == 1u ? r0 : (row == 2u ? r1 : r2));
    return op(v, carry);
}

// q = c / d and r = c % d for a divisor shared by the wave: m = floor(2^32 / d)
// (0xFFFFFFFF for d = 1) underestimates the quotient by at most one.
struct Div32 {
    uint32_t d, m;
    __device__ __forceinline__ static uint32_t multiplier(uint32_t dv) {
        return dv <= 1u ? 0xFFFFFFFFu : (uint32_t)(0x100000000ull / dv);
    }
    __device__ __forceinline__ Div32(uint32_t dv, uint32_t mv) : d(dv), m(mv) {}
    __device__ __forceinline__ uint32_t div(uint32_t c, uint32_t& r) const {
        uint32_t q = __umulhi(c, m);
        r = c - q * d;
        if (r >= d) {
            ++q;
            r -= d;
        }
        return q;
    }
};

// Wave-uniform copy of lane k's value (v_readlane into a scalar register).
template <typename T>
__device__ __forceinline__ T lane_value(T v, int k) {
    static_assert(sizeof(T) == 4, "32-bit values");
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), k));
}

// The same walk with a whole wavefront per particle and the cells of the
// search box flattened into one stream. Waves take batches of WAVE_BATCH
// consecutive particles (grid-stride; small batches keep the uneven cost of
// particles spread over the waves: 2.99 ms per water-drop walk at 4, 3.10 at
// 8, 3.60 at 16, 3.15 at 2; a queue counter shared by all waves was slower
// still, its atomics serialising). The per-particle set-up (box, exact
// distance bounds, culling constants) runs for a batch at once, one particle
// per lane; the wave then takes the particles one by one with their values
// read from those lanes. 64 cells (in the reference order: x fastest,
// then y, then z) are looked up at a time, one per lane, culled and scanned
// into stream offsets; the wave then tests 64 consecutive candidates of that
// stream per step across cell ends, so the lanes stay busy whatever the
// cells hold. Cells whose particles lie back to back in memory form one run;
// a lane takes its candidate's run from the run holding the step's first
// candidate and the (few) runs starting inside the step, read from their
// lanes. Outputs, order and clamps are those of k_walk_rows.
constexpr uint32_t WAVE_BATCH = 4;
// Candidates per lane and step (the step covers 64 * WAVE_CANDS consecutive
// candidates): the run lookup is shared and the loads of all of them are in
// flight together (water drop: 2.82 ms per walk at 2 and 3, 2.99 at 1, 3.08
// at 4).
constexpr uint32_t WAVE_CANDS = 2;
template <int MODE, bool SPREAD, bool SEG>
__global__ __launch_bounds__(BLOCK) void k_walk_wave(
    const int4* __restrict__ pos, const float* __restrict__ kw, const float* __restrict__ radius,
    const float* __restrict__ trad, const uint32_t* __restrict__ cs, const uint32_t* __restrict__ ce,
    uint32_t* __restrict__ counts, WalkArgs wa, uint32_t* __restrict__ cnt_raw, uint32_t* __restrict__ cnt_f,
    const uint32_t* __restrict__ row_raw, const uint32_t* __restrict__ row_f, uint32_t* __restrict__ col,
    uint32_t* __restrict__ ukw) {
    constexpr bool FILL = MODE == WALK_FILL;
    constexpr bool PAD = MODE == WALK_PAD;
    bool only_long_rows = false;
    if (FILL && wa.gate) {
        if (wa.gate[C_FALLBACK] == 0u) {
            if (wa.gate[C_ROWFIX] == 0u) return;
            only_long_rows = true;
        }
    }
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    auto vadd = [](uint32_t a, uint32_t b) { return a + b; };
    const uint32_t n = counts[C_N];
    const uint32_t stride = wa.stride ? slot_stride(wa.pcap, n) : 0u;
    const uint32_t waves = BLOCK / 64;
    for (uint32_t base = (blockIdx.x * waves + threadIdx.x / 64) * WAVE_BATCH; base < n;
         base += gridDim.x * waves * WAVE_BATCH) {
        // ---- set-up of particle base + lane, lanes < WAVE_BATCH
        const uint32_t ip = base + lane;
        bool act = lane < WAVE_BATCH && ip < n;
        if (act && only_long_rows && row_f[ip + 1] - row_f[ip] <= stride) act = false;
        uint32_t s_sc = 0, s_lo[3] = {0, 0, 0}, s_n[3] = {0, 0, 0}, s_umax = 0, s_cull = 0, s_braw = 0, s_bf = 0;
        uint32_t s_mx = 0, s_my = 0;  // Div32 multipliers of the box's x and y sizes
        int32_t s_pi[3] = {0, 0, 0};
        float s_trange = 0.0f, s_okw = 0.0f, s_tcut = 0.0f, s_tokw = 0.0f, s_thr = 0.0f;
        if (act) {
            s_sc = SEG ? wa.seg.sid[ip] : 0u;
            const GridDesc& g = SEG ? wa.seg.grids[s_sc] : wa.g;
            const int4 pi = pos[ip];
            const v3f p = pos_world(pi);
            const float kwi = kw[ip];
            const float range = kwi * wa.range_scale;
            uint32_t lo[3], hi[3];
            lo[0] = grid_cell(p.x - range, g.mn[0], g.mx[0], g.res);
            lo[1] = grid_cell(p.y - range, g.mn[1], g.mx[1], g.res);
            lo[2] = grid_cell(p.z - range, g.mn[2], g.mx[2], g.res);
            hi[0] = grid_cell(p.x + range, g.mn[0], g.mx[0], g.res);
            hi[1] = grid_cell(p.y + range, g.mn[1], g.mx[1], g.res);
            hi[2] = grid_cell(p.z + range, g.mn[2], g.mx[2], g.res);
            if (g.dims == 2u) lo[2] = hi[2] = 0u;  // neighborhood_green.comp:36
            const uint32_t span = 4u << g.res;
            for (int d = 0; d < 3; ++d) {
                if (hi[d] - lo[d] >= span && hi[d] >= lo[d]) hi[d] = lo[d] + span - 1u;
                s_lo[d] = lo[d];
                s_n[d] = hi[d] >= lo[d] ? hi[d] - lo[d] + 1u : 0u;
            }
            s_mx = Div32::multiplier(s_n[0] ? s_n[0] : 1u);
            s_my = Div32::multiplier(s_n[1] ? s_n[1] : 1u);
            s_pi[0] = pi.x;
            s_pi[1] = pi.y;
            s_pi[2] = pi.z;
            // the tests of k_walk_rows, on unscaled coordinates
            s_trange = ((range != range) ? __builtin_inff() : sqrt_le_bound(range)) * 0x1p36f;
            if (SPREAD) {
                const float tr = wa.on_target ? trad[ip] : 0.0f;
                s_okw = fmax_g(radius[ip], tr) * KERNEL_SCALE;
                s_tcut = sqrt_le_bound(fmax_g(s_okw, kwi)) * 0x1p36f;
                s_tokw = sqrt_le_bound(s_okw) * 0x1p36f;
                s_umax = to_u32(s_okw * KW_RES);
            }
            const uint32_t last_cell = (1u << g.res) - 1u;
            const bool cull = s_trange < __builtin_inff() && range >= 0.0f && hi[0] <= last_cell &&
                              hi[1] <= last_cell && hi[2] <= last_cell;
            s_cull = cull ? 1u : 0u;
            s_thr = cull ? range * range * 1.001f + 1e-30f : __builtin_inff();
            if (FILL) {
                s_braw = row_raw[ip];
                s_bf = SPREAD ? row_f[ip] : s_braw;
            }
        }
        // ---- the batch's particles one at a time
        for (uint64_t am = __ballot(act); am; am &= am - 1ull) {
            const int k = (int)__builtin_ctzll(am);
            const uint32_t i = base + (uint32_t)k;
            const uint32_t sc = lane_value(s_sc, k);
            const GridDesc& g = SEG ? wa.seg.grids[sc] : wa.g;
            const uint32_t cbase = SEG ? (sc << wa.seg.shift) : 0u;
            const int4 pi = make_int4(lane_value(s_pi[0], k), lane_value(s_pi[1], k), lane_value(s_pi[2], k), 0);
            const v3f p = pos_world(pi);
            const v3f pr = vec_i(pi.x, pi.y, pi.z);
            const float t_range = lane_value(s_trange, k);
            const float okw = lane_value(s_okw, k), t_cut = lane_value(s_tcut, k), t_okw = lane_value(s_tokw, k);
            const uint32_t umax = lane_value(s_umax, k);
            const bool cull = lane_value(s_cull, k) != 0u;
            const float thr = lane_value(s_thr, k);
            const uint32_t lo0 = lane_value(s_lo[0], k), lo1 = lane_value(s_lo[1], k), lo2 = lane_value(s_lo[2], k);
            const uint32_t nx = lane_value(s_n[0], k), ny = lane_value(s_n[1], k), nz = lane_value(s_n[2], k);
            auto spread_to = [&](uint32_t j, float d2r, uint32_t cur) {  // kernel_width.comp:41-53
                if (umax <= cur) return;
                uint32_t uval;
                if (d2r <= t_okw) {
                    uval = umax;
                } else {
                    const float dfk = __builtin_sqrtf(d2r * 0x1p-36f) - okw;
                    const float infl = fmax_g(0.0f, 1.0f - fmax_g(0.0f, dfk / (okw * KW_PROPAGATION)));
                    uval = to_u32(okw * infl * KW_RES);
                }
                if (uval > cur) atomicMax(&ukw[j], uval);
            };
            // the cell culling of k_walk_rows (the same cell bounds and margins)
            const uint32_t last_cell = (1u << g.res) - 1u;
            float cw[3], cm[3];
            for (int d = 0; d < 3; ++d) {
                cw[d] = (g.mx[d] - g.mn[d]) / (float)(1u << g.res);
                cm[d] = 1e-4f * cw[d] + 1e-6f * (__builtin_fabsf(g.mn[d]) + __builtin_fabsf(g.mx[d]));
            }
            auto axis_d2 = [&](int d, uint32_t c, float pc) -> float {
                if (!cull || c == 0u || c >= last_cell) return 0.0f;
                const float a = g.mn[d] + (float)c * cw[d] - cm[d];
                const float b = g.mn[d] + (float)(c + 1u) * cw[d] + cm[d];
                const float t = pc < a ? a - pc : (pc > b ? pc - b : 0.0f);
                return t * t;
            };
            uint32_t craw = 0, cf = 0, braw = 0, bf = 0;
            uint64_t tested = 0;
            bool more = true;
            if (FILL) {
                braw = lane_value(s_braw, k);
                bf = lane_value(s_bf, k);
                more = braw < wa.pcap;
            }
            const size_t slot = PAD ? (size_t)i * stride : 0;
            const uint64_t ncell = (uint64_t)nx * ny * nz;
            const bool planar = g.dims == 2u;
            const Div32 divx(nx ? nx : 1u, lane_value(s_mx, k)), divy(ny ? ny : 1u, lane_value(s_my, k));
            for (uint64_t c0 = 0; more && c0 < ncell; c0 += 64) {
                // one cell per lane
                uint32_t cstart = 0, clen = 0;
                const uint64_t c = c0 + lane;
                if (c < ncell) {
                    uint32_t x, y, z;
                    if (ncell <= 0xFFFFFFFFull) {
                        const uint32_t q = divx.div((uint32_t)c, x);
                        z = divy.div(q, y);
                    } else {  // a box of more than 2^32 cells (unbounded ranges at high resolutions)
                        x = (uint32_t)(c % nx);
                        y = (uint32_t)((c / nx) % ny);
                        z = (uint32_t)(c / ((uint64_t)nx * ny));
                    }
                    x += lo0;
                    y += lo1;
                    z += lo2;
                    const float d2 = (axis_d2(2, z, p.z) + axis_d2(1, y, p.y)) + axis_d2(0, x, p.x);
                    if (!(d2 > thr)) {
                        const uint32_t h = cbase + (planar ? (zbits(x, 0u, g.res, 2u) | zbits(y, 1u, g.res, 2u))
                                                           : (zbits(x, 0u, g.res, 3u) | zbits(y, 1u, g.res, 3u) |
                                                              zbits(z, 2u, g.res, 3u)));
                        cstart = cs[h];
                        const uint32_t cend = ce[h];
                        clen = cend > cstart ? cend - cstart : 0u;
                    }
                }
                const uint32_t incl = wave_scan(clen, vadd);
                const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
                const uint32_t off = incl - clen;  // the cell's first candidate in the stream
                // candidate = stream position + delta inside the cell; a cell whose
                // particles directly follow those of the previous non-empty cell
                // (the same delta) continues its run
                const uint32_t delta = cstart - off;
                const uint64_t nonempty = __ballot(clen > 0u);
                const uint64_t prev_mask = nonempty & below;
                const uint32_t prev = prev_mask ? 63u - (uint32_t)__clzll(prev_mask) : lane;
                const uint32_t prev_delta = (uint32_t)__shfl((int)delta, (int)prev, 64);
                const bool start = clen > 0u && (prev_mask == 0ull || prev_delta != delta);
                if (MODE == WALK_COUNT) tested += total;
                for (uint32_t w0 = 0; more && w0 < total; w0 += 64u * WAVE_CANDS) {
                    // WAVE_CANDS candidates per lane: w0 + 64 c + lane. The delta of
                    // each: that of the run containing w0, replaced by each run
                    // starting inside the step from its first candidate on
                    const uint64_t le = __ballot(start && off <= w0);
                    const int r0 = 63 - (int)__clzll(le);  // le != 0: w0 < total
                    uint32_t d[WAVE_CANDS];
                    d[0] = (uint32_t)__builtin_amdgcn_readlane((int)delta, r0);
#pragma unroll
                    for (uint32_t c = 1; c < WAVE_CANDS; ++c) d[c] = d[0];
                    uint64_t inside = __ballot(start && off > w0 && off - w0 < 64u * WAVE_CANDS);
                    while (inside) {
                        const int q = (int)__builtin_ctzll(inside);
                        inside &= inside - 1ull;
                        const uint32_t o = (uint32_t)__builtin_amdgcn_readlane((int)off, q) - w0;
                        const uint32_t dq = (uint32_t)__builtin_amdgcn_readlane((int)delta, q);
#pragma unroll
                        for (uint32_t c = 0; c < WAVE_CANDS; ++c)
                            if (lane + 64u * c >= o) d[c] = dq;
                    }
                    // every candidate's loads first, then the tests in stream order
                    uint32_t jc[WAVE_CANDS], uc[WAVE_CANDS];
                    int4 pc[WAVE_CANDS];
                    bool vc[WAVE_CANDS];
#pragma unroll
                    for (uint32_t c = 0; c < WAVE_CANDS; ++c) {
                        const uint32_t t = w0 + 64u * c + lane;
                        jc[c] = t + d[c];
                        vc[c] = t < total && jc[c] != i;
                        pc[c] = make_int4(0, 0, 0, 0);
                        uc[c] = 0u;
                        if (vc[c]) {
                            pc[c] = pos[jc[c]];
                            if (SPREAD && MODE != WALK_COUNT) uc[c] = ukw[jc[c]];
                        }
                    }
#pragma unroll
                    for (uint32_t c = 0; c < WAVE_CANDS; ++c) {
                        if (c > 0 && !(w0 + 64u * c < total && more)) break;
                        const uint32_t j = jc[c];
                        const int4 pj = pc[c];
                        const uint32_t uj = uc[c];
                        bool green = false, keep = false;
                        float d2 = 0.0f;
                        if (vc[c]) {
                            const v3f u = pr - vec_i(pj.x, pj.y, pj.z);
                            green = !(dot(u, u) > t_range);
                            if (SPREAD && green) {
                                const v3f v = vec_i(isub(pj.x, pi.x), isub(pj.y, pi.y), isub(pj.z, pi.z));
                                d2 = dot(v, v);
                                keep = d2 <= t_cut;
                            }
                        }
                        const uint64_t rbal = __ballot(green);
                        if (PAD) {
                            if (SPREAD && green) spread_to(j, d2, uj);
                            const bool kept = SPREAD ? (green && keep) : green;
                            const uint64_t kbal = SPREAD ? __ballot(kept) : rbal;
                            const uint32_t kk = cf + (uint32_t)__popcll(kbal & below);
                            if (kept && kk < stride) col[slot + kk] = j;
                            cf += (uint32_t)__popcll(kbal);
                            craw += (uint32_t)__popcll(rbal);
                        } else if (!FILL) {
                            craw += (uint32_t)__popcll(rbal);
                            if (SPREAD) cf += (uint32_t)__popcll(__ballot(keep));
                        } else {
                            const uint32_t ridx = braw + craw + (uint32_t)__popcll(rbal & below);
                            const bool incl_r = green && ridx < wa.pcap;
                            if (SPREAD) {
                                if (incl_r) spread_to(j, d2, uj);
                                const bool kept = incl_r && keep;
                                const uint64_t fbal = __ballot(kept);
                                if (kept) col[bf + cf + (uint32_t)__popcll(fbal & below)] = j;
                                cf += (uint32_t)__popcll(fbal);
                            } else if (incl_r) {
                                col[ridx] = j;
                            }
                            craw += (uint32_t)__popcll(rbal);
                            more = braw + craw < wa.pcap;
                        }
                    }
                }
            }
            if (lane == 0) {
                if (PAD) {
                    cnt_raw[i] = craw;
                    if (SPREAD) cnt_f[i] = cf;
                    if (cf > stride) atomicMax(&counts[C_MAXROW], cf);
                } else if (!FILL) {
                    cnt_raw[i] = craw;
                    if (SPREAD) cnt_f[i] = cf;
                    if (wa.cand) atomicAdd(wa.cand, (unsigned long long)tested);
                } else if (SPREAD && braw <= wa.pcap && row_raw[i + 1] > wa.pcap) {
                    counts[C_CLAMP_ROW] = i;
                    counts[C_CLAMP_KEPT] = cf;
                }
            }
        }
    }
    if (!FILL && blockIdx.x == 0 && threadIdx.x == 0) {
        cnt_raw[n] = 0;
        if (SPREAD) cnt_f[n] = 0;
    }
}

// After a PAD walk and the scan of its raw counts: take the exact path when a
// row outgrew its slot or the list exceeds the capacity.
__global__ void k_walk_decide(uint32_t* __restrict__ counts, uint32_t pcap) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t stride = slot_stride(pcap, counts[C_N]);
    counts[C_FALLBACK] = counts[C_PAIRS_RAW] > pcap ? 1u : 0u;  // every row exactly (clamped list)
    counts[C_ROWFIX] = counts[C_MAXROW] > stride ? 1u : 0u;     // only rows longer than their slot
    counts[C_MAXROW_LAST] = counts[C_MAXROW];
    counts[C_MAXROW] = 0u;
}

// Packs the PAD slots into the list at the scanned offsets (16 lanes per row,
// coalesced); rows longer than their slot are left to the exact fill, and
// nothing is packed when the exact fill rewrites every row.
__global__ __launch_bounds__(BLOCK) void k_pack_rows(const uint32_t* __restrict__ counts,
                                                     const uint32_t* __restrict__ row,
                                                     const uint32_t* __restrict__ slots, uint32_t pcap,
                                                     uint32_t* __restrict__ col) {
    if (counts[C_FALLBACK]) return;
    constexpr int L = 16, B = 8;  // a lane has B loads in flight before its stores
    const uint32_t n = counts[C_N];
    const uint32_t stride = slot_stride(pcap, n);
    const int gl = threadIdx.x & (L - 1);
    for (uint32_t base = blockIdx.x * (BLOCK / L); base < n; base += gridDim.x * (BLOCK / L)) {
        const uint32_t i = base + threadIdx.x / L;
        if (i >= n) continue;
        const uint32_t b = row[i], len = row[i + 1] - b;
        if (len > stride) continue;
        const uint32_t* src = slots + (size_t)i * stride;
        for (uint32_t k0 = 0; k0 < len; k0 += (uint32_t)(B * L)) {
            uint32_t v[B];
#pragma unroll
            for (int t = 0; t < B; ++t) {
                const uint32_t k = k0 + (uint32_t)gl + (uint32_t)(t * L);
                v[t] = k < len ? src[k] : 0u;
            }
#pragma unroll
            for (int t = 0; t < B; ++t) {
                const uint32_t k = k0 + (uint32_t)gl + (uint32_t)(t * L);
                if (k < len) col[b + k] = v[t];
            }
        }
    }
}

// After a clamped spreading walk: rows after the cut row are empty, so their
// offsets all equal the end of the cut row's kept entries.
__global__ __launch_bounds__(BLOCK) void k_fix_clamped_rows(const uint32_t* __restrict__ counts, uint32_t pcap,
                                                            uint32_t* __restrict__ row_f) {
    if (counts[C_PAIRS_RAW] <= pcap) return;
    const uint32_t n = counts[C_N];
    const uint32_t p = counts[C_CLAMP_ROW];
    const uint32_t end = row_f[p] + counts[C_CLAMP_KEPT];
    for (uint32_t k = p + 1 + blockIdx.x * blockDim.x + threadIdx.x; k <= n; k += gridDim.x * blockDim.x)
        row_f[k] = end;
}

// neighborhood_brute_force.comp:31-50: particle i against every particle in
// index order, a whole wavefront per particle (64 candidates per step); the
// distance test and the capacity clamp are those of the cell walk, so the
// rows come out in the order a sequential run appends them. COUNT writes the
// row lengths, FILL the rows at the scanned offsets.
template <bool FILL>
__global__ __launch_bounds__(BLOCK) void k_brute_rows(const int4* __restrict__ pos, const float* __restrict__ kw,
                                                      const uint32_t* __restrict__ counts, float range_scale,
                                                      uint32_t pcap, uint32_t* __restrict__ cnt,
                                                      const uint32_t* __restrict__ row, uint32_t* __restrict__ col) {
    constexpr uint32_t L = 64;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t n = counts[C_N];
    const uint32_t waves = BLOCK / 64;
    for (uint32_t i = blockIdx.x * waves + threadIdx.x / 64; i < n; i += gridDim.x * waves) {
        const v3f p = pos_world(pos[i]);
        const float range = kw[i] * range_scale;
        const float t_range = (range != range) ? __builtin_inff() : sqrt_le_bound(range);
        const uint32_t b0 = FILL ? row[i] : 0u;
        uint32_t c = 0;
        bool more = !FILL || b0 < pcap;
        for (uint32_t j0 = 0; more && j0 < n; j0 += L) {
            const uint32_t j = j0 + lane;
            bool green = false;
            if (j < n && j != i) {
                const v3f d = p - pos_world(pos[j]);
                green = !(dot(d, d) > t_range);
            }
            const uint64_t bal = __ballot(green);
            if (FILL) {
                const uint32_t ridx = b0 + c + (uint32_t)__popcll(bal & below);
                if (green && ridx < pcap) col[ridx] = j;
            }
            c += (uint32_t)__popcll(bal);
            if (FILL) more = b0 + c < pcap;
        }
        if (!FILL && lane == 0) cnt[i] = c;
    }
    if (!FILL && blockIdx.x == 0 && threadIdx.x == 0) cnt[n] = 0;
}

void launch_brute(bool fill, const int4* pos, const float* kw, const uint32_t* counts, uint32_t cap,
                  float range_scale, uint32_t pcap, uint32_t* cnt, const uint32_t* row, uint32_t* col,
                  hipStream_t s) {
    uint32_t g = div_up(cap ? cap : 1, BLOCK / 64);
    g = g < 16384u ? g : 16384u;
    if (fill)
        hipLaunchKernelGGL(k_brute_rows<true>, dim3(g), dim3(BLOCK), 0, s, pos, kw, counts, range_scale, pcap, cnt,
                           row, col);
    else
        hipLaunchKernelGGL(k_brute_rows<false>, dim3(g), dim3(BLOCK), 0, s, pos, kw, counts, range_scale, pcap, cnt,
                           row, col);
}

static uint32_t walk_grid(uint32_t cap, uint32_t lanes) {
    const uint32_t groups = BLOCK / lanes;
    uint32_t g = div_up(cap ? cap : 1, groups);
    return g < 16384u ? g : 16384u;
}

void launch_walk(int mode, bool spread, const int4* pos, const float* kw, const float* radius, const float* trad,
                 int on_target, const uint32_t* cs, const uint32_t* ce, uint32_t* counts, uint32_t cap,
                 float range_scale, const GridDesc& g, uint32_t pcap, uint32_t* cnt_raw, uint32_t* cnt_f,
                 const uint32_t* row_raw, const uint32_t* row_f, uint32_t* col, uint32_t* ukw, hipStream_t s,
                 int lanes, uint32_t stride, const uint32_t* gate, const SegmentTable* seg,
                 unsigned long long* cand) {
    WalkArgs wa;
    wa.range_scale = range_scale;
    wa.g = g;
    wa.on_target = on_target;
    wa.pcap = pcap;
    wa.stride = stride;
    wa.gate = gate;
    wa.cand = cand;
    wa.seg = seg ? *seg : SegmentTable{nullptr, nullptr, nullptr, nullptr, 0u};
    const bool packed = seg && seg->sid;
    if (lanes != 4 && lanes != 16 && lanes != 64) lanes = WALK_LANES;
    if (packed && lanes != 64) lanes = WALK_LANES;  // packed batches: the default group width or whole waves
    const dim3 grid(walk_grid(cap, (uint32_t)lanes)), block(BLOCK);
    if (lanes == 64) {
        const dim3 wgrid(walk_grid(div_up(cap ? cap : 1u, WAVE_BATCH), 64u));
#define WAVE(F, S, G)                                                                                                  \
    hipLaunchKernelGGL((k_walk_wave<F, S, G>), wgrid, block, 0, s, pos, kw, radius, trad, cs, ce, counts, wa, cnt_raw, \
                       cnt_f, row_raw, row_f, col, ukw)
#define WAVE_G(G)                         \
    if (mode == WALK_PAD) {               \
        if (spread)                       \
            WAVE(WALK_PAD, true, G);      \
        else                              \
            WAVE(WALK_PAD, false, G);     \
    } else if (mode == WALK_FILL) {       \
        if (spread)                       \
            WAVE(WALK_FILL, true, G);     \
        else                              \
            WAVE(WALK_FILL, false, G);    \
    } else {                              \
        if (spread)                       \
            WAVE(WALK_COUNT, true, G);    \
        else                              \
            WAVE(WALK_COUNT, false, G);   \
    }
        if (packed) {
            WAVE_G(true)
        } else {
            WAVE_G(false)
        }
#undef WAVE_G
#undef WAVE
        return;
    }
#define WALK(L, F, S, G)                                                                                                 \
    hipLaunchKernelGGL((k_walk_rows<L, F, S, G>), grid, block, 0, s, pos, kw, radius, trad, cs, ce, counts, wa, cnt_raw, \
                       cnt_f, row_raw, row_f, col, ukw)
#define WALK_L(L, G)                       \
    if (mode == WALK_PAD) {                \
        if (spread)                        \
            WALK(L, WALK_PAD, true, G);    \
        else                               \
            WALK(L, WALK_PAD, false, G);   \
    } else if (mode == WALK_FILL) {        \
        if (spread)                        \
            WALK(L, WALK_FILL, true, G);   \
        else                               \
            WALK(L, WALK_FILL, false, G);  \
    } else {                               \
        if (spread)                        \
            WALK(L, WALK_COUNT, true, G);  \
        else                               \
            WALK(L, WALK_COUNT, false, G); \
    }
    if (packed) {
        WALK_L(WALK_LANES, true)
    } else if (lanes == 4) {
        WALK_L(4, false)
    } else if (lanes == 16) {
        WALK_L(16, false)
    } else {
        WALK_L(WALK_LANES, false)
    }
#undef WALK_L
#undef WALK
}

void launch_walk_decide(uint32_t* counts, uint32_t pcap, hipStream_t s) {
    hipLaunchKernelGGL(k_walk_decide, dim3(1), dim3(64), 0, s, counts, pcap);
}
void launch_pack_rows(const uint32_t* counts, uint32_t cap, const uint32_t* row, const uint32_t* slots,
                      uint32_t pcap, uint32_t* col, hipStream_t s) {
    uint32_t g = div_up(cap ? cap : 1, BLOCK / 16);
    hipLaunchKernelGGL(k_pack_rows, dim3(g < 8192u ? g : 8192u), dim3(BLOCK), 0, s, counts, row, slots, pcap, col);
}

void launch_fix_clamped_rows(const uint32_t* counts, uint32_t cap, uint32_t pcap, uint32_t* row_f, hipStream_t s) {
    hipLaunchKernelGGL(k_fix_clamped_rows, dim3(stride_grid(cap + 1)), dim3(BLOCK), 0, s, counts, pcap, row_f);
}

}  // namespace apbf
