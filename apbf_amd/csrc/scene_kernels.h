// Device state of one scene and the kernels of a substep.
#pragma once
#include "apbf_internal.h"
#include "apbf_sph.h"

namespace apbf {

// Device-resident list lengths and flags (the reference's per-list 4-byte
// length buffers, gathered into one block).
enum CountSlot : uint32_t {
    C_N = 0,           // particles (hidden list length == fluid list length)
    C_N_PLUS_1 = 1,    // n + 1, length of the CSR offset arrays
    C_PAIRS = 2,       // forward neighbour pairs (after the capacity clamp)
    C_TRANSFERS = 3,   // hidden transfer list length
    C_SPLITS = 4,      // split list length
    C_PROPOSALS = 5,   // merge/split proposals of find_split_and_merge_3
    C_OVERFLOW = 6,    // bit 0 pairs, bit 1 transfers, bit 2 particles
    C_PAIRS_RAW = 7,   // neighbour pairs before the clamp
    C_KEEP = 8,        // particles surviving a compaction
    C_TKEEP = 9,       // transfers surviving a compaction
    C_UNDECIDED = 10,  // claims left after the parallel rounds
    C_MERGES = 11,     // merges won in find_split_and_merge_3
    C_TMP0 = 12,
    C_TMP1 = 13,
    C_ROWS = 14,       // particles covered by the CSR rows (n when the list was built)
    C_ERR_SUBSTEPS = 15,  // substeps accumulated in the density-error statistics
    C_N4 = 16,         // 4 n (words of the solver's shift accumulator)
    C_PAIRS_PEAK = 17, // largest neighbour-walk pair count (before the clamp) since the scene was created
    C_CLAMP_ROW = 18,  // row cut by the pair capacity in a spreading walk
    C_CLAMP_KEPT = 19, // entries of that row kept in the short list
    C_MAXROW = 20,     // single-pass walk: longest row that outgrew its slot (0: none)
    C_FALLBACK = 21,   // single-pass walk: 1 when the exact fill must rewrite every row (clamped list)
    C_ROWFIX = 22,     // single-pass walk: 1 when some rows outgrew their slot (exact fill of those rows)
    C_SOLVE_DONE = 23,     // tolerance mode: the density error reached the tolerance; later iterations do nothing
    C_SOLVE_ACTION = 24,   // what the current iteration's apply pass does (SolveAction)
    C_SOLVE_ITERS = 25,    // solver iterations whose shifts were applied in this substep
    C_WARM_FALLBACK = 26,  // 1 when this substep's warm start was undone (the cold start judged better)
    C_MAXROW_LAST = 27,    // single-pass walk: longest row that outgrew its slot in the last walk (0: none)
    C_CANDIDATES = 28,     // 28-29: candidates tested by the diagnostic counting walks (a 64-bit sum)
    C_LONG = 30,           // rows in the solver's first long-row list (longer than a row group's cache)
    C_LONG2 = 31,          // rows in its second list (longer than a 32-lane group's cache)
    C_COUNT = 32
};

// The apply pass of one solver iteration, as decided from that iteration's
// density error (tolerance mode, warm-start fallback).
enum SolveAction : uint32_t {
    SOLVE_APPLY = 0,    // add the shifts
    SOLVE_SKIP = 1,     // leave the positions (the error is within the tolerance, or the solve is done)
    SOLVE_RESTORE = 2,  // put back the positions saved before the warm-start shift
};

struct GridDesc {
    float mn[3];
    float mx[3];
    uint32_t res;
    uint32_t dims;  // DIMENSIONS (cpu_gpu_shared_config.h:2): 2 or 3
};

// The scene's colliders: up to APBF_MAX_BOXES boxes (box_collision) and one
// sphere (sphere_collision); `sphere` = 0 when the scene has none.
struct BoxSet {
    uint32_t n;
    float mn[APBF_MAX_BOXES][3];
    float mx[APBF_MAX_BOXES][3];
    uint32_t sphere;
    int32_t hollow;
    float sc[3];
    float sr;
};

// Several independent scenes packed into one particle list (a packed batch):
// each particle carries its scene's index (sid), the scenes' hash grids and
// colliders are tables, every scene owns a block of 2^shift cells of the
// cell tables, and seg_start[s] is the first list position of scene s after
// the cell sort (list positions inside a scene, the reference's invocation
// ids, are position - seg_start). All null / 0 for a single scene.
struct SegmentTable {
    const GridDesc* grids;
    const BoxSet* boxes;
    const uint32_t* seg_start;
    const uint32_t* sid;
    uint32_t shift;
};

struct SolverArgs {
    int hid, gid, method, update_b;
    float b_speed, b_sgl_factor, b_under_factor, lambda_scale_base;
    float dim;  // DIMENSIONS, 2 or 3
    double* it_err;  // when set: += mean density error of the positions this iteration starts from
    const float* lam_override;  // when set: shift with these lambdas, no boundariness update or filter
    float* lam_acc;             // when set: += min(lambda, 0) per particle
    double* tol_sum;            // when set: += sum of max(0, rho/rho0 - 1) over the positions the pass starts from
    const uint32_t* skip;       // when set and *skip != 0: the pass does nothing (the solve is done)
    uint32_t split_len;         // when nonzero: rows of more entries are left to a pass over `list`
    const uint32_t* list;       // when set: the pass takes the rows listed here (counts[list_slot] of them)
    uint32_t list_slot;         // C_LONG or C_LONG2
};

// -------------------------------------------------------- shared device code
APBF_HD uint32_t grid_cell(float p, float lo, float hi, uint32_t res) {
    return to_u32((p - lo) / (hi - lo) * (float)(1u << res));
}
APBF_HD uint32_t spread_bits3(uint32_t v) {  // bit i -> bit 3i, for the low 10 bits
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
APBF_HD uint32_t spread_bits2(uint32_t v) {  // bit i -> bit 2i, for the low 16 bits
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}
// The z-curve of calculate_position_hash.comp:29-36 (bit i of axis d -> bit
// D i + d for i < res, d < D), split by axis: the bits of cell coordinate c
// on axis d. The hash of a cell is the OR over its axes.
APBF_HD uint32_t zbits(uint32_t c, uint32_t d, uint32_t res, uint32_t dims) {
    const uint32_t m = (res >= 32u) ? 0xFFFFFFFFu : ((1u << res) - 1u);
    return dims == 2u ? spread_bits2(c & m) << d : spread_bits3(c & m) << d;
}
// The z-curve hash of a cell; in 2D the z coordinate is not part of it.
APBF_HD uint32_t zhash(uint32_t gx, uint32_t gy, uint32_t gz, uint32_t res, uint32_t dims) {
    const uint32_t h = zbits(gx, 0u, res, dims) | zbits(gy, 1u, res, dims);
    return dims == 2u ? h : h | zbits(gz, 2u, res, dims);
}
APBF_HD v3f pos_world(int4 p) { return vec_i(p.x, p.y, p.z) / POS_RES; }

// box_collision.comp:15-20 ("Hash without Sine", MIT licence)
APBF_HD v3f hash31(float p) {
    v3f p3 = mk(fract(p * 0.1031f), fract(p * 0.1030f), fract(p * 0.0973f));
    float d = dot(p3, mk(p3.y + 33.33f, p3.z + 33.33f, p3.x + 33.33f));
    p3 = mk(p3.x + d, p3.y + d, p3.z + d);
    return mk(fract((p3.x + p3.y) * p3.z), fract((p3.x + p3.z) * p3.y), fract((p3.y + p3.z) * p3.x));
}
APBF_HD v3f smallest_mask(v3f v) {  // box_collision.comp:22-28
    bool b1 = v.x <= v.y, b2 = v.x <= v.z, b3 = v.y <= v.z;
    return mk((b1 && b2) ? 1.0f : 0.0f, (!b1 && b3) ? 1.0f : 0.0f, (!b2 && !b3) ? 1.0f : 0.0f);
}
// box_collision.comp:36-54 for the particle with list position `id`.
APBF_HD int4 box_collide(uint32_t id, int4 p, float r, const BoxSet& boxes) {
    v3f pos = pos_world(p);
    for (uint32_t i = 0; i < boxes.n; ++i) {
        v3f jmin = hash31((float)id * pos.x - pos.y - pos.z) * 0.05f;
        v3f bmin = (mk(boxes.mn[i][0], boxes.mn[i][1], boxes.mn[i][2]) - mk(r, r, r)) - jmin;
        v3f jmax = hash31((float)id * pos.y + pos.x + pos.z) * 0.05f;
        v3f bmax = (mk(boxes.mx[i][0], boxes.mx[i][1], boxes.mx[i][2]) + mk(r, r, r)) + jmax;
        v3f toMin = bmin - pos, toMax = bmax - pos;
        toMin = toMin * smallest_mask(-toMin);
        toMax = toMax * smallest_mask(toMax);
        float dMin = -dot(toMin, mk(1, 1, 1));
        float dMax = dot(toMax, mk(1, 1, 1));
        if (dMin <= 0.0f || dMax <= 0.0f) continue;
        pos = pos + (dMin < dMax ? toMin : toMax);
    }
    return make_int4(to_i32(pos.x * POS_RES), to_i32(pos.y * POS_RES), to_i32(pos.z * POS_RES), p.w);
}

// sphere_collision.comp:24-30 ("Hash without Sine", MIT licence)
APBF_HD float hash11(float p) {
    p = fract(p * 0.1031f);
    p *= p + 33.33f;
    p *= p + p;
    return fract(p);
}
// sphere_collision.comp:42-60 for the particle with list position `id`: a
// hollow sphere keeps particles inside (radius - r - jitter), a solid one
// outside; a particle already on its side is left untouched.
APBF_HD int4 sphere_collide(uint32_t id, int4 p, float r, const BoxSet& c) {
    v3f pos = pos_world(p);
    const float dir = c.hollow ? -1.0f : 1.0f;
    const float offset = dir * (r + hash11((float)id * pos.y + pos.x + pos.z) * 0.05f);
    const v3f center = mk(c.sc[0], c.sc[1], c.sc[2]);
    const v3f rel = pos - center;
    const float d = length(rel);
    const float target = c.sr + offset;
    if ((d < target) == (c.hollow != 0)) return p;
    pos = center + rel * (target / d);
    return make_int4(to_i32(pos.x * POS_RES), to_i32(pos.y * POS_RES), to_i32(pos.z * POS_RES), p.w);
}

// incompressibility_2.comp:37-41
APBF_HD float move_towards_add(float oldv, float newv, float max_step) {
    float step = newv - oldv;
    return oldv + fmin_g(max_step, fmax_g(-max_step, step));
}
// uint_to_float_but_gradual.comp:18-24 / uint_to_float_with_indexed_lower_bound.comp:20-26
APBF_HD float move_towards_mult(float oldv, float newv, float step) {
    float dir = sign_f(newv - oldv);
    float result = oldv * (1.0f + step * dir);
    bool reached = (oldv < newv) != (result < newv);
    return reached ? newv : result;
}

}  // namespace apbf
