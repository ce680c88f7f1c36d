// Kernels of the plain (non-adaptive) part of a substep.
//
// Particles are stored so that fluid id == hidden slot (the fluid index list
// of the reference is the identity once every hidden particle is a fluid
// particle and index lists are rebuilt in ascending hidden-slot order), so the
// per-fluid-id and per-hidden-slot arrays share one index.
//
// Neighbour lists are CSR (row offsets + neighbour ids) instead of the
// reference's unordered uvec2 pair list: row i holds the neighbours of
// particle i in the order of the reference's cell walk, which makes every
// per-pair shader a gather over the row (no atomics) and keeps the sums exact
// because the reference accumulates per-pair truncated integers.
#include "plain_kernels.h"

namespace apbf {


// velocity_handling::apply (velocity_handling.cpp:15-31): infer_velocity.comp:25-26,
// position backup, apply_acceleration.comp:23-24, apply_velocity.comp:24-25, fused.
__global__ __launch_bounds__(BLOCK) void k_integrate(int4* __restrict__ pos, int4* __restrict__ backup,
                                                     float4* __restrict__ vel,
                                                     const uint32_t* __restrict__ counts, float last_dt,
                                                     float dt, float ax, float ay, float az) {
    const uint32_t n = counts[C_N];
    const float denom = last_dt * POS_RES;
    const float s = dt * POS_RES;
    GRID_STRIDE(i, n) {
        int4 p = pos[i];
        int4 b = backup[i];
        float4 v = vel[i];
        v.x = (float)isub(p.x, b.x) / denom;
        v.y = (float)isub(p.y, b.y) / denom;
        v.z = (float)isub(p.z, b.z) / denom;
        backup[i] = p;
        v.x += ax;
        v.y += ay;
        v.z += az;
        vel[i] = v;
        p.x = iadd(p.x, to_i32(v.x * s));
        p.y = iadd(p.y, to_i32(v.y * s));
        p.z = iadd(p.z, to_i32(v.z * s));
        pos[i] = p;
    }
}

__global__ __launch_bounds__(BLOCK) void k_copy_int4(int4* __restrict__ dst, const int4* __restrict__ src,
                                                     const uint32_t* __restrict__ counts) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) dst[i] = src[i];
}

// calculate_position_hash.comp:39-46 over every hidden particle, with the
// index sequence of neighborhood_green.cpp:75.
__global__ __launch_bounds__(BLOCK) void k_hash_iota(const int4* __restrict__ pos,
                                                     const uint32_t* __restrict__ d_len, GridDesc g,
                                                     uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals) {
    const uint32_t n = *d_len;
    GRID_STRIDE(i, n) {
        v3f p = pos_world(pos[i]);
        uint32_t gx = grid_cell(p.x, g.mn[0], g.mx[0], g.res);
        uint32_t gy = grid_cell(p.y, g.mn[1], g.mx[1], g.res);
        uint32_t gz = grid_cell(p.z, g.mn[2], g.mx[2], g.res);
        keys[i] = zhash(gx, gy, gz, g.res, g.dims);
        if (vals) vals[i] = i;
    }
}
// A packed batch: the hash in the particle's own scene grid, above it the
// scene index, so one stable sort orders by scene, then by cell.
__global__ __launch_bounds__(BLOCK) void k_hash_iota_seg(const int4* __restrict__ pos,
                                                         const uint32_t* __restrict__ d_len, SegmentTable seg,
                                                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t n = *d_len;
    GRID_STRIDE(i, n) {
        const uint32_t sc = seg.sid[i];
        const GridDesc& g = seg.grids[sc];
        v3f p = pos_world(pos[i]);
        uint32_t gx = grid_cell(p.x, g.mn[0], g.mx[0], g.res);
        uint32_t gy = grid_cell(p.y, g.mn[1], g.mx[1], g.res);
        uint32_t gz = grid_cell(p.z, g.mn[2], g.mx[2], g.res);
        keys[i] = (sc << seg.shift) | zhash(gx, gy, gz, g.res, g.dims);
        if (vals) vals[i] = i;
    }
}
__global__ __launch_bounds__(BLOCK) void k_segment_starts(const uint32_t* __restrict__ sid,
                                                          const uint32_t* __restrict__ counts,
                                                          uint32_t* __restrict__ seg_start) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) {
        const uint32_t s = sid[i];
        if (i == 0 || sid[i - 1] != s) seg_start[s] = i;
    }
}

// calculate_position_code.comp:23-62: bit b of the 96-bit code is bit b/3 of
// coordinate b%3 (positions reinterpreted as uint).
__global__ __launch_bounds__(BLOCK) void k_position_code(const int4* __restrict__ pos,
                                                         const uint32_t* __restrict__ d_len,
                                                         uint32_t section, uint32_t* __restrict__ out) {
    const uint32_t n = *d_len;
    GRID_STRIDE(i, n) {
        int4 p = pos[i];
        uint32_t u[3] = {(uint32_t)p.x, (uint32_t)p.y, (uint32_t)p.z};
        uint32_t r = 0;
#pragma unroll
        for (uint32_t k = 0; k < 32u; ++k) {
            uint32_t b = section * 32u + k;
            r |= ((u[b % 3u] >> (b / 3u)) & 1u) << k;
        }
        out[i] = r;
    }
}

// Gather of every particle/fluid array by the sort permutation (the
// hidden_list().apply_edit of neighborhood_green.cpp:54 plus the owner
// rebuild of indexed_list.h:289-308, which with fluid id == hidden slot is the
// same permutation for the fluid attributes).
__global__ __launch_bounds__(BLOCK) void k_permute(const uint32_t* __restrict__ perm,
                                                   const uint32_t* __restrict__ counts, ParticleArrays src,
                                                   ParticleArrays dst) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) {
        uint32_t o = perm[i];
        dst.pos[i] = src.pos[o];
        dst.vel[i] = src.vel[o];
        dst.inv_mass[i] = src.inv_mass[o];
        dst.radius[i] = src.radius[o];
        dst.backup[i] = src.backup[o];
        dst.transferring[i] = src.transferring[o];
        dst.target_radius[i] = src.target_radius[o];
        dst.kernel_width[i] = src.kernel_width[o];
        dst.boundariness[i] = src.boundariness[o];
        dst.boundary_distance[i] = src.boundary_distance[o];
        if (src.lam_prev) dst.lam_prev[i] = src.lam_prev[o];
        if (src.sid) dst.sid[i] = src.sid[o];
        if (src.prev_old) dst.prev_old[i] = o;
    }
}

// find_value_ranges.comp:16-31; particle 0 has no predecessor.
__global__ __launch_bounds__(BLOCK) void k_ranges(const uint32_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ d_len,
                                                  uint32_t* __restrict__ start, uint32_t* __restrict__ end) {
    const uint32_t n = *d_len;
    GRID_STRIDE(i, n) {
        uint32_t cur = keys[i];
        if (i == 0) {
            start[cur] = 0;
        } else {
            uint32_t prev = keys[i - 1];
            if (cur != prev) {
                start[cur] = i;
                end[prev] = i;
            }
        }
        if (i == n - 1) end[cur] = i + 1;
    }
}

// Pair totals after the row-offset scan: clamp to the capacity and flag overflow.
__global__ void k_pairs_total(uint32_t* __restrict__ counts, const uint32_t* __restrict__ row,
                              uint32_t pair_cap) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t n = counts[C_N];
    uint32_t total = row[n];
    counts[C_ROWS] = n;
    counts[C_PAIRS_RAW] = total;
    if (total > counts[C_PAIRS_PEAK]) counts[C_PAIRS_PEAK] = total;
    if (total > pair_cap) {
        counts[C_PAIRS] = pair_cap;
        counts[C_OVERFLOW] |= 1u;
    } else {
        counts[C_PAIRS] = total;
    }
}

__global__ void k_set_n_plus_1(uint32_t* __restrict__ counts) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        counts[C_N_PLUS_1] = counts[C_N] + 1u;
        counts[C_N4] = 4u * counts[C_N];
    }
}

// Kernel constants per particle (depend on the kernel width only).
// Per particle: the Gaussian constants, the rest-density factor
// pow(2 r, D) * inverse mass of incompressibility_2.comp:77, and the
// particle's own density term of incompressibility_0.comp:43 (as integer bits).
__global__ __launch_bounds__(BLOCK) void k_kernel_consts(const float* __restrict__ kw,
                                                         const float* __restrict__ radius,
                                                         const float* __restrict__ inv_mass,
                                                         const uint32_t* __restrict__ counts, int hid, int gid,
                                                         float dim, float4* __restrict__ gc) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) {
        const float k = kw[i];
        const float im = inv_mass[i];
        const GaussConst g = gauss_const(k, dim);
        float w0;
        v3f unused;
        kernel_eval(hid, gid, dim, mk(0, 0, 0), k, g, true, false, &w0, &unused);
        const uint32_t self = to_u32(w0 / im * INCOMP_RES);
        gc[i] = make_float4(g.ivd, g.norm, pow_f(2.0f * radius[i], dim) * im, __builtin_bit_cast(float, self));
    }
}

// box_collision.comp:33-55 over every particle.
__global__ __launch_bounds__(BLOCK) void k_box(int4* __restrict__ pos, const float* __restrict__ radius,
                                               const uint32_t* __restrict__ counts, BoxSet boxes) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) pos[i] = box_collide(i, pos[i], radius[i], boxes);
}
// A packed batch: each particle against its own scene's boxes, with its list
// position inside that scene.
__global__ __launch_bounds__(BLOCK) void k_box_seg(int4* __restrict__ pos, const float* __restrict__ radius,
                                                   const uint32_t* __restrict__ counts, SegmentTable seg) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) {
        const uint32_t sc = seg.sid[i];
        const BoxSet& b = seg.boxes[sc];
        if (b.n) pos[i] = box_collide(i - seg.seg_start[sc], pos[i], radius[i], b);
    }
}
// sphere_collision::apply (sphere_collision.cpp): one pass over the particles.
__global__ __launch_bounds__(BLOCK) void k_sphere(int4* __restrict__ pos, const float* __restrict__ radius,
                                                  const uint32_t* __restrict__ counts, BoxSet c) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) pos[i] = sphere_collide(i, pos[i], radius[i], c);
}

// Adds one substep's mean and max compression error to the running statistics
// (stats[0] += mean, stats[1] = max(stats[1], max)); stats[2] keeps the last mean.
__global__ __launch_bounds__(BLOCK) void k_density_error_finish(const double* __restrict__ part_sum,
                                                                const float* __restrict__ part_max,
                                                                uint32_t blocks, uint32_t* __restrict__ counts,
                                                                double* __restrict__ stats) {
    __shared__ double ssum[BLOCK];
    __shared__ double sabs[BLOCK];
    __shared__ double smax[BLOCK];
    double a = 0.0, ab = 0.0, m = 0.0;
    for (uint32_t b = threadIdx.x; b < blocks; b += BLOCK) {
        a += part_sum[b];
        ab += part_sum[blocks + b];
        m = m > (double)part_max[b] ? m : (double)part_max[b];
    }
    ssum[threadIdx.x] = a;
    sabs[threadIdx.x] = ab;
    smax[threadIdx.x] = m;
    __syncthreads();
    for (int s = BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            ssum[threadIdx.x] += ssum[threadIdx.x + s];
            sabs[threadIdx.x] += sabs[threadIdx.x + s];
            smax[threadIdx.x] = smax[threadIdx.x] > smax[threadIdx.x + s] ? smax[threadIdx.x] : smax[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // stats: sum of per-substep means, max, last mean, last max, sum of
        // per-substep absolute means, last absolute mean
        uint32_t n = counts[C_N];
        double mean = n ? ssum[0] / (double)n : 0.0;
        double mean_abs = n ? sabs[0] / (double)n : 0.0;
        stats[0] += mean;
        stats[1] = stats[1] > smax[0] ? stats[1] : smax[0];
        stats[2] = mean;
        stats[3] = smax[0];
        stats[4] += mean_abs;
        stats[5] = mean_abs;
        counts[C_ERR_SUBSTEPS] += 1u;
    }
}

__global__ void k_count_particle_substeps(const uint32_t* __restrict__ counts,
                                          unsigned long long* __restrict__ acc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        acc[0] += (unsigned long long)counts[C_ROWS];
        acc[1] += (unsigned long long)counts[C_PAIRS];
        acc[2] += (unsigned long long)counts[C_PAIRS_RAW];
    }
}

// ------------------------------------------------------------ launch wrappers
#define LAUNCH(kern, grid, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, s, __VA_ARGS__)

void launch_integrate(int4* pos, int4* backup, float4* vel, const uint32_t* counts, uint32_t cap,
                      float last_dt, float dt, const float acc[3], hipStream_t s) {
    LAUNCH(k_integrate, stride_grid(cap), pos, backup, vel, counts, last_dt, dt, acc[0], acc[1], acc[2]);
}
void launch_copy_int4(int4* dst, const int4* src, const uint32_t* counts, uint32_t cap, hipStream_t s) {
    LAUNCH(k_copy_int4, stride_grid(cap), dst, src, counts);
}
void launch_hash_iota(const int4* pos, const uint32_t* d_len, uint32_t cap, const GridDesc& g,
                      uint32_t* keys, uint32_t* vals, hipStream_t s, const SegmentTable* seg) {
    if (seg && seg->sid)
        LAUNCH(k_hash_iota_seg, stride_grid(cap), pos, d_len, *seg, keys, vals);
    else
        LAUNCH(k_hash_iota, stride_grid(cap), pos, d_len, g, keys, vals);
}
void launch_segment_starts(const uint32_t* sid, const uint32_t* counts, uint32_t cap, uint32_t* seg_start,
                           hipStream_t s) {
    LAUNCH(k_segment_starts, stride_grid(cap), sid, counts, seg_start);
}
void launch_position_code(const int4* pos, const uint32_t* d_len, uint32_t cap, uint32_t section,
                          uint32_t* out, hipStream_t s) {
    LAUNCH(k_position_code, stride_grid(cap), pos, d_len, section, out);
}
void launch_permute(const uint32_t* perm, const uint32_t* counts, uint32_t cap,
                    const ParticleArrays& src, const ParticleArrays& dst, hipStream_t s) {
    LAUNCH(k_permute, stride_grid(cap), perm, counts, src, dst);
}
void launch_ranges(const uint32_t* keys, const uint32_t* d_len, uint32_t cap, uint32_t* start,
                   uint32_t* end, hipStream_t s) {
    LAUNCH(k_ranges, stride_grid(cap), keys, d_len, start, end);
}
void launch_pairs_total(uint32_t* counts, const uint32_t* row, uint32_t pair_cap, hipStream_t s) {
    hipLaunchKernelGGL(k_pairs_total, dim3(1), dim3(64), 0, s, counts, row, pair_cap);
}
void launch_set_n_plus_1(uint32_t* counts, hipStream_t s) {
    hipLaunchKernelGGL(k_set_n_plus_1, dim3(1), dim3(64), 0, s, counts);
}
void launch_kernel_consts(const float* kw, const float* radius, const float* inv_mass, const uint32_t* counts,
                          uint32_t cap, int hid, int gid, float dim, float4* gc,
                          hipStream_t s) {
    LAUNCH(k_kernel_consts, stride_grid(cap), kw, radius, inv_mass, counts, hid, gid, dim, gc);
}
void launch_box(int4* pos, const float* radius, const uint32_t* counts, uint32_t cap,
                const BoxSet& boxes, hipStream_t s, const SegmentTable* seg) {
    if (seg && seg->sid)
        LAUNCH(k_box_seg, stride_grid(cap), pos, radius, counts, *seg);
    else
        LAUNCH(k_box, stride_grid(cap), pos, radius, counts, boxes);
}
void launch_sphere(int4* pos, const float* radius, const uint32_t* counts, uint32_t cap, const BoxSet& c,
                   hipStream_t s) {
    LAUNCH(k_sphere, stride_grid(cap), pos, radius, counts, c);
}
void launch_count_particle_substeps(const uint32_t* counts, unsigned long long* acc, hipStream_t s) {
    hipLaunchKernelGGL(k_count_particle_substeps, dim3(1), dim3(64), 0, s, counts, acc);
}
void launch_density_error_finish(const double* part_sum, const float* part_max, uint32_t blocks,
                                 uint32_t* counts, double* stats, hipStream_t s) {
    hipLaunchKernelGGL(k_density_error_finish, dim3(1), dim3(BLOCK), 0, s, part_sum, part_max, blocks, counts,
                       stats);
}

}  // namespace apbf
