// Stage operators over caller-owned lists (include/apbf.h, "stage operators").
//
// The reference's subsystems bind to lists (incompressibility::set_data(fluid*,
// neighbors*), update_transfers::set_data(fluid*, neighbors*, transfers*), ...)
// and run their shaders on those lists' buffers. Here a context holds the
// temporaries of every stage for given capacities (the reference draws them
// from its list pool inside apply(), gpu_list_data.cpp:6-45), and each
// operator runs one stage on the caller's buffers:
//  - the caller's particle, transfer and neighbour arrays are put in place of
//    the context's first buffer set for the duration of the call;
//  - the stage runs the same kernel sequence as a scene's stage;
//  - stages that gather out of place (the neighbourhood's permutation, the
//    compactions of particle_transfer, the spreading's new list) leave their
//    result in the context's second set, which is copied back into the
//    caller's arrays at the end of the call;
//  - list lengths are read from and written back to the caller's device
//    length words; nothing is read back to the host.
#include <algorithm>
#include <new>

#include "scene.h"

namespace apbf {

// Count block from the caller's device lengths (particles, transfers, pairs,
// rows covered by the neighbour list).
__global__ void k_bind_counts(uint32_t* __restrict__ counts, const uint32_t* __restrict__ d_n,
                              const uint32_t* __restrict__ d_nt, const uint32_t* __restrict__ d_np,
                              const uint32_t* __restrict__ d_rows) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t n = *d_n;
    counts[C_N] = n;
    counts[C_N_PLUS_1] = n + 1u;
    counts[C_N4] = 4u * n;
    counts[C_TRANSFERS] = d_nt ? *d_nt : 0u;
    counts[C_PAIRS] = d_np ? *d_np : 0u;
    counts[C_ROWS] = d_rows ? *d_rows : 0u;
}
__global__ void k_unbind_counts(const uint32_t* __restrict__ counts, uint32_t* __restrict__ d_n,
                                uint32_t* __restrict__ d_nt, uint32_t* __restrict__ d_np,
                                uint32_t* __restrict__ d_rows) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *d_n = counts[C_N];
    if (d_nt) *d_nt = counts[C_TRANSFERS];
    if (d_np) *d_np = counts[C_PAIRS];
    if (d_rows) *d_rows = counts[C_ROWS];
}
__global__ __launch_bounds__(BLOCK) void k_copy_particles(const uint32_t* __restrict__ counts, ParticleArrays src,
                                                          ParticleArrays dst) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) {
        dst.pos[i] = src.pos[i];
        dst.vel[i] = src.vel[i];
        dst.inv_mass[i] = src.inv_mass[i];
        dst.radius[i] = src.radius[i];
        dst.backup[i] = src.backup[i];
        dst.transferring[i] = src.transferring[i];
        dst.target_radius[i] = src.target_radius[i];
        dst.kernel_width[i] = src.kernel_width[i];
        dst.boundariness[i] = src.boundariness[i];
        dst.boundary_distance[i] = src.boundary_distance[i];
        dst.lam_prev[i] = src.lam_prev[i];
        if (src.sid) dst.sid[i] = src.sid[i];
    }
}
__global__ __launch_bounds__(BLOCK) void k_copy_transfers(const uint32_t* __restrict__ counts,
                                                          const uint32_t* __restrict__ ss, const uint32_t* __restrict__ st,
                                                          const float* __restrict__ sl, uint32_t* __restrict__ ds,
                                                          uint32_t* __restrict__ dt, float* __restrict__ dl) {
    const uint32_t n = counts[C_TRANSFERS];
    GRID_STRIDE(i, n) {
        ds[i] = ss[i];
        dt[i] = st[i];
        dl[i] = sl[i];
    }
}
__global__ __launch_bounds__(BLOCK) void k_copy_rows(const uint32_t* __restrict__ counts, const uint32_t* __restrict__ srow,
                                                     const uint32_t* __restrict__ scol, uint32_t* __restrict__ drow,
                                                     uint32_t* __restrict__ dcol) {
    const uint32_t rows = counts[C_ROWS], np = counts[C_PAIRS];
    GRID_STRIDE(i, rows + 1u) drow[i] = srow[i];
    GRID_STRIDE(k, np) dcol[k] = scol[k];
}
// CSR rows -> (first, second) pairs, the reference's uvec2 neighbour list
// (list_definitions.h:19, neighbor_add.glsl:20-23): pair k of row i is
// (i, col[k]).
__global__ __launch_bounds__(BLOCK) void k_rows_to_pairs(const uint32_t* __restrict__ row,
                                                         const uint32_t* __restrict__ col,
                                                         const uint32_t* __restrict__ d_rows,
                                                         const uint32_t* __restrict__ d_np, uint32_t pair_cap,
                                                         uint2* __restrict__ pairs) {
    const uint32_t rows = *d_rows, np = min(*d_np, pair_cap);
    GRID_STRIDE(i, rows) {
        const uint32_t b = min(row[i], np), e = min(row[i + 1], np);
        for (uint32_t k = b; k < e; ++k) pairs[k] = make_uint2(i, col[k]);
    }
}
// Pairs grouped by first id (after a stable sort by it) -> CSR: counts per
// row, the second ids in pair order.
__global__ __launch_bounds__(BLOCK) void k_pair_firsts(const uint2* __restrict__ pairs, const uint32_t* __restrict__ d_np,
                                                       uint32_t* __restrict__ first, uint32_t* __restrict__ iota) {
    const uint32_t np = *d_np;
    GRID_STRIDE(k, np) {
        first[k] = pairs[k].x;
        iota[k] = k;
    }
}
__global__ __launch_bounds__(BLOCK) void k_pair_row_counts(const uint32_t* __restrict__ sorted_first,
                                                           const uint32_t* __restrict__ d_np, uint32_t rows_cap,
                                                           uint32_t* __restrict__ cnt) {
    const uint32_t np = *d_np;
    GRID_STRIDE(k, np) {
        const uint32_t f = sorted_first[k];
        if (f < rows_cap) atomicAdd(&cnt[f], 1u);
    }
}
__global__ __launch_bounds__(BLOCK) void k_pair_cols(const uint2* __restrict__ pairs, const uint32_t* __restrict__ order,
                                                     const uint32_t* __restrict__ d_np, uint32_t* __restrict__ col) {
    const uint32_t np = *d_np;
    GRID_STRIDE(k, np) col[k] = pairs[order[k]].y;
}

}  // namespace apbf

using namespace apbf;

struct apbf_context {
    Scene* sc = nullptr;
};

namespace {

// The context scene's own first-set buffers while the caller's are bound.
struct Stash {
    ParticleArrays p;
    uint32_t *ts, *tt;
    float* tl;
    uint32_t *row, *col;
    uint32_t pcap;
};

bool fits(const Scene* s, const apbf_particle_lists* P, const apbf_transfer_lists* T, const apbf_neighbor_lists* N) {
    if (!P || !P->d_length || P->capacity != s->cap || !P->position || !P->velocity || !P->inverse_mass ||
        !P->radius || !P->pos_backup || !P->transferring || !P->target_radius || !P->kernel_width ||
        !P->boundariness || !P->boundary_distance)
        return false;
    if (T && (!T->d_length || T->capacity != s->tcap || !T->source || !T->target || !T->time_left)) return false;
    if (N && (!N->d_length || !N->d_rows || N->capacity != s->pcap || !N->row_offset || !N->neighbors))
        return false;
    return true;
}

Stash bind(Scene* s, const apbf_particle_lists* P, const apbf_transfer_lists* T, const apbf_neighbor_lists* N) {
    Stash k;
    k.p = s->arrays(0);
    k.ts = s->t_source[0];
    k.tt = s->t_target[0];
    k.tl = s->t_time_left[0];
    k.row = s->row;
    k.col = s->col;
    k.pcap = s->pcap;
    s->cur = 0;
    s->tcur = 0;
    s->pos[0] = reinterpret_cast<int4*>(P->position);
    s->vel[0] = reinterpret_cast<float4*>(P->velocity);
    s->inv_mass[0] = P->inverse_mass;
    s->radius[0] = P->radius;
    s->backup[0] = reinterpret_cast<int4*>(P->pos_backup);
    s->transferring[0] = P->transferring;
    s->target_radius[0] = P->target_radius;
    s->kernel_width[0] = P->kernel_width;
    s->boundariness[0] = P->boundariness;
    s->boundary_distance[0] = P->boundary_distance;
    if (T) {
        s->t_source[0] = T->source;
        s->t_target[0] = T->target;
        s->t_time_left[0] = T->time_left;
    }
    if (N) {
        s->row = N->row_offset;
        s->col = N->neighbors;
    }
    s->pairs_valid = N != nullptr;
    hipLaunchKernelGGL(k_bind_counts, dim3(1), dim3(64), 0, s->stream, s->counts, P->d_length,
                       T ? T->d_length : nullptr, N ? N->d_length : nullptr, N ? N->d_rows : nullptr);
    return k;
}

// Results that a stage left in the second buffer set (or in the context's own
// neighbour arrays) go back into the caller's arrays; then the context's own
// buffers are put back.
void unbind(Scene* s, const Stash& k, const apbf_particle_lists* P, const apbf_transfer_lists* T,
            const apbf_neighbor_lists* N) {
    hipStream_t st = s->stream;
    if (s->cur == 1) {
        ParticleArrays dst = s->arrays(0);  // the caller's arrays (and the context's lambda history)
        hipLaunchKernelGGL(k_copy_particles, dim3(stride_grid(s->cap)), dim3(BLOCK), 0, st, s->counts, s->arrays(1),
                           dst);
    }
    if (T && s->tcur == 1)
        hipLaunchKernelGGL(k_copy_transfers, dim3(stride_grid(s->tcap)), dim3(BLOCK), 0, st, s->counts,
                           s->t_source[1], s->t_target[1], s->t_time_left[1], T->source, T->target, T->time_left);
    if (N && (s->row != N->row_offset || s->col != N->neighbors)) {
        // the spreading wrote its kept list into the context's arrays (pointer swaps)
        hipLaunchKernelGGL(k_copy_rows, dim3(stride_grid(s->cap + 1)), dim3(BLOCK), 0, st, s->counts, s->row, s->col,
                           N->row_offset, N->neighbors);
        if (s->row != N->row_offset) std::swap(s->row, s->u0);
        if (s->col != N->neighbors) std::swap(s->col, s->col_tmp);
    }
    hipLaunchKernelGGL(k_unbind_counts, dim3(1), dim3(64), 0, st, s->counts, P->d_length, T ? T->d_length : nullptr,
                       N ? N->d_length : nullptr, N ? N->d_rows : nullptr);
    s->cur = 0;
    s->tcur = 0;
    s->pos[0] = k.p.pos;
    s->vel[0] = k.p.vel;
    s->inv_mass[0] = k.p.inv_mass;
    s->radius[0] = k.p.radius;
    s->backup[0] = k.p.backup;
    s->transferring[0] = k.p.transferring;
    s->target_radius[0] = k.p.target_radius;
    s->kernel_width[0] = k.p.kernel_width;
    s->boundariness[0] = k.p.boundariness;
    s->boundary_distance[0] = k.p.boundary_distance;
    s->t_source[0] = k.ts;
    s->t_target[0] = k.tt;
    s->t_time_left[0] = k.tl;
    s->row = k.row;
    s->col = k.col;
    s->pcap = k.pcap;
}

template <class Stage>
apbf_status run(apbf_context* c, apbf_stream stream, const apbf_particle_lists* P, const apbf_transfer_lists* T,
                const apbf_neighbor_lists* N, Stage&& stage) {
    if (!c || !c->sc) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = c->sc;
    if (!fits(s, P, T, N)) return APBF_ERR_INVALID_ARGUMENT;
    s->stream = static_cast<hipStream_t>(stream);
    const Stash k = bind(s, P, T, N);
    stage(s);
    unbind(s, k, P, T, N);
    return hipGetLastError() == hipSuccess ? APBF_OK : APBF_ERR_HIP;
}

}  // namespace

extern "C" {

apbf_status apbf_context_create(const apbf_scene_params* params, apbf_context** out) {
    if (!params || !out) return APBF_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    void* scene = nullptr;
    apbf_status st = apbf_scene_create(params, &scene);
    if (st != APBF_OK) return st;
    apbf_context* c = new (std::nothrow) apbf_context();
    if (!c) {
        apbf_scene_destroy(scene);
        return APBF_ERR_INVALID_ARGUMENT;
    }
    c->sc = static_cast<Scene*>(scene);
    c->sc->exact_walks = true;  // the caller's pair capacity holds the list itself; no walk slots
    *out = c;
    return APBF_OK;
}

apbf_status apbf_context_destroy(apbf_context* c) {
    if (!c) return APBF_ERR_INVALID_ARGUMENT;
    apbf_status st = c->sc ? apbf_scene_destroy(c->sc) : APBF_OK;
    delete c;
    return st;
}

apbf_status apbf_op_velocity_handling(apbf_context* c, const apbf_particle_lists* p, float* last_delta_time,
                                      float delta_time, const float* acceleration3, apbf_stream stream) {
    if (!last_delta_time || !acceleration3) return APBF_ERR_INVALID_ARGUMENT;
    return run(c, stream, p, nullptr, nullptr, [&](Scene* s) {
        for (int d = 0; d < 3; ++d) s->P.gravity[d] = acceleration3[d];
        s->last_dt = *last_delta_time;
        s->velocity_handling(delta_time);
        *last_delta_time = s->last_dt;
    });
}

apbf_status apbf_op_particle_transfer(apbf_context* c, const apbf_particle_lists* p, const apbf_transfer_lists* t,
                                      float delta_time, apbf_stream stream) {
    if (!t) return APBF_ERR_INVALID_ARGUMENT;
    return run(c, stream, p, t, nullptr, [&](Scene* s) { s->particle_transfer(delta_time); });
}

apbf_status apbf_op_kernel_width_from_boundary_distance(apbf_context* c, const apbf_particle_lists* p,
                                                        apbf_stream stream) {
    return run(c, stream, p, nullptr, nullptr, [&](Scene* s) { s->kernel_width_from_bd(); });
}

apbf_status apbf_op_neighborhood(apbf_context* c, const apbf_particle_lists* p, const apbf_transfer_lists* t,
                                 const apbf_neighbor_lists* nb, int32_t method, float range_scale,
                                 const float* min3, const float* max3, uint32_t res_log2, apbf_stream stream) {
    if (!nb || method < 0 || method > 2 || !min3 || !max3 || res_log2 == 0) return APBF_ERR_INVALID_ARGUMENT;
    if (c && c->sc && !hash_res_ok(res_log2, c->sc->dims())) return APBF_ERR_INVALID_ARGUMENT;
    if (c && c->sc && (1u << (c->sc->dims() * res_log2)) != c->sc->n_cells) return APBF_ERR_CAPACITY;  // cell tables
    return run(c, stream, p, t, nb, [&](Scene* s) {
        for (int d = 0; d < 3; ++d) {
            s->grid.mn[d] = min3[d];
            s->grid.mx[d] = max3[d];
        }
        s->grid.res = res_log2;
        if (method == 1)
            s->neighborhood_bs(range_scale);
        else if (method == 2)
            s->neighborhood_bf(range_scale);
        else
            s->neighborhood_green(range_scale, false);
    });
}

apbf_status apbf_op_spread_kernel_width(apbf_context* c, const apbf_particle_lists* p, const apbf_neighbor_lists* nb,
                                        apbf_stream stream) {
    if (!nb) return APBF_ERR_INVALID_ARGUMENT;
    return run(c, stream, p, nullptr, nb, [&](Scene* s) { s->spread_kernel_width(); });
}

apbf_status apbf_op_box_collision(apbf_context* c, const apbf_particle_lists* p, uint32_t n_boxes,
                                  const float* box_min4, const float* box_max4, apbf_stream stream) {
    if (n_boxes > APBF_MAX_BOXES || (n_boxes && (!box_min4 || !box_max4))) return APBF_ERR_INVALID_ARGUMENT;
    return run(c, stream, p, nullptr, nullptr, [&](Scene* s) {
        BoxSet b = s->boxes;
        b.n = n_boxes;
        for (uint32_t i = 0; i < n_boxes; ++i)
            for (int d = 0; d < 3; ++d) {
                b.mn[i][d] = box_min4[4 * i + d];
                b.mx[i][d] = box_max4[4 * i + d];
            }
        launch_box(s->pos[s->cur], s->radius[s->cur], s->counts, s->cap, b, s->stream);
    });
}

apbf_status apbf_op_sphere_collision(apbf_context* c, const apbf_particle_lists* p, const float* center3,
                                     float radius, int32_t hollow, apbf_stream stream) {
    if (!center3) return APBF_ERR_INVALID_ARGUMENT;
    return run(c, stream, p, nullptr, nullptr, [&](Scene* s) {
        BoxSet b = s->boxes;
        b.sphere = 1u;
        b.hollow = hollow ? 1 : 0;
        for (int d = 0; d < 3; ++d) b.sc[d] = center3[d];
        b.sr = radius;
        launch_sphere(s->pos[s->cur], s->radius[s->cur], s->counts, s->cap, b, s->stream);
    });
}

apbf_status apbf_op_incompressibility(apbf_context* c, const apbf_particle_lists* p, const apbf_neighbor_lists* nb,
                                      apbf_stream stream) {
    if (!nb) return APBF_ERR_INVALID_ARGUMENT;
    return run(c, stream, p, nullptr, nb, [&](Scene* s) { s->incompressibility(); });
}

apbf_status apbf_op_update_transfers(apbf_context* c, const apbf_particle_lists* p, const apbf_transfer_lists* t,
                                     const apbf_neighbor_lists* nb, apbf_stream stream) {
    if (!t || !nb) return APBF_ERR_INVALID_ARGUMENT;
    return run(c, stream, p, t, nb, [&](Scene* s) { s->update_transfers(); });
}

apbf_status apbf_neighbor_pairs(const uint32_t* row_offset, const uint32_t* neighbors, const uint32_t* d_rows,
                                const uint32_t* d_n_pairs, uint32_t rows_capacity, uint32_t pair_capacity,
                                uint32_t* out_pairs, apbf_stream stream) {
    if (!row_offset || !neighbors || !d_rows || !d_n_pairs || !out_pairs) return APBF_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(k_rows_to_pairs, dim3(stride_grid(rows_capacity)), dim3(BLOCK), 0,
                       static_cast<hipStream_t>(stream), row_offset, neighbors, d_rows, d_n_pairs, pair_capacity,
                       reinterpret_cast<uint2*>(out_pairs));
    return hipGetLastError() == hipSuccess ? APBF_OK : APBF_ERR_HIP;
}

size_t apbf_neighbor_rows_from_pairs_workspace_size(uint32_t pair_capacity, uint32_t rows) {
    const size_t inner = std::max(sort_workspace_bytes(pair_capacity), scan_workspace_bytes(rows + 1));
    return (size_t)(pair_capacity + 1) * 4 * 4 + inner + 256;
}

apbf_status apbf_neighbor_rows_from_pairs(const uint32_t* pairs, const uint32_t* d_n_pairs, uint32_t pair_capacity,
                                          uint32_t rows, uint32_t* row_offset, uint32_t* neighbors, void* workspace,
                                          size_t workspace_bytes, apbf_stream stream) {
    if (!pairs || !d_n_pairs || !row_offset || !neighbors || !workspace) return APBF_ERR_INVALID_ARGUMENT;
    if (workspace_bytes < apbf_neighbor_rows_from_pairs_workspace_size(pair_capacity, rows)) return APBF_ERR_WORKSPACE;
    const uint32_t rows_capacity = rows;
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t* first = static_cast<uint32_t*>(workspace);
    uint32_t* iota = first + pair_capacity + 1;
    uint32_t* sfirst = iota + pair_capacity + 1;
    uint32_t* order = sfirst + pair_capacity + 1;
    void* inner = order + pair_capacity + 1;
    const uint2* pr = reinterpret_cast<const uint2*>(pairs);
    hipLaunchKernelGGL(k_pair_firsts, dim3(stride_grid(pair_capacity)), dim3(BLOCK), 0, s, pr, d_n_pairs, first, iota);
    // stable: pairs of one particle keep their order
    radix_sort_pairs(first, iota, d_n_pairs, pair_capacity, 0xFFFFFFFFu, sfirst, order, inner, s);
    fill_u32(row_offset, 0u, (uint64_t)rows_capacity + 1, s);
    hipLaunchKernelGGL(k_pair_row_counts, dim3(stride_grid(pair_capacity)), dim3(BLOCK), 0, s, sfirst, d_n_pairs,
                       rows_capacity, row_offset);
    // exclusive scan over rows + 1 entries: row_offset[rows] = pairs of the first `rows` particles
    scan_u32_n(row_offset, row_offset, rows_capacity + 1, false, nullptr, inner, s);
    hipLaunchKernelGGL(k_pair_cols, dim3(stride_grid(pair_capacity)), dim3(BLOCK), 0, s, pr, order, d_n_pairs,
                       neighbors);
    return hipGetLastError() == hipSuccess ? APBF_OK : APBF_ERR_HIP;
}

}  // extern "C"
