// Host side of a scene: device buffers, the subsystem apply() sequence of one
// substep, and the C ABI of include/apbf.h.
//
// The subsystem methods mirror the reference's L4 classes and keep their
// order and flags (source/pool.cpp:67-106):
//   velocity_handling::apply      -> Scene::velocity_handling
//   particle_transfer::apply      -> Scene::particle_transfer      (adaptive.hip)
//   kernel width from boundary d. -> Scene::kernel_width_from_bd   (adaptive.hip)
//   neighborhood_green::apply     -> Scene::neighborhood_green
//   spread_kernel_width::apply    -> Scene::spread_kernel_width    (adaptive.hip)
//   box_collision::apply          -> Scene::box_collision
//   incompressibility::apply      -> Scene::incompressibility
//   update_transfers::apply       -> Scene::update_transfers       (adaptive.hip)
// No device memory is allocated after construction and nothing is read back
// to the host inside a substep; all list lengths stay on the device.
#include "scene.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

namespace apbf {

#define HIP_OK(x)                                  \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) return APBF_ERR_HIP; \
    } while (0)

template <class T>
static apbf_status dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc((void**)p, n * sizeof(T)) != hipSuccess) return APBF_ERR_HIP;
    return APBF_OK;
}

ParticleArrays Scene::arrays(int set) {
    ParticleArrays a;
    a.pos = pos[set];
    a.vel = vel[set];
    a.inv_mass = inv_mass[set];
    a.radius = radius[set];
    a.backup = backup[set];
    a.transferring = transferring[set];
    a.target_radius = target_radius[set];
    a.kernel_width = kernel_width[set];
    a.boundariness = boundariness[set];
    a.boundary_distance = boundary_distance[set];
    a.lam_prev = lam_prev[set];
    a.prev_old = nullptr;
    a.sid = sid[set];
    return a;
}

apbf_status Scene::init(const apbf_scene_params& p) {
    P = p;
    if (P.dimensions == 0u) P.dimensions = 3u;
    if (P.dimensions != 2u && P.dimensions != 3u) return APBF_ERR_INVALID_ARGUMENT;
    if (p.cap_particles == 0 || !hash_res_ok(p.hash_res_log2, dims()) || p.n_boxes > APBF_MAX_BOXES)
        return APBF_ERR_INVALID_ARGUMENT;
    cap = p.cap_particles;
    tcap = std::max<uint32_t>(p.cap_transfers, 1);
    pcap = std::max<uint32_t>(p.cap_pairs, 1);
    n_cells = 1u << (dims() * p.hash_res_log2);
    grid.res = p.hash_res_log2;
    grid.dims = dims();
    for (int d = 0; d < 3; ++d) {
        grid.mn[d] = p.hash_min[d];
        grid.mx[d] = p.hash_max[d];
    }
    boxes.n = p.n_boxes;
    for (uint32_t i = 0; i < p.n_boxes; ++i)
        for (int d = 0; d < 3; ++d) {
            boxes.mn[i][d] = p.box_min[i][d];
            boxes.mx[i][d] = p.box_max[i][d];
        }
    no_boxes.n = 0;
    HIP_OK(hipGetDevice(&device));
    apbf_status st;
#define A(ptr, n)                               \
    if ((st = dalloc(&(ptr), (n))) != APBF_OK) \
        return st;
    for (int s = 0; s < 2; ++s) {
        A(pos[s], cap);
        A(vel[s], cap);
        A(inv_mass[s], cap);
        A(radius[s], cap);
        A(backup[s], cap);
        A(transferring[s], cap);
        A(target_radius[s], cap);
        A(kernel_width[s], cap);
        A(boundariness[s], cap);
        A(boundary_distance[s], cap);
        A(lam_prev[s], cap);
        A(t_source[s], tcap);
        A(t_target[s], tcap);
        A(t_time_left[s], tcap);
    }
    A(counts, C_COUNT);
    A(keys, cap);
    A(vals, cap);
    A(skeys, cap);
    A(perm, cap);
    A(cell_start, n_cells);
    A(cell_end, n_cells);
    A(cnt, cap + 1);
    A(row, cap + 1);
    A(col, pcap);
    A(col_tmp, pcap);
    A(gc, cap);
    A(long_list, cap);
    A(long_list2, cap);
    A(lambda, cap);
    A(shift, cap);
    A(ukw, cap);
    A(u0, cap + 1);
    A(u1, cap + 1);
    A(u2, cap + 1);
    A(u3, cap + 1);
    A(u4, cap + 1);
    A(small, 64);
    A(a_type, cap + 1);
    A(a_s, cap + 1);
    A(a_t, cap + 1);
    A(a_state, cap + 1);
    A(a_owner, cap + 1);
    A(a_owner2, cap + 1);
    A(a_flag, cap + 1);
    A(a_scan, cap + 1);
    A(a_list, cap + 1);
    A(a_t0, cap + 1);
    A(t_flag, tcap + 1);
    A(t_scan, tcap + 1);
    err_blocks = density_rows_blocks(cap);
    A(part_sum, 2 * err_blocks);
    A(part_max, err_blocks);
    A(err_stats, 6);
    A(iter_err, MAX_ITER_RECORD);
    A(feat, (size_t)cap * MLP_IN);
    A(lam_acc, cap);
    A(lam_hat, cap);
    A(particle_substeps, 5);
    A(tol_err, MAX_ITER_RECORD + 1);
    A(pos_save, cap);
    A(solve_stats, 4);
#undef A
    size_t ws = std::max(sort_workspace_bytes_any(cap), scan_workspace_bytes(cap + 1));
    ws = std::max(ws, scan_workspace_bytes(pcap));
    ws = std::max(ws, scan_workspace_bytes(tcap + 1));
    ws_bytes = ws;
    HIP_OK(hipMalloc(&ws_mem, ws_bytes));
    HIP_OK(hipMemset(counts, 0, C_COUNT * 4));
    HIP_OK(hipMemset(err_stats, 0, 6 * sizeof(double)));
    HIP_OK(hipMemset(iter_err, 0, MAX_ITER_RECORD * sizeof(double)));
    HIP_OK(hipMemset(particle_substeps, 0, 5 * sizeof(unsigned long long)));
    HIP_OK(hipMemset(solve_stats, 0, 4 * sizeof(unsigned long long)));
    HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&h_counts), C_COUNT * 4, hipHostMallocDefault));
    std::memset(h_counts, 0, C_COUNT * 4);
    HIP_OK(hipEventCreateWithFlags(&h_counts_ready, hipEventDisableTiming));
    HIP_OK(hipDeviceSynchronize());
    return APBF_OK;
}

Scene::~Scene() {
    scene_release_graph(this);
    for (int s = 0; s < 2; ++s) {
        hipFree(pos[s]); hipFree(vel[s]); hipFree(inv_mass[s]); hipFree(radius[s]);
        hipFree(backup[s]); hipFree(transferring[s]); hipFree(target_radius[s]);
        hipFree(kernel_width[s]); hipFree(boundariness[s]); hipFree(boundary_distance[s]); hipFree(lam_prev[s]);
        hipFree(t_source[s]); hipFree(t_target[s]); hipFree(t_time_left[s]);
    }
    void* ps_seg[] = {d_grids, d_boxes, seg_start, sid[0], sid[1]};
    for (void* q : ps_seg) hipFree(q);
    void* ps[] = {counts, keys, vals, skeys, perm, cell_start, cell_end, cnt, row, col, col_tmp, gc, long_list, long_list2,
                  lambda, shift, ukw, u0, u1, u2, u3, u4, small, part_sum,
                  part_max, ws_mem, a_type, a_s, a_t, a_state, a_owner, a_owner2, a_flag, a_scan,
                  a_list, a_t0, t_flag, t_scan, err_stats, iter_err, feat, lam_acc, lam_hat, particle_substeps,
                  tol_err, pos_save, solve_stats};
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    for (void* q : ps) hipFree(q);
    if (h_counts) (void)hipHostFree(h_counts);
    if (h_counts_ready) (void)hipEventDestroy(h_counts_ready);
}

// ------------------------------------------------------ velocity_handling
// velocity_handling.cpp:15-31
void Scene::velocity_handling(float dt) {
    if (dt == 0.0f) {
        launch_copy_int4(backup[cur], pos[cur], counts, cap, stream);
        return;
    }
    float acc[3] = {P.gravity[0] * dt, P.gravity[1] * dt, P.gravity[2] * dt};
    launch_integrate(pos[cur], backup[cur], vel[cur], counts, cap, last_dt, dt, acc, stream);
    last_dt = dt;
}

// A packed batch: the scenes' grids (each with its own resolution; the cell
// tables give every scene a block of 2^(3 max res) cells) and boxes.
apbf_status Scene::init_segments(uint32_t n, const apbf_segment_desc* segs) {
    if (n == 0 || !segs) return APBF_ERR_INVALID_ARGUMENT;
    uint32_t max_res = 0;
    std::vector<GridDesc> g(n);
    std::vector<BoxSet> b(n);
    bool any_box = false;
    for (uint32_t k = 0; k < n; ++k) {
        const apbf_segment_desc& d = segs[k];
        if (!hash_res_ok(d.hash_res_log2, dims()) || d.n_boxes > APBF_MAX_BOXES) return APBF_ERR_INVALID_ARGUMENT;
        max_res = std::max(max_res, d.hash_res_log2);
        g[k].res = d.hash_res_log2;
        g[k].dims = dims();
        std::memset(&b[k], 0, sizeof(BoxSet));
        b[k].n = d.n_boxes;
        any_box = any_box || d.n_boxes;
        for (int a = 0; a < 3; ++a) {
            g[k].mn[a] = d.hash_min[a];
            g[k].mx[a] = d.hash_max[a];
        }
        for (uint32_t i = 0; i < d.n_boxes; ++i)
            for (int a = 0; a < 3; ++a) {
                b[k].mn[i][a] = d.box_min[i][a];
                b[k].mx[i][a] = d.box_max[i][a];
            }
    }
    uint32_t bits = 0;
    while ((1u << bits) < n) ++bits;
    if (dims() * max_res + bits > 31u) return APBF_ERR_CAPACITY;  // scene index above the cell hash in a u32 key
    n_seg = n;
    seg_shift = dims() * max_res;
    n_cells = n << seg_shift;
    boxes.n = any_box ? 1u : 0u;  // the per-scene boxes are in the table; this only says "run the colliders"
    hipFree(cell_start);
    hipFree(cell_end);
    cell_start = cell_end = nullptr;
    apbf_status st;
    if ((st = dalloc(&cell_start, n_cells)) != APBF_OK || (st = dalloc(&cell_end, n_cells)) != APBF_OK ||
        (st = dalloc(&d_grids, n)) != APBF_OK || (st = dalloc(&d_boxes, n)) != APBF_OK ||
        (st = dalloc(&seg_start, n)) != APBF_OK || (st = dalloc(&sid[0], cap)) != APBF_OK ||
        (st = dalloc(&sid[1], cap)) != APBF_OK)
        return st;
    HIP_OK(hipMemcpy(d_grids, g.data(), n * sizeof(GridDesc), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_boxes, b.data(), n * sizeof(BoxSet), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(seg_start, 0, n * 4));
    HIP_OK(hipMemset(sid[0], 0, (size_t)cap * 4));
    HIP_OK(hipMemset(sid[1], 0, (size_t)cap * 4));
    return APBF_OK;
}

// ----------------------------------------------------- neighborhood_green
// neighborhood_green.cpp:27-77: hash every particle, stable sort by hash,
// permute the particle SoA, cell tables, pair search. In a packed batch the
// key holds the scene index above the hash, so the scenes stay apart and
// every scene's particles keep their own sorted order.
void Scene::neighborhood_green(float range_scale, bool spread) {
    hipStream_t s = stream;
    const SegmentTable seg0 = seg_table();
    uint32_t upper = packed() ? (n_seg << seg_shift) : (1u << (grid.res * dims()));
    launch_hash_iota(pos[cur], counts + C_N, cap, grid, keys, vals, s, &seg0);
    radix_sort_pairs(keys, vals, counts + C_N, cap, upper, skeys, perm, ws_mem, s, sort_shape);
    ParticleArrays src = arrays(cur), dst = arrays(1 - cur);
    bool tk = tbegin(K_PERMUTE);
    launch_permute(perm, counts, cap, src, dst, s);
    tend(tk);
    cur = 1 - cur;
    if (packed()) launch_segment_starts(sid[cur], counts, cap, seg_start, s);
    // The transfer list follows its particles (indexed_list owner rebuild).
    remap_transfers_after_permute();
    fill_u32(cell_start, 0u, n_cells, s);
    fill_u32(cell_end, 0u, n_cells, s);
    launch_ranges(skeys, counts + C_N, cap, cell_start, cell_end, s);
    build_pairs(range_scale, spread);
}

// The pair build. With `spread` the walk also carries the kernel-width
// spreading (spread_kernel_width.cpp:12-26, kernel_width.comp:32-60) and writes only its short list, so
// neighborhood_green followed by spread_kernel_width costs two walks.
void Scene::build_pairs(float range_scale, bool spread) {
    hipStream_t s = stream;
    const SegmentTable seg = seg_table();
    if (walk_auto) {
        // short rows: 4 lanes per particle fill a wave better (large pool: 12.6 ms per substep against 13.5
        // with 8 and 15.3 with whole waves); long rows: whole waves per particle (water drop: 6.79 against
        // 7.15 ms with 8)
        refresh_seen_counts();
        walk_lanes = (seen_n != 0u && (uint64_t)seen_pairs < 62ull * (uint64_t)seen_n) ? 4 : 64;
    }
    const int on_t = P.s.mBaseKernelWidthOnTargetRadius;
    uint32_t* row_raw = spread ? u0 : row;  // offsets of the full (unfiltered) list
    const uint32_t stride = exact_walks ? 0u : pcap / (cap ? cap : 1u);
    launch_set_n_plus_1(counts, s);
    if (count_candidates)  // diagnostic: the counting walk visits the cells and candidates of every walk
        launch_walk(0, spread, pos[cur], kernel_width[cur], radius[cur], target_radius[cur], on_t, cell_start,
                    cell_end, counts, cap, range_scale, grid, pcap, cnt, spread ? u1 : nullptr, nullptr, nullptr,
                    nullptr, nullptr, s, walk_lanes, 0, nullptr, &seg,
                    reinterpret_cast<unsigned long long*>(counts + C_CANDIDATES));
    if (spread) spread_begin();
    if (stride == 0) {  // no room for slots: the two exact walks
        bool tk = tbegin(K_WALK_COUNT);
        launch_walk(0, spread, pos[cur], kernel_width[cur], radius[cur], target_radius[cur], on_t, cell_start,
                    cell_end, counts, cap, range_scale, grid, pcap, cnt, u1, nullptr, nullptr, nullptr, nullptr, s,
                    walk_lanes, 0, nullptr, &seg);
        tend(tk);
        scan_u32(cnt, row_raw, counts + C_N_PLUS_1, cap + 1, false, nullptr, ws_mem, s);
        launch_pairs_total(counts, row_raw, pcap, s);
        if (spread) scan_u32(u1, row, counts + C_N_PLUS_1, cap + 1, false, nullptr, ws_mem, s);
        tk = tbegin(K_WALK_FILL);
        launch_walk(1, spread, pos[cur], kernel_width[cur], radius[cur], target_radius[cur], on_t, cell_start,
                    cell_end, counts, cap, range_scale, grid, pcap, nullptr, nullptr, row_raw, row, col, ukw, s,
                    walk_lanes, 0, nullptr, &seg);
        tend(tk);
    } else {
        // one walk into fixed-stride slots (col_tmp), counts of both lists and
        // the spreading; then either packing, or on overflow the exact walk
        bool tk = tbegin(K_WALK_FILL);
        launch_walk(2, spread, pos[cur], kernel_width[cur], radius[cur], target_radius[cur], on_t, cell_start,
                    cell_end, counts, cap, range_scale, grid, pcap, cnt, spread ? u1 : nullptr, nullptr, nullptr,
                    col_tmp, ukw, s, walk_lanes, stride, nullptr, &seg);
        tend(tk);
        scan_u32(cnt, row_raw, counts + C_N_PLUS_1, cap + 1, false, nullptr, ws_mem, s);
        launch_pairs_total(counts, row_raw, pcap, s);
        if (spread) scan_u32(u1, row, counts + C_N_PLUS_1, cap + 1, false, nullptr, ws_mem, s);
        launch_walk_decide(counts, pcap, s);
        launch_pack_rows(counts, cap, row, col_tmp, pcap, col, s);
        // exact fill: every row when the list is clamped (after restarting the
        // spreading), the rows longer than their slot otherwise, else nothing
        if (spread) spread_reset_if_fallback();
        launch_walk(1, spread, pos[cur], kernel_width[cur], radius[cur], target_radius[cur], on_t, cell_start,
                    cell_end, counts, cap, range_scale, grid, pcap, nullptr, nullptr, row_raw, row, col, ukw, s,
                    walk_lanes, stride, counts, &seg);
    }
    if (spread) spread_end();
    pairs_valid = true;
}

// neighborhood_brute_force.cpp:19-42: every particle against every other, in
// the current particle order (no sort, no permutation).
void Scene::neighborhood_bf(float range_scale) {
    hipStream_t s = stream;
    launch_set_n_plus_1(counts, s);
    bool tk = tbegin(K_WALK_COUNT);
    launch_brute(false, pos[cur], kernel_width[cur], counts, cap, range_scale, pcap, cnt, nullptr, nullptr, s);
    tend(tk);
    scan_u32(cnt, row, counts + C_N_PLUS_1, cap + 1, false, nullptr, ws_mem, s);
    launch_pairs_total(counts, row, pcap, s);
    tk = tbegin(K_WALK_FILL);
    launch_brute(true, pos[cur], kernel_width[cur], counts, cap, range_scale, pcap, nullptr, row, col, s);
    tend(tk);
    pairs_valid = true;
}

// --------------------------------------------------------- box_collision
void Scene::box_collision() {
    const SegmentTable seg = seg_table();
    launch_box(pos[cur], radius[cur], counts, cap, boxes, stream, &seg);
}
// sphere_collision::apply (spherical_pool.cpp:86-89 runs it in place of the
// box collision)
void Scene::sphere_collision() {
    if (boxes.sphere) launch_sphere(pos[cur], radius[cur], counts, cap, boxes, stream);
}
// The colliders that open a solver iteration: the boxes (when the scene has
// any), then the sphere (when set).
void Scene::colliders() {
    if (boxes.n) box_collision();
    sphere_collision();
}

// ----------------------------------------------------- incompressibility
// incompressibility.cpp:12-45: one k_solve_rows pass (sums, lambda,
// boundariness, shifts into the accumulator) and one k_solve_apply pass per
// iteration. `next_boxes` fuses the box collision that opens the next solver
// iteration into the apply pass.
SolverArgs Scene::solver_args() const {
    SolverArgs a;
    a.hid = P.s.mHeightKernelId;
    a.gid = P.s.mGradientKernelId;
    a.method = P.s.mBoundarinessCalculationMethod;
    a.update_b = P.s.mUpdateBoundariness;
    a.b_speed = P.s.mBoundarinessAdaptionSpeed;
    a.b_sgl_factor = P.s.mBoundarinessSelfGradLengthFactor;
    a.b_under_factor = P.s.mBoundarinessUnderpressureFactor;
    a.lambda_scale_base = pow_f(2.0f * P.s.mSmallestTargetRadius, dim());
    a.dim = dim();
    a.it_err = nullptr;
    a.lam_override = nullptr;
    a.lam_acc = nullptr;
    a.tol_sum = nullptr;
    a.skip = nullptr;
    a.split_len = 0;
    a.list = nullptr;
    a.list_slot = C_LONG;
    return a;
}

// The counts of the last substep whose copy to the host has completed (the
// copy is stream-ordered and never waited for; until one completes, the
// previous values stay). Inside stream capture no copy is recorded, so a
// captured graph keeps the launch choices of the state it was captured from.
void Scene::refresh_seen_counts() {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    if (hipEventQuery(h_counts_ready) != hipSuccess) return;  // still in flight (or never recorded)
    seen_n = h_counts[C_N];
    seen_pairs = h_counts[C_PAIRS];
}

void Scene::prepare_solver() {
    refresh_seen_counts();
    const bool short_rows = seen_n != 0u && (uint64_t)seen_pairs < 62ull * (uint64_t)seen_n;
    if (long_rows_mode >= 0) {
        long_rows = long_rows_mode == 1;
    } else {
        long_rows = (uint64_t)seen_pairs > 62ull * (uint64_t)seen_n;
    }
    // short rows: 8 lanes per row (large pool, 26 entries: 1.58 against 2.14 ms per launch with 16);
    // long rows: 16 (water drop: 0.66 against 1.01 ms with 8 and 0.80 with 32)
    if (solve_auto) solve_lanes = short_rows ? 8 : 16;
    if (P.s.mHeightKernelId == 1 && P.s.mGradientKernelId == 1)
        launch_long_rows(counts, cap, row, pcap, solve_lanes, long_list, long_list2, stream);
    launch_kernel_consts(kernel_width[cur], radius[cur], inv_mass[cur], counts, cap, P.s.mHeightKernelId,
                         P.s.mGradientKernelId, dim(), gc, stream);
    fill_u32_dev(reinterpret_cast<uint32_t*>(shift), 0u, counts + C_N4, 4 * cap, stream);
}

// `decided`: the iteration's row pass accumulates its density error into
// tol_err, a decision kernel picks the apply pass's action (tolerance mode and
// the warm-start fallback), and a finished solve skips the pass.
void Scene::incompressibility_iteration(const BoxSet& next_boxes, int it, const float* lam_override, bool decided) {
    hipStream_t s = stream;
    SolverArgs a = solver_args();
    if (measure_error && it >= 0 && it < MAX_ITER_RECORD) a.it_err = iter_err + it;
    if (ws_mode != 0 && it >= 0) a.lam_acc = lam_acc;
    a.lam_override = lam_override;
    if (lam_override) a.lam_acc = lam_acc;
    if (decided) {
        a.tol_sum = tol_err + (lam_override ? 0 : 1 + it);
        if (!lam_override) a.skip = counts + C_SOLVE_DONE;
    }
    bool tk = tbegin(K_SOLVE_ROWS);
    launch_solve_rows(pos[cur], inv_mass[cur], radius[cur], kernel_width[cur], gc, counts, cap, row, col, pcap, a,
                      boundariness[cur], lambda, shift, s, solve_lanes, long_rows, long_list, long_list2);
    tend(tk);
    const bool judge = decided && !lam_override;
    if (judge) launch_solve_decide(counts, tol_err, (uint32_t)it, tol, ws_mode == 2 && ws_model && it == 0, s);
    tk = tbegin(K_SOLVE_APPLY);
    const SegmentTable seg = seg_table();
    launch_solve_apply(pos[cur], shift, radius[cur], counts, cap, next_boxes, s,
                       judge ? counts + C_SOLVE_ACTION : nullptr, judge ? pos_save : nullptr, &seg);
    tend(tk);
}

void Scene::incompressibility() {
    prepare_solver();
    incompressibility_iteration(no_boxes, -1);
}

// The solver loop of pool.cpp:92-95 (box collision + incompressibility per
// iteration). Extensions without a counterpart in the reference:
//  - tolerance mode (tol > 0): up to tol_max_iters iterations; iteration k's
//    row pass measures the mean density error e_k of the positions it starts
//    from, and when e_k <= tol the iteration applies nothing and the solve ends;
//  - learned warm start (mode 2): the features of the starting state go
//    through the network and one shift pass with the predicted lambdas runs
//    before the iterations; the plain solver judges it: when the error after
//    the warm shift (iteration 0's e_0) exceeds the error before it, the
//    positions from before the warm shift are put back and the plain
//    iterations continue from there (cold start).
// Every launch is fixed, so a substep stays graph-capturable; decisions are
// device words.
void Scene::solver_loop() {
    const bool tolerance = tol > 0.0f;
    const int iters = tolerance ? tol_max_iters : P.solver_iterations;
    if (iters <= 0) return;
    const bool warm = ws_mode == 2 && ws_model;
    const bool decided = tolerance || warm;
    prepare_solver();
    if (decided) launch_solve_begin(counts, tol_err, MAX_ITER_RECORD + 1, stream);
    colliders();
    if (ws_mode != 0) {
        // learned warm start: features of the state the solver starts from, then
        // (mode 2) one shift pass with the predicted lambdas before the iterations
        fill_u32_dev(reinterpret_cast<uint32_t*>(lam_acc), 0u, counts + C_N, cap, stream);
        launch_features_rows(pos[cur], vel[cur], inv_mass[cur], radius[cur], kernel_width[cur], gc,
                             boundariness[cur], lam_prev[cur], counts, cap, row, col, pcap, solver_args(), last_dt,
                             feat, stream);
        if (warm) {
            launch_mlp_infer(ws_model->params, ws_model->norm, feat, counts + C_N, cap, lam_hat, stream);
            copy_u32_dev(reinterpret_cast<uint32_t*>(pos_save), reinterpret_cast<const uint32_t*>(pos[cur]),
                         counts + C_N4, 4 * cap, stream);
            incompressibility_iteration(boxes, -1, lam_hat, true);
        }
    }
    for (int it = 0; it < iters; ++it) {
        bool last = it == iters - 1;
        incompressibility_iteration(last ? no_boxes : boxes, it, nullptr, decided);
    }
    if (ws_mode != 0)
        copy_u32_dev(reinterpret_cast<uint32_t*>(lam_prev[cur]), reinterpret_cast<const uint32_t*>(lam_acc),
                     counts + C_N, cap, stream);
    launch_solve_stats(counts, solve_stats, decided ? 0u : (uint32_t)iters, stream);
}

// ----------------------------------------------------------------- update
void Scene::update(float dt) {
    const apbf_settings& s = P.s;
    velocity_handling(dt);
    if (!P.basic_pbf && (s.mMerge || s.mSplit)) particle_transfer(dt);
    if (!P.basic_pbf && s.mBaseKernelWidthOnBoundaryDistance) kernel_width_from_bd();
    const bool spread = !P.basic_pbf && !s.mBaseKernelWidthOnBoundaryDistance;
    const float scale = range_scale_override >= 0.0f ? range_scale_override : (spread ? 1.5f : 1.0f);
    if (nb_method == 1 || nb_method == 2) {
        if (nb_method == 1)
            neighborhood_bs(scale);
        else
            neighborhood_bf(scale);
        if (spread) spread_kernel_width();
    } else {
        neighborhood_green(scale, spread);  // spreading fused into the walk
    }
    solver_loop();
    launch_count_particle_substeps(counts, particle_substeps, stream);
    // the count block to the host for the launch choices of later substeps;
    // not recorded into captured graphs
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone) {
        refresh_seen_counts();  // read a completed copy before the next one is queued
        if (hipMemcpyAsync(h_counts, counts, C_COUNT * 4, hipMemcpyDeviceToHost, stream) == hipSuccess)
            (void)hipEventRecord(h_counts_ready, stream);
    }
    if (measure_error && !P.basic_pbf) {
        update_transfers(true);  // the error pass rides on the split/merge find
    } else {
        if (measure_error) accumulate_density_error();
        if (!P.basic_pbf) update_transfers();
    }
}

void Scene::accumulate_density_error() {  // after the solver loop of update()
    hipStream_t s = stream;
    if (P.solver_iterations <= 0)  // otherwise the solver loop computed gc from the same state
        launch_kernel_consts(kernel_width[cur], radius[cur], inv_mass[cur], counts, cap, P.s.mHeightKernelId,
                             P.s.mGradientKernelId, dim(), gc, s);
    launch_density_rows(pos[cur], inv_mass[cur], radius[cur], kernel_width[cur], gc, counts, cap, row, col, pcap,
                        P.s.mHeightKernelId, P.s.mGradientKernelId, dim(), part_sum, part_max, s);
    launch_density_error_finish(part_sum, part_max, err_blocks, counts, err_stats, s);
}

apbf_status Scene::stage(int st, float dt) {
    const apbf_settings& s = P.s;
    switch (st) {
        case APBF_STAGE_VELOCITY: velocity_handling(dt); break;
        case APBF_STAGE_PARTICLE_TRANSFER: particle_transfer(dt); break;
        case APBF_STAGE_KERNEL_WIDTH_BD: kernel_width_from_bd(); break;
        case APBF_STAGE_NEIGHBORHOOD: {
            const float scale = range_scale_override >= 0.0f ? range_scale_override
                                : (P.basic_pbf || s.mBaseKernelWidthOnBoundaryDistance) ? 1.0f
                                                                                        : 1.5f;
            if (nb_method == 1)
                neighborhood_bs(scale);
            else if (nb_method == 2)
                neighborhood_bf(scale);
            else
                neighborhood_green(scale, false);
            break;
        }
        case APBF_STAGE_SPREAD: spread_kernel_width(); break;
        case APBF_STAGE_BOX: box_collision(); break;
        case APBF_STAGE_SPHERE: sphere_collision(); break;
        case APBF_STAGE_INCOMPRESSIBILITY: incompressibility(); break;
        case APBF_STAGE_UPDATE_TRANSFERS: update_transfers(); break;
        default: return APBF_ERR_INVALID_ARGUMENT;
    }
    return hipGetLastError() == hipSuccess ? APBF_OK : APBF_ERR_HIP;
}

apbf_status Scene::density_error(double* mean_err, double* max_err) {
    if (!pairs_valid) return APBF_ERR_INVALID_ARGUMENT;
    hipStream_t s = stream;
    launch_kernel_consts(kernel_width[cur], radius[cur], inv_mass[cur], counts, cap, P.s.mHeightKernelId,
                         P.s.mGradientKernelId, dim(), gc, s);
    launch_density_rows(pos[cur], inv_mass[cur], radius[cur], kernel_width[cur], gc, counts, cap, row, col, pcap,
                        P.s.mHeightKernelId, P.s.mGradientKernelId, dim(), part_sum, part_max, s);
    std::vector<double> hs(err_blocks);
    std::vector<float> hm(err_blocks);
    uint32_t n = 0;
    HIP_OK(hipMemcpyAsync(hs.data(), part_sum, err_blocks * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(hm.data(), part_max, err_blocks * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(&n, counts + C_N, 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    double sum = 0.0;
    float mx = 0.0f;
    for (uint32_t b = 0; b < err_blocks; ++b) {
        sum += hs[b];
        mx = fmax_g(mx, hm[b]);
    }
    *mean_err = n ? sum / (double)n : 0.0;
    *max_err = (double)mx;
    return APBF_OK;
}

}  // namespace apbf

// ======================================================================= C ABI
using apbf::Scene;

extern "C" {

apbf_status apbf_scene_create(const apbf_scene_params* params, void** out_scene) {
    if (!params || !out_scene) return APBF_ERR_INVALID_ARGUMENT;
    *out_scene = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return APBF_ERR_NO_DEVICE;
    Scene* s = new (std::nothrow) Scene();
    if (!s) return APBF_ERR_INVALID_ARGUMENT;
    apbf_status st = s->init(*params);
    if (st != APBF_OK) {
        delete s;
        return st;
    }
    *out_scene = s;
    return APBF_OK;
}

apbf_status apbf_scene_count_walk_candidates(void* scene, int32_t enable) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    s->count_candidates = enable != 0;
    apbf::fill_u32(s->counts + apbf::C_CANDIDATES, 0u, 2, s->stream);
    return hipGetLastError() == hipSuccess ? APBF_OK : APBF_ERR_HIP;
}

apbf_status apbf_scene_create_packed(const apbf_scene_params* params, uint32_t n_segments,
                                     const apbf_segment_desc* segments, void** out_scene) {
    if (!params || !segments || !out_scene || n_segments == 0) return APBF_ERR_INVALID_ARGUMENT;
    *out_scene = nullptr;
    apbf_scene_params p = *params;
    p.hash_res_log2 = segments[0].hash_res_log2 ? segments[0].hash_res_log2 : 1u;  // replaced by the segments
    p.n_boxes = 0;
    void* h = nullptr;
    apbf_status st = apbf_scene_create(&p, &h);
    if (st != APBF_OK) return st;
    Scene* s = static_cast<Scene*>(h);
    st = s->init_segments(n_segments, segments);
    if (st == APBF_OK) st = hipDeviceSynchronize() == hipSuccess ? APBF_OK : APBF_ERR_HIP;
    if (st != APBF_OK) {
        apbf_scene_destroy(h);
        return st;
    }
    *out_scene = h;
    return APBF_OK;
}

apbf_status apbf_scene_set_segment_ids(void* scene, const uint32_t* host_ids) {
    if (!scene || !host_ids) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    if (!s->packed()) return APBF_ERR_INVALID_ARGUMENT;
    uint32_t n = 0;
    if (hipMemcpyAsync(&n, s->counts + apbf::C_N, 4, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    for (uint32_t i = 0; i < n; ++i)
        if (host_ids[i] >= s->n_seg) return APBF_ERR_INVALID_ARGUMENT;
    if (n && hipMemcpy(s->sid[s->cur], host_ids, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess)
        return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_scene_get_segment_ids(void* scene, uint32_t* host_ids) {
    if (!scene || !host_ids) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    if (!s->packed()) return APBF_ERR_INVALID_ARGUMENT;
    uint32_t n = 0;
    if (hipMemcpyAsync(&n, s->counts + apbf::C_N, 4, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    if (n && hipMemcpy(host_ids, s->sid[s->cur], (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_scene_destroy(void* scene) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    hipDeviceSynchronize();
    delete static_cast<Scene*>(scene);
    return APBF_OK;
}

apbf_status apbf_scene_set_stream(void* scene, apbf_stream stream) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    return apbf::scene_switch_stream(static_cast<Scene*>(scene), static_cast<hipStream_t>(stream));
}

apbf_status apbf_scene_set_state(void* scene, uint32_t n, const int32_t* h_pos, const float* h_vel,
                                 const float* h_im, const float* h_r, const int32_t* h_backup,
                                 const uint32_t* h_tr, const uint32_t* h_fidx, const float* h_trad,
                                 const float* h_kw, const float* h_b, const uint32_t* h_bd,
                                 uint32_t nt, const uint32_t* h_ts, const uint32_t* h_tt,
                                 const float* h_ttl, float last_delta_time) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    return s->set_state(n, h_pos, h_vel, h_im, h_r, h_backup, h_tr, h_fidx, h_trad, h_kw, h_b, h_bd, nt,
                        h_ts, h_tt, h_ttl, last_delta_time);
}

apbf_status apbf_scene_get_counts(void* scene, uint32_t* n, uint32_t* nt, uint32_t* np) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    uint32_t c[apbf::C_COUNT];
    if (hipMemcpyAsync(c, s->counts, sizeof(c), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    if (n) *n = c[apbf::C_N];
    if (nt) *nt = c[apbf::C_TRANSFERS];
    if (np) *np = c[apbf::C_PAIRS];
    return APBF_OK;
}

apbf_status apbf_scene_get_state(void* scene, int32_t* h_pos, float* h_vel, float* h_im, float* h_r,
                                 int32_t* h_backup, uint32_t* h_tr, float* h_trad, float* h_kw,
                                 float* h_b, uint32_t* h_bd, uint32_t* h_ts, uint32_t* h_tt,
                                 float* h_ttl, uint32_t* h_pairs, float* h_ldt) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    return static_cast<Scene*>(scene)->get_state(h_pos, h_vel, h_im, h_r, h_backup, h_tr, h_trad, h_kw,
                                                 h_b, h_bd, h_ts, h_tt, h_ttl, h_pairs, h_ldt);
}

apbf_status apbf_scene_get_lambda(void* scene, float* h_lambda) {
    if (!scene || !h_lambda) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    uint32_t n = 0;
    if (hipMemcpyAsync(&n, s->counts + apbf::C_N, 4, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    if (hipMemcpy(h_lambda, s->lambda, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_scene_stage(void* scene, int32_t stage, float dt) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    return static_cast<Scene*>(scene)->stage(stage, dt);
}

apbf_status apbf_scene_update(void* scene, float dt) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    s->update(dt);
    return hipGetLastError() == hipSuccess ? APBF_OK : APBF_ERR_HIP;
}

apbf_status apbf_scene_density_error(void* scene, double* mean_err, double* max_err) {
    if (!scene || !mean_err || !max_err) return APBF_ERR_INVALID_ARGUMENT;
    return static_cast<Scene*>(scene)->density_error(mean_err, max_err);
}

apbf_status apbf_scene_overflow_flags(void* scene, uint32_t* flags) {
    if (!scene || !flags) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    if (hipMemcpyAsync(flags, s->counts + apbf::C_OVERFLOW, 4, hipMemcpyDeviceToHost, s->stream) !=
            hipSuccess ||
        hipMemsetAsync(s->counts + apbf::C_OVERFLOW, 0, 4, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_scene_set_error_measurement(void* scene, int32_t enable) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    static_cast<Scene*>(scene)->measure_error = enable != 0;
    return APBF_OK;
}

apbf_status apbf_scene_error_stats(void* scene, double* mean_of_means, double* max_err, double* last_mean,
                                   double* mean_abs_of_means, uint32_t* substeps, int32_t reset) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    double st[6];
    uint32_t n = 0;
    if (hipMemcpyAsync(st, s->err_stats, sizeof(st), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipMemcpyAsync(&n, s->counts + apbf::C_ERR_SUBSTEPS, 4, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    if (mean_of_means) *mean_of_means = n ? st[0] / (double)n : 0.0;
    if (max_err) *max_err = st[1];
    if (last_mean) *last_mean = st[2];
    if (mean_abs_of_means) *mean_abs_of_means = n ? st[4] / (double)n : 0.0;
    if (substeps) *substeps = n;
    if (reset) {
        if (hipMemsetAsync(s->err_stats, 0, sizeof(st), s->stream) != hipSuccess ||
            hipMemsetAsync(s->counts + apbf::C_ERR_SUBSTEPS, 0, 4, s->stream) != hipSuccess ||
            hipStreamSynchronize(s->stream) != hipSuccess)
            return APBF_ERR_HIP;
    }
    return APBF_OK;
}

apbf_status apbf_scene_set_neighborhood_method(void* scene, int32_t method) {
    if (!scene || method < 0 || method > 2) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    if (s->packed() && method != 0) return APBF_ERR_INVALID_ARGUMENT;  // packed batches: the hash grid only
    s->nb_method = method;
    scene_drop_graph(s);
    return APBF_OK;
}

apbf_status apbf_scene_set_sphere(void* scene, const float* center3, float radius, int32_t hollow, int32_t enable) {
    if (!scene || (enable && !center3)) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    if (s->packed() && enable) return APBF_ERR_INVALID_ARGUMENT;  // packed batches: boxes only
    s->boxes.sphere = enable ? 1u : 0u;
    s->boxes.hollow = hollow ? 1 : 0;
    for (int d = 0; d < 3; ++d) s->boxes.sc[d] = enable ? center3[d] : 0.0f;
    s->boxes.sr = enable ? radius : 0.0f;
    scene_drop_graph(s);
    return APBF_OK;
}

apbf_status apbf_scene_set_gravity(void* scene, const float* gravity3) {
    if (!scene || !gravity3) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    for (int d = 0; d < 3; ++d) s->P.gravity[d] = gravity3[d];
    scene_drop_graph(s);
    return APBF_OK;
}

apbf_status apbf_scene_set_range_scale(void* scene, float scale) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    s->range_scale_override = scale;
    scene_drop_graph(s);
    return APBF_OK;
}

apbf_status apbf_scene_set_position_range(void* scene, const float* min3, const float* max3, uint32_t res_log2) {
    if (!scene || !min3 || !max3) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    if (!apbf::hash_res_ok(res_log2, s->dims())) return APBF_ERR_INVALID_ARGUMENT;
    if (s->packed()) return APBF_ERR_INVALID_ARGUMENT;  // packed batches: the grids of their segments
    const uint32_t cells = 1u << (s->dims() * res_log2);
    if (cells != s->n_cells) {  // new cell tables (outside any substep)
        if (hipStreamSynchronize(s->stream) != hipSuccess) return APBF_ERR_HIP;
        uint32_t *a = nullptr, *b = nullptr;
        if (hipMalloc(&a, (size_t)cells * 4) != hipSuccess || hipMalloc(&b, (size_t)cells * 4) != hipSuccess) {
            if (a) (void)hipFree(a);
            return APBF_ERR_HIP;
        }
        (void)hipFree(s->cell_start);
        (void)hipFree(s->cell_end);
        s->cell_start = a;
        s->cell_end = b;
        s->n_cells = cells;
    }
    for (int d = 0; d < 3; ++d) {
        s->grid.mn[d] = s->P.hash_min[d] = min3[d];
        s->grid.mx[d] = s->P.hash_max[d] = max3[d];
    }
    s->grid.res = s->P.hash_res_log2 = res_log2;
    s->pairs_valid = false;
    scene_drop_graph(s);
    return APBF_OK;
}

apbf_status apbf_scene_set_warm_start(void* scene, apbf_mlp* model, int32_t mode) {
    if (!scene || mode < 0 || mode > 2 || (mode == 2 && !model)) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    if (s->ws_mode == 0 && mode != 0) {  // the lambda history starts empty
        if (hipMemsetAsync(s->lam_prev[s->cur], 0, (size_t)s->cap * 4, s->stream) != hipSuccess) return APBF_ERR_HIP;
    }
    s->ws_mode = mode;
    s->ws_model = model;
    return APBF_OK;
}

apbf_status apbf_scene_harvest(void* scene, float* features, float* targets, uint32_t capacity, uint32_t* count) {
    if (!scene || !count) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    uint32_t n = 0;
    if (hipMemcpyAsync(&n, s->counts + apbf::C_N, 4, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    *count = n;
    if (s->ws_mode == 0) return APBF_ERR_INVALID_ARGUMENT;
    if (n > capacity) return APBF_ERR_CAPACITY;
    if ((features && hipMemcpyAsync(features, s->feat, (size_t)n * apbf::MLP_IN * 4, hipMemcpyDeviceToDevice,
                                    s->stream) != hipSuccess) ||
        (targets && hipMemcpyAsync(targets, s->lam_acc, (size_t)n * 4, hipMemcpyDeviceToDevice, s->stream) !=
                        hipSuccess) ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_scene_iteration_errors(void* scene, double* out, uint32_t max_iters, int32_t reset) {
    if (!scene || (!out && max_iters)) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    double e[apbf::MAX_ITER_RECORD];
    uint32_t n = 0;
    if (hipMemcpyAsync(e, s->iter_err, sizeof(e), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipMemcpyAsync(&n, s->counts + apbf::C_ERR_SUBSTEPS, 4, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    for (uint32_t k = 0; k < max_iters; ++k) out[k] = (k < apbf::MAX_ITER_RECORD && n) ? e[k] / (double)n : 0.0;
    if (reset && (hipMemsetAsync(s->iter_err, 0, sizeof(e), s->stream) != hipSuccess ||
                  hipStreamSynchronize(s->stream) != hipSuccess))
        return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_scene_get_warm_lambdas(void* scene, float* host_lambda) {
    if (!scene || !host_lambda) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    uint32_t n = 0;
    if (hipMemcpyAsync(&n, s->counts + apbf::C_ROWS, 4, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    if (n && hipMemcpy(host_lambda, s->lam_hat, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_scene_set_solver_tolerance(void* scene, float tolerance, int32_t max_iterations) {
    if (!scene || (tolerance > 0.0f && (max_iterations <= 0 || max_iterations > apbf::MAX_ITER_RECORD)))
        return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    s->tol = tolerance > 0.0f ? tolerance : 0.0f;
    s->tol_max_iters = tolerance > 0.0f ? max_iterations : 0;
    apbf::scene_drop_graph(s);
    return APBF_OK;
}

apbf_status apbf_scene_solver_stats(void* scene, uint64_t* iterations, uint64_t* substeps, uint64_t* fallbacks,
                                    int32_t reset) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    unsigned long long v[4] = {0, 0, 0, 0};
    if (hipMemcpyAsync(v, s->solve_stats, sizeof(v), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    if (iterations) *iterations = v[0];
    if (substeps) *substeps = v[1];
    if (fallbacks) *fallbacks = v[2];
    if (reset && (hipMemsetAsync(s->solve_stats, 0, sizeof(v), s->stream) != hipSuccess ||
                  hipStreamSynchronize(s->stream) != hipSuccess))
        return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_scene_particle_substeps(void* scene, uint64_t* total, uint64_t* pairs_total, int32_t reset) {
    if (!scene || !total) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    unsigned long long v[2] = {0, 0};
    if (hipMemcpyAsync(v, s->particle_substeps, 16, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    *total = v[0];
    if (pairs_total) *pairs_total = v[1];
    if (reset && (hipMemsetAsync(s->particle_substeps, 0, 5 * sizeof(unsigned long long), s->stream) != hipSuccess ||
                  hipStreamSynchronize(s->stream) != hipSuccess))
        return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_scene_time_kernels(void* scene, uint32_t kernel_mask, uint32_t max_launches) {
    if (!scene || kernel_mask >= (1u << apbf::K_COUNT_IDS)) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return APBF_ERR_HIP;
    for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
    s->events.clear();
    s->event_ids.clear();
    s->events_used = 0;
    s->timed_mask = kernel_mask;
    if (!kernel_mask) return APBF_OK;
    s->events.resize((size_t)max_launches * 2);
    s->event_ids.resize(max_launches);
    for (auto& e : s->events)
        if (hipEventCreate(&e) != hipSuccess) return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_scene_time_kernel(void* scene, int32_t kernel_id, uint32_t max_launches) {
    if (!scene || kernel_id >= apbf::K_COUNT_IDS) return APBF_ERR_INVALID_ARGUMENT;
    return apbf_scene_time_kernels(scene, kernel_id < 0 ? 0u : (1u << kernel_id), max_launches);
}

apbf_status apbf_scene_kernel_time_of(void* scene, int32_t kernel_id, double* total_ms, uint32_t* launches) {
    if (!scene || !total_ms || !launches) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return APBF_ERR_HIP;
    double t = 0.0;
    uint32_t n = 0;
    for (size_t i = 0; i + 1 < s->events_used; i += 2) {
        if (kernel_id >= 0 && s->event_ids[i / 2] != kernel_id) continue;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, s->events[i], s->events[i + 1]) != hipSuccess) return APBF_ERR_HIP;
        t += ms;
        ++n;
    }
    *total_ms = t;
    *launches = n;
    return APBF_OK;
}

apbf_status apbf_scene_kernel_time(void* scene, double* total_ms, uint32_t* launches) {
    apbf_status st = apbf_scene_kernel_time_of(scene, -1, total_ms, launches);
    if (st == APBF_OK) static_cast<Scene*>(scene)->events_used = 0;
    return st;
}

apbf_status apbf_scene_substep_totals(void* scene, uint64_t* out, uint32_t n_out, int32_t reset) {
    if (!scene || (!out && n_out)) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    unsigned long long v[5] = {0, 0, 0, 0, 0};
    if (hipMemcpyAsync(v, s->particle_substeps, sizeof(v), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    for (uint32_t k = 0; k < n_out; ++k) out[k] = k < 5 ? v[k] : 0;
    if (reset && (hipMemsetAsync(s->particle_substeps, 0, sizeof(v), s->stream) != hipSuccess ||
                  hipStreamSynchronize(s->stream) != hipSuccess))
        return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_scene_diagnostics(void* scene, uint32_t* host_counts, uint32_t n_words) {
    if (!scene || !host_counts) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    uint32_t c[apbf::C_COUNT];
    if (hipMemcpyAsync(c, s->counts, sizeof(c), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return APBF_ERR_HIP;
    for (uint32_t k = 0; k < n_words && k < apbf::C_COUNT; ++k) host_counts[k] = c[k];
    return APBF_OK;
}

const char* apbf_version(void) { return "apbf-mi355x 0.1 (gfx950)"; }

int32_t apbf_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"
