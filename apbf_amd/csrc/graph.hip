// Substeps as HIP graphs, batches of independent scenes, and the launch-shape
// autotuner.
//
// A substep is a fixed sequence of kernels whose sizes come from the device
// count block, so stream capture records it once and a graph launch replays
// it without host work. The host-side buffer roles (current particle and
// transfer buffer sets) flip a fixed number of times per substep, so two
// substeps always return them to their starting roles: the captured graph
// holds two substeps and every replay advances the scene by two.
//
// A batch steps many scenes, each on a stream of its own so the device runs
// scenes of uneven size side by side. Captured, every scene's two substeps
// become a graph of their own, and a replay launches all of them on the
// scenes' streams. (One graph holding every scene as a parallel branch was
// measured slower than the eager streams: its branches did not overlap.)
#include <vector>

#include "scene.h"

using namespace apbf;

namespace {

struct HostRoles {
    int cur, tcur;
    const void *row, *col, *u0, *col_tmp;
    bool operator==(const HostRoles& o) const {
        return cur == o.cur && tcur == o.tcur && row == o.row && col == o.col && u0 == o.u0 && col_tmp == o.col_tmp;
    }
};
HostRoles roles(const Scene* s) { return HostRoles{s->cur, s->tcur, s->row, s->col, s->u0, s->col_tmp}; }

apbf_status ensure_stream(Scene* s) {
    if (s->stream) return APBF_OK;
    if (!s->own_stream && hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking) != hipSuccess)
        return APBF_ERR_HIP;
    // the new stream does not wait for the null stream by itself
    return scene_switch_stream(s, s->own_stream);
}

void drop_graph(Scene* s) {
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->graph_exec = nullptr;
    s->graph = nullptr;
}

}  // namespace

namespace apbf {
// Work already queued on the scene's current stream finishes before anything
// queued on `next` afterwards: an event recorded on the old stream, waited
// for by the new one (no host synchronisation).
apbf_status scene_switch_stream(Scene* s, hipStream_t next) {
    if (next == s->stream) return APBF_OK;
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return APBF_ERR_HIP;
    apbf_status st = APBF_OK;
    if (hipEventRecord(ev, s->stream) != hipSuccess || hipStreamWaitEvent(next, ev, 0) != hipSuccess)
        st = APBF_ERR_HIP;
    (void)hipEventDestroy(ev);
    if (st == APBF_OK) s->stream = next;
    return st;
}
void scene_drop_graph(Scene* s) { drop_graph(s); }
void scene_release_graph(Scene* s) {
    drop_graph(s);
    if (s->own_stream) {
        if (s->stream == s->own_stream) s->stream = nullptr;
        (void)hipStreamDestroy(s->own_stream);
    }
    s->own_stream = nullptr;
}
}  // namespace apbf

struct apbf_batch {
    std::vector<Scene*> scenes;
    bool captured = false;
};

extern "C" {

apbf_status apbf_scene_capture(void* scene, float dt) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    if (s->timed_mask) return APBF_ERR_INVALID_ARGUMENT;  // event timing is not captured
    apbf_status st = ensure_stream(s);
    if (st != APBF_OK) return st;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return APBF_ERR_HIP;
    drop_graph(s);
    const HostRoles before = roles(s);
    if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return APBF_ERR_HIP;
    for (int k = 0; k < 2; ++k) s->update(dt);
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(s->stream, &g) != hipSuccess || !g) return APBF_ERR_HIP;
    if (!(roles(s) == before)) {  // cannot happen for the substep sequence; refuse rather than replay wrongly
        (void)hipGraphDestroy(g);
        return APBF_ERR_INVALID_ARGUMENT;
    }
    s->graph = g;
    if (hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
        drop_graph(s);
        return APBF_ERR_HIP;
    }
    s->graph_dt = dt;
    return APBF_OK;
}

apbf_status apbf_scene_replay(void* scene, uint32_t times) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    if (!s->graph_exec) return APBF_ERR_INVALID_ARGUMENT;
    for (uint32_t k = 0; k < times; ++k)
        if (hipGraphLaunch(s->graph_exec, s->stream) != hipSuccess) return APBF_ERR_HIP;
    s->pairs_valid = true;
    return APBF_OK;
}

apbf_status apbf_batch_create(void* const* scenes, uint32_t n, apbf_batch** out) {
    if (!scenes || !n || !out) return APBF_ERR_INVALID_ARGUMENT;
    apbf_batch* b = new apbf_batch();
    for (uint32_t k = 0; k < n; ++k) {
        Scene* s = static_cast<Scene*>(scenes[k]);
        if (!s) {
            apbf_batch_destroy(b);
            return APBF_ERR_INVALID_ARGUMENT;
        }
        if (!s->stream) {  // every scene gets a stream of its own so the batch can run them side by side
            if (ensure_stream(s) != APBF_OK) {
                apbf_batch_destroy(b);
                return APBF_ERR_HIP;
            }
        }
        b->scenes.push_back(s);
    }
    *out = b;
    return APBF_OK;
}

apbf_status apbf_batch_destroy(apbf_batch* b) {
    if (!b) return APBF_ERR_INVALID_ARGUMENT;
    delete b;
    return APBF_OK;
}

// One substep of every scene, each on its own stream (no graph).
apbf_status apbf_batch_update(apbf_batch* b, float dt) {
    if (!b) return APBF_ERR_INVALID_ARGUMENT;
    for (Scene* s : b->scenes) s->update(dt);
    return hipGetLastError() == hipSuccess ? APBF_OK : APBF_ERR_HIP;
}

apbf_status apbf_batch_capture(apbf_batch* b, float dt) {
    if (!b) return APBF_ERR_INVALID_ARGUMENT;
    for (Scene* s : b->scenes) {
        apbf_status st = apbf_scene_capture(s, dt);
        if (st != APBF_OK) return st;
    }
    b->captured = true;
    return APBF_OK;
}

apbf_status apbf_batch_replay(apbf_batch* b, uint32_t times) {
    if (!b || !b->captured) return APBF_ERR_INVALID_ARGUMENT;
    for (uint32_t k = 0; k < times; ++k)
        for (Scene* s : b->scenes) {
            if (!s->graph_exec) return APBF_ERR_INVALID_ARGUMENT;  // a scene's graph was dropped since
            if (hipGraphLaunch(s->graph_exec, s->stream) != hipSuccess) return APBF_ERR_HIP;
        }
    for (Scene* s : b->scenes) s->pairs_valid = true;
    return APBF_OK;
}

apbf_status apbf_batch_synchronize(apbf_batch* b) {
    if (!b) return APBF_ERR_INVALID_ARGUMENT;
    for (Scene* s : b->scenes)
        if (hipStreamSynchronize(s->stream) != hipSuccess) return APBF_ERR_HIP;
    return APBF_OK;
}

apbf_status apbf_scene_set_launch_shape(void* scene, int32_t walk_lanes, int32_t solve_lanes) {
    if (!scene || (walk_lanes != 0 && walk_lanes != 4 && walk_lanes != 8 && walk_lanes != 16 && walk_lanes != 64) ||
        (solve_lanes != 0 && solve_lanes != 8 && solve_lanes != 16 && solve_lanes != 32))
        return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    s->walk_auto = walk_lanes == 0;  // 0: lanes from a recent mean row length again
    if (walk_lanes != 0) s->walk_lanes = walk_lanes;
    s->solve_auto = solve_lanes == 0;
    if (solve_lanes != 0) s->solve_lanes = solve_lanes;
    drop_graph(s);  // a captured graph holds the old shapes
    return APBF_OK;
}

apbf_status apbf_scene_set_sort_shape(void* scene, int32_t digit_bits, int32_t items_per_thread) {
    if (!scene || digit_bits <= 0 || items_per_thread <= 0) return APBF_ERR_INVALID_ARGUMENT;
    const SortShape sh{(uint32_t)digit_bits, (uint32_t)items_per_thread};
    if (!sort_shape_valid(sh)) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    s->sort_shape = sh;
    drop_graph(s);
    return APBF_OK;
}

apbf_status apbf_scene_sort_shape(void* scene, int32_t* digit_bits, int32_t* items_per_thread) {
    if (!scene) return APBF_ERR_INVALID_ARGUMENT;
    const Scene* s = static_cast<const Scene*>(scene);
    if (digit_bits) *digit_bits = (int32_t)s->sort_shape.bits;
    if (items_per_thread) *items_per_thread = (int32_t)s->sort_shape.items;
    return APBF_OK;
}

apbf_status apbf_scene_set_long_rows(void* scene, int32_t mode) {
    if (!scene || mode < -1 || mode > 1) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    s->long_rows_mode = mode;
    drop_graph(s);  // a captured graph holds the old variant
    return APBF_OK;
}

// Times `substeps` substeps under every launch shape, each from the same saved
// state, and keeps the fastest; the scene is put back afterwards. All shapes
// compute identical results.
apbf_status apbf_scene_autotune(void* scene, float dt, uint32_t substeps, int32_t* best_walk, int32_t* best_solve,
                                float* best_ms) {
    if (!scene || substeps == 0) return APBF_ERR_INVALID_ARGUMENT;
    Scene* s = static_cast<Scene*>(scene);
    // Every shape runs the same substeps from the same state: the scene is
    // saved first and restored before each trial and at the end, together
    // with its statistics, so that shapes tried later do not time a scene
    // that has moved on, and tuning leaves the scene as it found it (except
    // the neighbour list, which a restore empties until the next substep).
    // Each trial runs one untimed substep, then times `substeps`; every shape
    // is tried twice and keeps its faster time.
    apbf_snapshot* snap = nullptr;
    apbf_status st = apbf_snapshot_create(scene, &snap);
    if (st != APBF_OK) return st;
    const size_t err_bytes = 6 * sizeof(double), iter_bytes = MAX_ITER_RECORD * sizeof(double);
    const size_t ps_bytes = 5 * sizeof(unsigned long long);
    char* stats = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipMalloc(&stats, err_bytes + iter_bytes + ps_bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        st = APBF_ERR_HIP;
    auto copy_stats = [&](bool save) {
        auto cp = [&](void* scene_buf, char* keep, size_t bytes) {
            if (hipMemcpyAsync(save ? keep : scene_buf, save ? scene_buf : keep, bytes, hipMemcpyDeviceToDevice,
                               s->stream) != hipSuccess)
                st = APBF_ERR_HIP;
        };
        cp(s->err_stats, stats, err_bytes);
        cp(s->iter_err, stats + err_bytes, iter_bytes);
        cp(s->particle_substeps, stats + err_bytes + iter_bytes, ps_bytes);
    };
    const uint32_t timed_mask = s->timed_mask;
    const int walk0 = s->walk_lanes, solve0 = s->solve_lanes;
    const bool auto0 = s->walk_auto, sauto0 = s->solve_auto;
    s->walk_auto = false;
    s->solve_auto = false;
    s->timed_mask = 0;
    if (st == APBF_OK) copy_stats(true);
    const int walks[3] = {4, 8, 64}, solves[3] = {8, 16, 32};
    float trial_ms[3][3];
    for (int r = 0; r < 2 && st == APBF_OK; ++r)
        for (int a = 0; a < 3 && st == APBF_OK; ++a)
            for (int b = 0; b < 3 && st == APBF_OK; ++b) {
                s->walk_lanes = walks[a];
                s->solve_lanes = solves[b];
                st = apbf_snapshot_restore(snap, scene);
                if (st != APBF_OK) break;
                s->update(dt);
                float ms = 0.0f;
                if (hipEventRecord(e0, s->stream) != hipSuccess) st = APBF_ERR_HIP;
                for (uint32_t k = 0; k < substeps; ++k) s->update(dt);
                if (hipEventRecord(e1, s->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                    st = APBF_ERR_HIP;
                trial_ms[a][b] = r == 0 ? ms : fminf(trial_ms[a][b], ms);
            }
    float best = 1e30f;
    int bw = walk0, bs = solve0;
    if (st == APBF_OK)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                if (trial_ms[a][b] < best) {
                    best = trial_ms[a][b];
                    bw = walks[a];
                    bs = solves[b];
                }
    // The hash sort's shape, timed alone on the keys of one substep from the
    // saved state (the sort is a small part of a substep, below the noise of
    // substep timings). Sorting the hash keys again only rewrites the sorted
    // copies, which the next neighbourhood build recomputes.
    SortShape best_sort = s->sort_shape;
    if (st == APBF_OK) st = apbf_snapshot_restore(snap, scene);
    if (st == APBF_OK) {
        s->walk_lanes = bw;
        s->solve_lanes = bs;
        s->update(dt);
        const uint32_t upper = s->packed() ? (s->n_seg << s->seg_shift) : (1u << (s->grid.res * s->dims()));
        float best_sort_ms = 1e30f;
        for (uint32_t bits : {4u, 8u, 11u})
            for (uint32_t items : {4u, 8u, 16u}) {
                const SortShape sh{bits, items};
                float ms = 0.0f;
                radix_sort_pairs(s->keys, s->vals, s->counts + C_N, s->cap, upper, s->skeys, s->perm, s->ws_mem,
                                 s->stream, sh);
                if (hipEventRecord(e0, s->stream) != hipSuccess) st = APBF_ERR_HIP;
                for (int k = 0; k < 5; ++k)
                    radix_sort_pairs(s->keys, s->vals, s->counts + C_N, s->cap, upper, s->skeys, s->perm, s->ws_mem,
                                     s->stream, sh);
                if (hipEventRecord(e1, s->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                    st = APBF_ERR_HIP;
                if (st == APBF_OK && ms < best_sort_ms) {
                    best_sort_ms = ms;
                    best_sort = sh;
                }
            }
    }
    // put the scene back even after a failed trial
    const apbf_status back = apbf_snapshot_restore(snap, scene);
    if (st == APBF_OK) st = back;
    if (stats) {
        copy_stats(false);
        (void)hipStreamSynchronize(s->stream);
        (void)hipFree(stats);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)apbf_snapshot_destroy(snap);
    s->timed_mask = timed_mask;
    s->walk_lanes = st == APBF_OK ? bw : walk0;
    s->walk_auto = st == APBF_OK ? false : auto0;
    s->solve_auto = st == APBF_OK ? false : sauto0;
    s->solve_lanes = st == APBF_OK ? bs : solve0;
    if (st == APBF_OK) s->sort_shape = best_sort;
    drop_graph(s);
    if (st != APBF_OK) return st;
    if (best_walk) *best_walk = bw;
    if (best_solve) *best_solve = bs;
    if (best_ms) *best_ms = best / (float)substeps;
    return APBF_OK;
}

}  // extern "C"
