// Device-resident state snapshots: the keyframes of the reference's time
// machine (time_machine.h:46-143), which copies the particle lists, the fluid
// lists and the transfer lists of a scene (pool.cpp:9) and puts them back
// when the user steps into the past. A snapshot here is a set of device
// buffers sized for the scene's capacities; saving and restoring are
// device-to-device copies on the scene's stream (no host synchronisation, so
// a keyframe every N substeps costs one copy of ~80 B per particle slot).
//
// Saved: the count block, every per-particle list of the current buffer set
// (positions, velocities, inverse masses, radii, position backups, transfer
// flags, target radii, kernel widths, boundariness, boundary distances, the
// warm start's previous lambda), the transfer lists and the last time step of
// velocity_handling. Not saved, as in the reference: the neighbour list,
// which every substep rebuilds; a restored scene reports no pairs until its
// next neighbourhood stage.
#include "scene.h"

namespace apbf {
namespace {

__global__ void k_clear_pair_counts(uint32_t* __restrict__ counts) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        counts[C_PAIRS] = 0u;
        counts[C_ROWS] = 0u;
    }
}

}  // namespace
}  // namespace apbf

using namespace apbf;

struct apbf_snapshot {
    uint32_t cap = 0, tcap = 0;
    int device = 0;
    void* mem = nullptr;
    float last_dt = 1.0f;
    // views into mem
    uint32_t* counts = nullptr;
    int4 *pos = nullptr, *backup = nullptr;
    float4* vel = nullptr;
    float *inv_mass = nullptr, *radius = nullptr, *target_radius = nullptr, *kernel_width = nullptr;
    float *boundariness = nullptr, *lam_prev = nullptr, *t_time_left = nullptr;
    uint32_t *transferring = nullptr, *boundary_distance = nullptr, *t_source = nullptr, *t_target = nullptr;
    uint32_t* sid = nullptr;  // scene index per particle (packed batches)
    bool packed = false;
};

namespace {

size_t carve(apbf_snapshot* s) {
    // 256-byte aligned sub-buffers; returns the total size (views set when mem is non-null)
    size_t off = 0;
    char* base = static_cast<char*>(s->mem);
    auto take = [&](auto*& p, size_t bytes) {
        off = (off + 255u) & ~(size_t)255u;
        if (base) p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(base + off);
        off += bytes;
    };
    const size_t c = s->cap, t = s->tcap;
    take(s->counts, (size_t)C_COUNT * 4);
    take(s->pos, c * 16);
    take(s->backup, c * 16);
    take(s->vel, c * 16);
    take(s->inv_mass, c * 4);
    take(s->radius, c * 4);
    take(s->target_radius, c * 4);
    take(s->kernel_width, c * 4);
    take(s->boundariness, c * 4);
    take(s->lam_prev, c * 4);
    take(s->transferring, c * 4);
    take(s->boundary_distance, c * 4);
    take(s->t_source, t * 4);
    take(s->t_target, t * 4);
    take(s->t_time_left, t * 4);
    take(s->sid, s->packed ? c * 4 : 0);
    return off + 256;
}

// Copies every saved list between the snapshot and the scene's current
// buffer sets; `to_scene` picks the direction.
apbf_status copy_state(apbf_snapshot* snap, Scene* sc, bool to_scene) {
    hipStream_t st = sc->stream;
    const int c = sc->cur, t = sc->tcur;
    bool ok = true;
    auto cp = [&](void* scene_buf, void* snap_buf, size_t bytes) {
        if (!bytes) return;
        void* dst = to_scene ? scene_buf : snap_buf;
        const void* src = to_scene ? snap_buf : scene_buf;
        ok = ok && hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st) == hipSuccess;
    };
    const size_t n = sc->cap, nt = sc->tcap;
    cp(sc->counts, snap->counts, (size_t)C_COUNT * 4);
    cp(sc->pos[c], snap->pos, n * 16);
    cp(sc->backup[c], snap->backup, n * 16);
    cp(sc->vel[c], snap->vel, n * 16);
    cp(sc->inv_mass[c], snap->inv_mass, n * 4);
    cp(sc->radius[c], snap->radius, n * 4);
    cp(sc->target_radius[c], snap->target_radius, n * 4);
    cp(sc->kernel_width[c], snap->kernel_width, n * 4);
    cp(sc->boundariness[c], snap->boundariness, n * 4);
    cp(sc->lam_prev[c], snap->lam_prev, n * 4);
    cp(sc->transferring[c], snap->transferring, n * 4);
    cp(sc->boundary_distance[c], snap->boundary_distance, n * 4);
    cp(sc->t_source[t], snap->t_source, nt * 4);
    cp(sc->t_target[t], snap->t_target, nt * 4);
    cp(sc->t_time_left[t], snap->t_time_left, nt * 4);
    if (snap->packed) cp(sc->sid[c], snap->sid, n * 4);
    if (to_scene) {
        sc->last_dt = snap->last_dt;
        hipLaunchKernelGGL(k_clear_pair_counts, dim3(1), dim3(64), 0, st, sc->counts);
        sc->pairs_valid = false;
    } else {
        snap->last_dt = sc->last_dt;
    }
    return ok && hipGetLastError() == hipSuccess ? APBF_OK : APBF_ERR_HIP;
}

}  // namespace

extern "C" {

apbf_status apbf_snapshot_create(void* scene, apbf_snapshot** out) {
    if (!scene || !out) return APBF_ERR_INVALID_ARGUMENT;
    Scene* sc = static_cast<Scene*>(scene);
    apbf_snapshot* s = new apbf_snapshot();
    s->cap = sc->cap;
    s->tcap = sc->tcap;
    s->device = sc->device;
    s->packed = sc->packed();
    const size_t bytes = carve(s);
    if (hipMalloc(&s->mem, bytes) != hipSuccess) {
        delete s;
        return APBF_ERR_HIP;
    }
    carve(s);
    const apbf_status st = copy_state(s, sc, false);
    if (st != APBF_OK) {
        (void)hipFree(s->mem);
        delete s;
        return st;
    }
    *out = s;
    return APBF_OK;
}

apbf_status apbf_snapshot_save(apbf_snapshot* snap, void* scene) {
    if (!snap || !scene) return APBF_ERR_INVALID_ARGUMENT;
    Scene* sc = static_cast<Scene*>(scene);
    if (sc->cap != snap->cap || sc->tcap != snap->tcap || sc->packed() != snap->packed) return APBF_ERR_INVALID_ARGUMENT;
    return copy_state(snap, sc, false);
}

apbf_status apbf_snapshot_restore(apbf_snapshot* snap, void* scene) {
    if (!snap || !scene) return APBF_ERR_INVALID_ARGUMENT;
    Scene* sc = static_cast<Scene*>(scene);
    if (sc->cap != snap->cap || sc->tcap != snap->tcap) return APBF_ERR_INVALID_ARGUMENT;
    return copy_state(snap, sc, true);
}

apbf_status apbf_snapshot_destroy(apbf_snapshot* snap) {
    if (!snap) return APBF_ERR_INVALID_ARGUMENT;
    if (snap->mem) (void)hipFree(snap->mem);
    delete snap;
    return APBF_OK;
}

}  // extern "C"
