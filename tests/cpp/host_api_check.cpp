// Test driver for include/apbf.hpp: runs the same substeps once through the
// subsystem classes in the order of pool.cpp:67-106 and once through
// scene::update, and requires identical states. Prints one line per case and
// exits non-zero on a mismatch. Built by __graft_entry__.build().
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/apbf.hpp"

using namespace apbf_pbd;

static apbf_settings reference_settings() {
    apbf_settings s{};
    s.mHeightKernelId = 1;
    s.mGradientKernelId = 1;
    s.mMerge = 1;
    s.mSplit = 1;
    s.mBaseKernelWidthOnTargetRadius = 1;
    s.mBaseKernelWidthOnBoundaryDistance = 0;
    s.mUpdateTargetRadius = 1;
    s.mUpdateBoundariness = 1;
    s.mNeighborListSorted = 0;
    s.mBoundarinessCalculationMethod = 2;
    s.mBoundarinessAdaptionSpeed = 0.5f;
    s.mKernelWidthAdaptionSpeed = 0.01f;
    s.mBoundarinessSelfGradLengthFactor = 2.0f;
    s.mBoundarinessUnderpressureFactor = 4.0f;
    s.mMergeDuration = 2.0f;
    s.mSmallestTargetRadius = 0.5f;
    s.mTargetRadiusOffset = 0.0f;
    s.mTargetRadiusScaleFactor = 0.05f;
    return s;
}

// Lattice block of n^3 particles of radius r at spacing 2r (plus a jitter of
// +-0.05 r from a fixed LCG), optionally with a second, finer block above it.
static state make_state(int n, float r, bool fine_drop) {
    state st;
    uint32_t lcg = 12345u;
    auto rnd = [&]() {
        lcg = lcg * 1664525u + 1013904223u;
        return (float)(lcg >> 8) / 16777216.0f - 0.5f;
    };
    auto add = [&](float x, float y, float z, float rad) {
        const float p[3] = {x, y, z};
        for (int d = 0; d < 3; ++d) st.position.push_back((int32_t)std::trunc((double)p[d] * 262144.0));
        st.position.push_back(0);
        for (int d = 0; d < 4; ++d) st.velocity.push_back(0.0f);
        st.inverse_mass.push_back(1.0f / ((2 * rad) * (2 * rad) * (2 * rad)));
        st.radius.push_back(rad);
        st.target_radius.push_back(rad);
        st.kernel_width.push_back(4.0f * rad);
        st.boundariness.push_back(1.0f);
        st.transferring.push_back(0u);
        st.boundary_distance.push_back((uint32_t)std::trunc((double)rad * 262144.0));
    };
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k)
                add(1.0f + r + 2 * r * i + 0.1f * r * rnd(), 1.0f + r + 2 * r * j + 0.1f * r * rnd(),
                    1.0f + r + 2 * r * k + 0.1f * r * rnd(), r);
    if (fine_drop)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n / 2; ++j)
                for (int k = 0; k < n; ++k)
                    add(1.0f + r + r * i + 0.05f * r * rnd(), 1.0f + 2 * r * n + 2 * r + r * j + 0.05f * r * rnd(),
                        1.0f + r + r * k + 0.05f * r * rnd(), 0.5f * r);
    st.pos_backup = st.position;
    return st;
}

static apbf_scene_params make_params(const state& st, bool basic, float extent) {
    apbf_scene_params p;
    std::memset(&p, 0, sizeof(p));
    p.s = reference_settings();
    p.basic_pbf = basic ? 1 : 0;
    p.solver_iterations = 3;
    p.split_duration = 0.0f;
    p.gravity[1] = -9.81f;
    for (int d = 0; d < 3; ++d) {
        p.hash_min[d] = 0.0f;
        p.hash_max[d] = extent;
    }
    p.hash_res_log2 = 4;
    // floor and four side walls, 4 thick (pool.cpp:30-40)
    const float lo = 0.0f, hi = extent, t = 4.0f;
    const float boxes[5][6] = {{lo - t, lo - t, lo - t, hi + t, lo, hi + t},
                               {lo - t, lo - t, lo - t, lo, hi + t, hi + t},
                               {hi, lo - t, lo - t, hi + t, hi + t, hi + t},
                               {lo - t, lo - t, lo - t, hi + t, hi + t, lo},
                               {lo - t, lo - t, hi, hi + t, hi + t, hi + t}};
    p.n_boxes = 5;
    for (int b = 0; b < 5; ++b)
        for (int d = 0; d < 3; ++d) {
            p.box_min[b][d] = boxes[b][d];
            p.box_max[b][d] = boxes[b][3 + d];
        }
    p.cap_particles = 4 * st.size();
    p.cap_transfers = st.size();
    p.cap_pairs = 400 * st.size();
    return p;
}

template <class T>
static bool same(const std::vector<T>& a, const std::vector<T>& b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

static bool same_state(const state& a, const state& b) {
    return same(a.position, b.position) && same(a.velocity, b.velocity) && same(a.inverse_mass, b.inverse_mass) &&
           same(a.radius, b.radius) && same(a.kernel_width, b.kernel_width) && same(a.boundariness, b.boundariness) &&
           same(a.boundary_distance, b.boundary_distance) && same(a.target_radius, b.target_radius) &&
           same(a.t_time_left, b.t_time_left) && a.t_source.size() == b.t_source.size();
}

static int run_case(const char* name, bool basic) {
    const state st = make_state(10, 0.5f, !basic);
    const apbf_scene_params p = make_params(st, basic, 16.0f);
    const float dt = 1.0f / 120.0f;
    scene whole(p), parts(p);
    whole.set_state(st);
    parts.set_state(st);
    velocity_handling velocity;
    particle_transfer transfer;
    neighborhood_green neighborhood;
    spread_kernel_width spread;
    box_collision box;
    incompressibility solve;
    update_transfers transfers;
    velocity.set_data(&parts).set_acceleration(vec3{0.0f, -9.81f, 0.0f});
    transfer.set_data(&parts);
    neighborhood.set_data(&parts).set_position_range(vec3{0, 0, 0}, vec3{16, 16, 16}, 4);
    spread.set_data(&parts);
    box.set_data(&parts);
    solve.set_data(&parts);
    transfers.set_data(&parts);
    for (int step = 0; step < 6; ++step) {
        whole.update(dt);
        // pool::update, pool.cpp:67-106
        velocity.apply(dt);
        if (!basic) transfer.apply(dt);
        neighborhood.set_range_scale(basic ? 1.0f : 1.5f).apply();
        if (!basic) spread.apply();
        for (int it = 0; it < p.solver_iterations; ++it) {
            box.apply();
            solve.apply();
        }
        if (!basic) transfers.apply();
        if (!same_state(whole.get_state(), parts.get_state())) {
            std::printf("%s: substep %d differs\n", name, step);
            return 1;
        }
    }
    std::printf("%s: ok, %u particles\n", name, whole.get_state().size());
    return 0;
}

// pool::update (pool.cpp:67-106) through subsystems bound to lists
// (pool.cpp:27-46: set_data(&mParticles), set_data(&mFluid, &mNeighborsFluid),
// set_data(&mFluid, &mNeighborsFluid, &mTransfers), ...), each apply() running
// its stage operator on the lists; the lists must equal scene::update's state
// after every substep, bit for bit.
static int run_list_case(const char* name, bool basic) {
    const state st = make_state(10, 0.5f, !basic);
    const apbf_scene_params p = make_params(st, basic, 16.0f);
    const float dt = 1.0f / 120.0f;
    scene whole(p);
    whole.set_state(st);
    context ctx(p);
    fluid particles(p.cap_particles);
    transfers trans(p.cap_transfers);
    neighbors nbs(p.cap_particles, p.cap_pairs);
    share_hidden_data(trans, particles);  // pool.cpp:18-19
    particles.upload(st);
    trans.upload(st);
    std::vector<float> bmin(4 * p.n_boxes), bmax(4 * p.n_boxes);
    for (uint32_t b = 0; b < p.n_boxes; ++b)
        for (int d = 0; d < 4; ++d) {
            bmin[4 * b + d] = p.box_min[b][d];
            bmax[4 * b + d] = p.box_max[b][d];
        }
    velocity_handling velocity;
    particle_transfer transfer;
    neighborhood_green neighborhood;
    spread_kernel_width spread;
    box_collision box;
    incompressibility solve;
    update_transfers transfers_update;
    velocity.set_data(&particles).set_acceleration(vec3{0.0f, -9.81f, 0.0f}).set_last_delta_time(st.last_delta_time);
    transfer.set_data(&particles, &trans);
    neighborhood.set_data(&particles, &nbs).set_position_range(vec3{0, 0, 0}, vec3{16, 16, 16}, 4);
    spread.set_data(&particles, &nbs);
    box.set_data(&particles, &bmin, &bmax);
    solve.set_data(&particles, &nbs);
    transfers_update.set_data(&particles, &nbs, &trans);
    for (int step = 0; step < 6; ++step) {
        whole.update(dt);
        velocity.apply(dt);
        if (!basic) transfer.apply(dt);
        neighborhood.set_range_scale(basic ? 1.0f : 1.5f).apply();
        if (!basic) spread.apply();
        for (int it = 0; it < p.solver_iterations; ++it) {
            box.apply();
            solve.apply();
        }
        if (!basic) transfers_update.apply();
        state got;
        particles.download(got);
        trans.download(got);
        const state want = whole.get_state();
        if (!same_state(want, got) || !same(want.t_source, got.t_source) || !same(want.t_target, got.t_target) ||
            !same(want.transferring, got.transferring) || !same(want.pos_backup, got.pos_backup)) {
            std::printf("%s: substep %d differs\n", name, step);
            return 1;
        }
    }
    // the neighbour list in the reference's uvec2 layout
    gpu_list<8> pairs = nbs.pairs();
    const std::vector<uint32_t> pr = pairs.read<uint32_t>();
    bool ordered = !pr.empty();
    for (size_t k = 2; k < pr.size(); k += 2) ordered = ordered && pr[k - 2] <= pr[k];
    if (!ordered) {
        std::printf("%s: pair list not in row order\n", name);
        return 1;
    }
    std::printf("%s: ok, %u particles, %zu pairs\n", name, whole.get_state().size(), pr.size() / 2);
    return 0;
}

// gpu_list (gpu_list.h) on the reference's list vectors: concatenation
// (test.cpp:47-57), append of an empty list (test.cpp:75-89), apply_edit
// (test.cpp:91-105), sort.
static int run_gpu_list_case() {
    bool pass = true;
    {
        std::vector<uint32_t> a = {3u, 64u, 12683u, 4294967295u}, b = {432587u, 0u, 5436u};
        auto la = gpu_list<4>::from_host(a, a.size() + b.size());
        auto lb = gpu_list<4>::from_host(b);
        la += lb;
        std::vector<uint32_t> want = a;
        want.insert(want.end(), b.begin(), b.end());
        pass = la.read<uint32_t>() == want && pass;
    }
    {
        std::vector<uint32_t> a = {3u, 64u, 12683u, 4294967295u};
        auto la = gpu_list<4>::from_host(a, a.size() + 5);
        auto lb = gpu_list<4>::from_host(std::vector<uint32_t>{}, a.size() + 5);
        la += lb;
        pass = la.read<uint32_t>() == a && pass;
    }
    {
        std::vector<uint32_t> v = {77u, 3u, 9999u, 4294967295u, 0u}, e = {1u, 3u, 1u, 4u};
        auto l = gpu_list<4>::from_host(v);
        l.apply_edit(gpu_list<4>::from_host(e));
        pass = l.read<uint32_t>() == std::vector<uint32_t>({3u, 4294967295u, 3u, 0u}) && pass;
    }
    {
        std::vector<uint32_t> v = {5u, 1u, 4u, 1u, 3u};
        auto l = gpu_list<4>::from_host(v);
        auto shared = l;  // lazy copy: sorting l leaves the copy as it was
        l.sort();
        pass = l.read<uint32_t>() == std::vector<uint32_t>({1u, 1u, 3u, 4u, 5u}) && pass;
        pass = shared.read<uint32_t>() == v && pass;
    }
    std::printf("gpu_list: %s\n", pass ? "ok" : "FAILED");
    return pass ? 0 : 1;
}

// spherical_pool::update (spherical_pool.cpp:67-106): the sphere collider in
// place of the boxes, here with the brute-force neighbour search.
static int run_sphere_case() {
    const state st = make_state(8, 0.5f, true);
    apbf_scene_params p = make_params(st, false, 16.0f);
    p.n_boxes = 0;
    const float dt = 1.0f / 120.0f;
    const vec3 center{5.0f, 6.0f, 5.0f};
    scene whole(p), parts(p);
    whole.set_state(st);
    parts.set_state(st);
    const float c[3] = {center.x, center.y, center.z};
    check(apbf_scene_set_sphere(whole.handle(), c, 7.0f, 1, 1), "set_sphere");
    check(apbf_scene_set_neighborhood_method(whole.handle(), 2), "set_neighborhood_method");
    velocity_handling velocity;
    particle_transfer transfer;
    neighborhood_brute_force neighborhood;
    spread_kernel_width spread;
    sphere_collision sphere;
    incompressibility solve;
    update_transfers transfers;
    velocity.set_data(&parts).set_acceleration(vec3{0.0f, -9.81f, 0.0f});
    transfer.set_data(&parts);
    neighborhood.set_data(&parts);
    spread.set_data(&parts);
    sphere.set_data(&parts).set_sphere(center, 7.0f, true);
    solve.set_data(&parts);
    transfers.set_data(&parts);
    for (int step = 0; step < 4; ++step) {
        whole.update(dt);
        velocity.apply(dt);
        transfer.apply(dt);
        neighborhood.set_range_scale(1.5f).apply();
        spread.apply();
        for (int it = 0; it < p.solver_iterations; ++it) {
            sphere.apply();
            solve.apply();
        }
        transfers.apply();
        if (!same_state(whole.get_state(), parts.get_state())) {
            std::printf("spherical pool: substep %d differs\n", step);
            return 1;
        }
    }
    std::printf("spherical pool: ok, %u particles\n", whole.get_state().size());
    return 0;
}

// time_machine: ten substeps in the present, back to the oldest keyframe,
// forward again through the past; every substep must equal its recording.
static int run_time_machine_case() {
    const state st = make_state(8, 0.5f, true);
    const apbf_scene_params p = make_params(st, false, 16.0f);
    scene s(p);
    s.set_state(st);
    time_machine tm(s);
    tm.set_max_keyframes(3).set_keyframe_interval(3).enable();
    std::vector<state> rec(11);
    rec[0] = s.get_state();
    auto dt_of = [](int k) { return (1.0f + 0.25f * (float)(k % 3)) / 120.0f; };
    for (int k = 1; k <= 10; ++k) {
        float dt = dt_of(k);
        if (tm.step_forward(dt)) return 1;
        s.update(dt);
        tm.save_state(dt);
        rec[k] = s.get_state();
    }
    int back = 0;
    while (tm.jump_back()) ++back;
    if (back != 3 || !same_state(s.get_state(), rec[3])) {
        std::printf("time machine: jump back %d\n", back);
        return 1;
    }
    for (int k = 4; k <= 10; ++k) {
        float dt = 1.0f;
        if (!tm.step_forward(dt)) {
            if (dt != dt_of(k)) return 1;
            s.update(dt);
            tm.save_state(dt);
        }
        if (!same_state(s.get_state(), rec[k])) {
            std::printf("time machine: replayed substep %d differs\n", k);
            return 1;
        }
    }
    std::printf("time machine: ok\n");
    return 0;
}

// The reference's own time-machine test (source/test.cpp:585-619): at most
// two keyframes, one every 3 steps; three steps in the present, jump_back,
// then step_forward four times with the expected returns. The reference's
// sparse list is the scene, its per-step value the recorded time step.
static int run_time_machine_reference_sequence() {
    const state st = make_state(6, 0.5f, false);
    const apbf_scene_params p = make_params(st, true, 16.0f);
    scene s(p);
    s.set_state(st);
    time_machine tm(s);
    tm.set_max_keyframes(2).set_keyframe_interval(3).enable(0.0f);
    std::vector<state> rec(5);
    rec[0] = s.get_state();
    auto dt_of = [](int k) { return (1.0f + 0.1f * (float)k) / 120.0f; };
    bool pass = true;
    for (int k = 1; k <= 3; ++k) {
        float dt = dt_of(k);
        pass = !tm.step_forward(dt) && pass;
        s.update(dt);
        tm.save_state(dt);
        rec[k] = s.get_state();
    }
    pass = tm.jump_back() && pass;
    pass = tm.current_dt() == 0.0f && same_state(s.get_state(), rec[0]) && pass;
    for (int k = 1; k <= 2; ++k) {
        float dt = 99.0f;
        pass = !tm.step_forward(dt) && pass;
        pass = dt == dt_of(k) && tm.current_dt() == dt_of(k) && pass;
        s.update(dt);
        tm.save_state(dt);
        pass = same_state(s.get_state(), rec[k]) && pass;
    }
    float dt = 99.0f;
    pass = tm.step_forward(dt) && pass;  // keyframe 3 loaded
    tm.save_state(dt);
    pass = tm.current_dt() == dt_of(3) && same_state(s.get_state(), rec[3]) && pass;
    dt = dt_of(4);
    pass = !tm.step_forward(dt) && !tm.in_past() && dt == dt_of(4) && pass;  // the present again
    std::printf("time machine reference sequence: %s\n", pass ? "ok" : "FAILED");
    return pass ? 0 : 1;
}

int main() {
    if (apbf_device_count() <= 0) {
        std::printf("no device\n");
        return 2;
    }
    try {
        return run_case("basic pbf", true) | run_case("adaptive", false) | run_sphere_case() |
               run_time_machine_case() | run_time_machine_reference_sequence() | run_list_case("lists basic pbf", true) |
               run_list_case("lists adaptive", false) | run_gpu_list_case();
    } catch (const error& e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
}
