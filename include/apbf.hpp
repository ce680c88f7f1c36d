/* C++ host interface with the shape of the reference's subsystem classes
 * (namespace pbd in source/*.h): each class binds to a scene with set_data
 * and runs its stage with apply(), so a caller written against
 * velocity_handling / particle_transfer / neighborhood_green /
 * spread_kernel_width / box_collision / incompressibility / update_transfers
 * keeps its call sequence (pool.cpp:67-106). Header-only over include/apbf.h;
 * errors become apbf_pbd::error exceptions (the reference asserts). */
#ifndef APBF_HPP
#define APBF_HPP

#include <algorithm>
#include <deque>
#include <stdexcept>
#include <string>
#include <vector>

#include "apbf.h"

namespace apbf_pbd {

struct error : std::runtime_error {
    apbf_status status;
    error(const char* what, apbf_status s) : std::runtime_error(std::string(what) + ": status " + std::to_string(s)), status(s) {}
};

inline void check(apbf_status s, const char* what) {
    if (s != APBF_OK) throw error(what, s);
}

struct vec3 {
    float x, y, z;
};

/* Host copy of one scene state (the lists of list_definitions.h:9-20 with the
 * fluid stored in hidden-slot order). */
struct state {
    std::vector<int32_t> position, pos_backup;  // 4 per particle
    std::vector<float> velocity;                // 4 per particle
    std::vector<float> inverse_mass, radius, target_radius, kernel_width, boundariness;
    std::vector<uint32_t> transferring, boundary_distance;
    std::vector<uint32_t> t_source, t_target;
    std::vector<float> t_time_left;
    float last_delta_time = 1.0f;
    uint32_t size() const { return (uint32_t)radius.size(); }
};

/* The device-side lists of one simulation (the pool / waterdrop members). */
class scene {
public:
    explicit scene(const apbf_scene_params& p) : params_(p) { check(apbf_scene_create(&p, &h_), "apbf_scene_create"); }
    ~scene() {
        if (h_) apbf_scene_destroy(h_);
    }
    scene(const scene&) = delete;
    scene& operator=(const scene&) = delete;

    void* handle() const { return h_; }
    const apbf_scene_params& params() const { return params_; }

    void set_state(const state& s) {
        check(apbf_scene_set_state(h_, s.size(), s.position.data(), s.velocity.data(), s.inverse_mass.data(),
                                   s.radius.data(), s.pos_backup.data(), s.transferring.data(), nullptr,
                                   s.target_radius.data(), s.kernel_width.data(), s.boundariness.data(),
                                   s.boundary_distance.data(), (uint32_t)s.t_source.size(), s.t_source.data(),
                                   s.t_target.data(), s.t_time_left.data(), s.last_delta_time),
              "apbf_scene_set_state");
    }
    state get_state() const {
        uint32_t n = 0, nt = 0, np = 0;
        check(apbf_scene_get_counts(h_, &n, &nt, &np), "apbf_scene_get_counts");
        state s;
        s.position.resize(4 * n);
        s.pos_backup.resize(4 * n);
        s.velocity.resize(4 * n);
        s.inverse_mass.resize(n);
        s.radius.resize(n);
        s.target_radius.resize(n);
        s.kernel_width.resize(n);
        s.boundariness.resize(n);
        s.transferring.resize(n);
        s.boundary_distance.resize(n);
        s.t_source.resize(nt);
        s.t_target.resize(nt);
        s.t_time_left.resize(nt);
        check(apbf_scene_get_state(h_, s.position.data(), s.velocity.data(), s.inverse_mass.data(), s.radius.data(),
                                   s.pos_backup.data(), s.transferring.data(), s.target_radius.data(),
                                   s.kernel_width.data(), s.boundariness.data(), s.boundary_distance.data(),
                                   s.t_source.data(), s.t_target.data(), s.t_time_left.data(), nullptr,
                                   &s.last_delta_time),
              "apbf_scene_get_state");
        return s;
    }
    /* pool::update / waterdrop::update (pool.cpp:67-106): one substep. */
    void update(float dt) { check(apbf_scene_update(h_, dt), "apbf_scene_update"); }
    void stage(int32_t id, float dt) { check(apbf_scene_stage(h_, id, dt), "apbf_scene_stage"); }

private:
    apbf_scene_params params_;
    void* h_ = nullptr;
};

/* velocity_handling.h:7-18 */
class velocity_handling {
public:
    velocity_handling& set_data(scene* s) {
        s_ = s;
        return *this;
    }
    velocity_handling& set_acceleration(const vec3& a = vec3{0, 0, 0}) {
        const float g[3] = {a.x, a.y, a.z};
        check(apbf_scene_set_gravity(s_->handle(), g), "set_acceleration");
        return *this;
    }
    void apply(float delta_time) { s_->stage(APBF_STAGE_VELOCITY, delta_time); }

private:
    scene* s_ = nullptr;
};

/* particle_transfer.h: set_data, apply(dt) */
class particle_transfer {
public:
    particle_transfer& set_data(scene* s) {
        s_ = s;
        return *this;
    }
    void apply(float delta_time) { s_->stage(APBF_STAGE_PARTICLE_TRANSFER, delta_time); }

private:
    scene* s_ = nullptr;
};

/* neighborhood_green.h:9-24 */
class neighborhood_green {
public:
    neighborhood_green& set_data(scene* s) {
        s_ = s;
        return *this;
    }
    neighborhood_green& set_range_scale(float scale) {
        check(apbf_scene_set_range_scale(s_->handle(), scale), "set_range_scale");
        return *this;
    }
    neighborhood_green& set_position_range(const vec3& mn, const vec3& mx, uint32_t resolution_log2) {
        const float a[3] = {mn.x, mn.y, mn.z}, b[3] = {mx.x, mx.y, mx.z};
        check(apbf_scene_set_position_range(s_->handle(), a, b, resolution_log2), "set_position_range");
        return *this;
    }
    void apply() { s_->stage(APBF_STAGE_NEIGHBORHOOD, 0.0f); }

private:
    scene* s_ = nullptr;
};

/* neighborhood_binary_search.h / neighborhood_brute_force.h: the scene's
 * search is one of the three (NEIGHBORHOOD_TYPE, cpu_gpu_shared_config.h:54-60);
 * these select it and run it, with the range scale of neighborhood_green. */
class neighborhood_binary_search : public neighborhood_green {
public:
    neighborhood_binary_search& set_data(scene* s) {
        neighborhood_green::set_data(s);
        check(apbf_scene_set_neighborhood_method(s->handle(), 1), "set_neighborhood_method");
        return *this;
    }
};
class neighborhood_brute_force : public neighborhood_green {
public:
    neighborhood_brute_force& set_data(scene* s) {
        neighborhood_green::set_data(s);
        check(apbf_scene_set_neighborhood_method(s->handle(), 2), "set_neighborhood_method");
        return *this;
    }
};

/* spread_kernel_width.h:10-11 */
class spread_kernel_width {
public:
    spread_kernel_width& set_data(scene* s) {
        s_ = s;
        return *this;
    }
    void apply() { s_->stage(APBF_STAGE_SPREAD, 0.0f); }

private:
    scene* s_ = nullptr;
};

/* box_collision.h */
class box_collision {
public:
    box_collision& set_data(scene* s) {
        s_ = s;
        return *this;
    }
    void apply() { s_->stage(APBF_STAGE_BOX, 0.0f); }

private:
    scene* s_ = nullptr;
};

/* sphere_collision.h: set_sphere(center, radius, hollow), apply() */
class sphere_collision {
public:
    sphere_collision& set_data(scene* s) {
        s_ = s;
        return *this;
    }
    sphere_collision& set_sphere(const vec3& center, float radius, bool hollow) {
        const float c[3] = {center.x, center.y, center.z};
        check(apbf_scene_set_sphere(s_->handle(), c, radius, hollow ? 1 : 0, 1), "set_sphere");
        return *this;
    }
    void apply() { s_->stage(APBF_STAGE_SPHERE, 0.0f); }

private:
    scene* s_ = nullptr;
};

/* incompressibility.h:9-18: one solver iteration per apply() */
class incompressibility {
public:
    incompressibility& set_data(scene* s) {
        s_ = s;
        return *this;
    }
    void apply() { s_->stage(APBF_STAGE_INCOMPRESSIBILITY, 0.0f); }

private:
    scene* s_ = nullptr;
};

/* update_transfers.h:11-12 */
class update_transfers {
public:
    update_transfers& set_data(scene* s) {
        s_ = s;
        return *this;
    }
    void apply() { s_->stage(APBF_STAGE_UPDATE_TRANSFERS, 0.0f); }

private:
    scene* s_ = nullptr;
};

/* time_machine.h:9-143 for one scene: keyframes every keyframe_interval
 * substeps as device snapshots (at most max_keyframes), the time step of every
 * substep on the host. As in the scenes' update (pool.cpp:67-106):
 *     if (!tm.step_forward(dt)) { s.update(dt); tm.save_state(dt); }
 * step_forward replaces dt by the recorded step in the past and returns true
 * when it loaded a keyframe (the substep is then skipped). */
class time_machine {
public:
    explicit time_machine(scene& s) : s_(s) {}
    ~time_machine() {
        for (apbf_snapshot* k : keys_) apbf_snapshot_destroy(k);
        for (apbf_snapshot* k : spare_) apbf_snapshot_destroy(k);
    }
    time_machine(const time_machine&) = delete;
    time_machine& operator=(const time_machine&) = delete;

    time_machine& set_max_keyframes(uint32_t k) {
        max_keyframes_ = k;
        return *this;
    }
    time_machine& set_keyframe_interval(uint32_t i) {
        interval_ = i;
        drop_front(0);
        dts_.clear();
        return *this;
    }
    time_machine& enable(float dt = 1.0f) {
        if (!enabled_) {
            keys_.push_back(take());
            dts_.push_back(dt);
        }
        enabled_ = true;
        return *this;
    }
    time_machine& disable() {
        drop_front(0);
        dts_.clear();
        enabled_ = false;
        current_ = present_ = 0;
        return *this;
    }
    bool enabled() const { return enabled_; }
    bool on_keyframe() const { return current_ % interval_ == 0u; }
    bool in_past() const { return current_ != present_; }
    uint32_t keyframe_step_for(uint32_t step) const { return step / interval_ * interval_; }
    uint32_t oldest_history_keyframe_step() const {
        return keyframe_step_for(present_) - (max_keyframes_ - 1u) * interval_;
    }
    uint32_t history_index() const { return std::min(current_, current_ - oldest_history_keyframe_step()); }
    uint32_t sparse_history_index() const { return history_index() / interval_; }

    bool step_forward(float& dt) {
        if (!enabled_) return false;
        const bool past = in_past();
        ++current_;
        present_ += past ? 0u : 1u;
        const bool loaded = past && on_keyframe();
        if (past) {
            dt = dts_[history_index()];
            if (loaded) load();
        } else if (on_keyframe()) {
            drop_front(max_keyframes_ - 1u);
            const size_t keep = (size_t)(max_keyframes_ - 1u) * interval_;
            while (dts_.size() > keep) dts_.pop_front();
        }
        return loaded;
    }
    void save_state(float dt) {
        if (!enabled_) return;
        if (keys_.size() <= sparse_history_index()) keys_.push_back(take());
        if (dts_.size() <= history_index()) dts_.push_back(dt);
    }
    bool jump_back() {
        if ((on_keyframe() && current_ + interval_ * (max_keyframes_ - 1u) <= present_) || current_ == 0u)
            return false;
        current_ = (current_ - 1u) / interval_ * interval_;
        load();
        return true;
    }

private:
    apbf_snapshot* take() {
        if (!spare_.empty()) {
            apbf_snapshot* k = spare_.back();
            spare_.pop_back();
            check(apbf_snapshot_save(k, s_.handle()), "apbf_snapshot_save");
            return k;
        }
        apbf_snapshot* k = nullptr;
        check(apbf_snapshot_create(s_.handle(), &k), "apbf_snapshot_create");
        return k;
    }
    void drop_front(size_t keep) {
        while (keys_.size() > keep) {
            spare_.push_back(keys_.front());
            keys_.pop_front();
        }
    }
    void load() { check(apbf_snapshot_restore(keys_[sparse_history_index()], s_.handle()), "apbf_snapshot_restore"); }

    scene& s_;
    uint32_t max_keyframes_ = 5u, interval_ = 120u, current_ = 0u, present_ = 0u;
    bool enabled_ = false;
    std::deque<apbf_snapshot*> keys_;
    std::vector<apbf_snapshot*> spare_;
    std::deque<float> dts_;
};

/* save_particle_info.h:8-23 (3D: the SVG dump writes measurements_<id>.txt only) */
class save_particle_info {
public:
    save_particle_info& set_data(scene* s) {
        s_ = s;
        return *this;
    }
    void apply(const char* folder = "particle_data") {
        check(apbf_scene_save_particle_info(s_->handle(), folder), "save_particle_info");
    }
    void save_as_svg(uint32_t svg_id, float step_ms = 0.0f, float neighborhood_ms = 0.0f, float solver_ms = 0.0f,
                     const char* folder = "particle_svg") {
        check(apbf_scene_save_measurements(s_->handle(), folder, svg_id, step_ms, neighborhood_ms, solver_ms),
              "save_measurements");
    }

private:
    scene* s_ = nullptr;
};

/* uint_to_float_with_indexed_lower_bound (pool.cpp:77-80): kernel width from
 * the boundary distance. */
inline void kernel_width_from_boundary_distance(scene& s) { s.stage(APBF_STAGE_KERNEL_WIDTH_BD, 0.0f); }

}  // namespace apbf_pbd

#endif
