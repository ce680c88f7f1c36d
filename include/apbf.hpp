/* C++ host interface with the shape of the reference's subsystem classes
 * (namespace pbd in source/*.h): each class binds to a scene with set_data
 * and runs its stage with apply(), so a caller written against
 * velocity_handling / particle_transfer / neighborhood_green /
 * spread_kernel_width / box_collision / incompressibility / update_transfers
 * keeps its call sequence (pool.cpp:67-106). Header-only over include/apbf.h;
 * errors become apbf_pbd::error exceptions (the reference asserts). */
#ifndef APBF_HPP
#define APBF_HPP

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

/* velocity_handling.h:9-20 */
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

/* uint_to_float_with_indexed_lower_bound (pool.cpp:77-80): kernel width from
 * the boundary distance. */
inline void kernel_width_from_boundary_distance(scene& s) { s.stage(APBF_STAGE_KERNEL_WIDTH_BD, 0.0f); }

}  // namespace apbf_pbd

#endif
