/* C++ host interface with the shape of the reference's subsystem classes
 * (namespace pbd in the source/ headers). Each class binds with set_data either to
 * the lists it works on (apbf_lists.hpp: fluid, transfers, neighbors - the
 * reference's set_data(fluid*, neighbors*, ...) signatures) and then runs its
 * stage operator on those lists, or to a whole scene and then runs the
 * scene's stage. A caller written against velocity_handling /
 * particle_transfer / neighborhood_green / spread_kernel_width /
 * box_collision / incompressibility / update_transfers keeps its call
 * sequence (pool.cpp:67-106). Header-only over include/apbf.h; errors become
 * apbf_pbd::error exceptions (the reference asserts). */
#ifndef APBF_HPP
#define APBF_HPP

#include <algorithm>
#include <deque>
#include <stdexcept>
#include <string>
#include <vector>

#include "apbf.h"
#include "apbf_lists.hpp"

namespace apbf_pbd {

struct vec3 {
    float x, y, z;
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

/* The temporaries of the stage operators (apbf_context) for the capacities,
 * settings and hash grid of a scene's parameters. The list-bound subsystems
 * below run on the context made current last (the reference's subsystems
 * share one global shader_provider and settings block in the same way,
 * shader_provider.cpp:5-8, settings.cpp:70-91). */
class context {
public:
    explicit context(const apbf_scene_params& p) : params_(p) {
        check(apbf_context_create(&p, &h_), "apbf_context_create");
        current() = this;
    }
    ~context() {
        if (current() == this) current() = nullptr;
        if (h_) apbf_context_destroy(h_);
    }
    context(const context&) = delete;
    context& operator=(const context&) = delete;
    apbf_context* handle() const { return h_; }
    const apbf_scene_params& params() const { return params_; }
    void make_current() { current() = this; }
    static context*& current() {
        static context* c = nullptr;
        return c;
    }
    static context& get() {
        if (!current()) throw error("no current apbf_pbd::context", APBF_ERR_INVALID_ARGUMENT);
        return *current();
    }

private:
    apbf_scene_params params_;
    apbf_context* h_ = nullptr;
};

/* The owner link of pool.cpp:18-19 (the transfer lists share the particles'
 * hidden list): reorderings and deletions of the particles are applied to the
 * transfer endpoints too. */
inline std::vector<std::pair<fluid*, transfers*>>& transfer_owners() {
    static std::vector<std::pair<fluid*, transfers*>> v;
    return v;
}
inline void share_hidden_data(transfers& t, fluid& f) {
    for (auto& e : transfer_owners())
        if (e.second == &t) {
            e.first = &f;
            return;
        }
    transfer_owners().push_back({&f, &t});
}
inline transfers* transfers_of(fluid* f) {
    for (auto& e : transfer_owners())
        if (e.first == f) return e.second;
    return nullptr;
}

/* velocity_handling.h:7-18 */
class velocity_handling {
public:
    velocity_handling& set_data(scene* s) {
        s_ = s;
        f_ = nullptr;
        return *this;
    }
    velocity_handling& set_data(fluid* particles) {
        f_ = particles;
        s_ = nullptr;
        return *this;
    }
    velocity_handling& set_acceleration(const vec3& a = vec3{0, 0, 0}) {
        acc_ = a;
        if (s_) {
            const float g[3] = {a.x, a.y, a.z};
            check(apbf_scene_set_gravity(s_->handle(), g), "set_acceleration");
        }
        return *this;
    }
    void apply(float delta_time) {
        if (s_) return s_->stage(APBF_STAGE_VELOCITY, delta_time);
        const float g[3] = {acc_.x, acc_.y, acc_.z};
        apbf_particle_lists p = f_->lists();
        check(apbf_op_velocity_handling(context::get().handle(), &p, &last_delta_time_, delta_time, g, list_stream()),
              "velocity_handling");
    }
    float last_delta_time() const { return last_delta_time_; }
    velocity_handling& set_last_delta_time(float t) {
        last_delta_time_ = t;
        return *this;
    }

private:
    scene* s_ = nullptr;
    fluid* f_ = nullptr;
    vec3 acc_{0, 0, 0};
    float last_delta_time_ = 1.0f;  // velocity_handling.h:17
};

/* particle_transfer.h:7-16 */
class particle_transfer {
public:
    particle_transfer& set_data(scene* s) {
        s_ = s;
        f_ = nullptr;
        return *this;
    }
    particle_transfer& set_data(fluid* f, transfers* t) {
        f_ = f;
        t_ = t;
        s_ = nullptr;
        return *this;
    }
    void apply(float delta_time) {
        if (s_) return s_->stage(APBF_STAGE_PARTICLE_TRANSFER, delta_time);
        apbf_particle_lists p = f_->lists();
        apbf_transfer_lists t = t_->lists();
        check(apbf_op_particle_transfer(context::get().handle(), &p, &t, delta_time, list_stream()), "particle_transfer");
        f_->sync_lengths();
        t_->sync_lengths();
    }

private:
    scene* s_ = nullptr;
    fluid* f_ = nullptr;
    transfers* t_ = nullptr;
};

/* neighborhood_green.h:8-24. Bound to lists, the range of each particle is
 * its kernel width (the reference passes the fluid's kernel_width list,
 * pool.cpp:46) and the particles' transfer owner follows the reordering. */
class neighborhood_green {
public:
    neighborhood_green& set_data(scene* s) {
        s_ = s;
        f_ = nullptr;
        return *this;
    }
    neighborhood_green& set_data(fluid* particles, neighbors* nb) {
        f_ = particles;
        nb_ = nb;
        s_ = nullptr;
        return *this;
    }
    neighborhood_green& set_range_scale(float scale) {
        scale_ = scale;
        if (s_) check(apbf_scene_set_range_scale(s_->handle(), scale), "set_range_scale");
        return *this;
    }
    neighborhood_green& set_position_range(const vec3& mn, const vec3& mx, uint32_t resolution_log2) {
        mn_ = mn;
        mx_ = mx;
        res_ = resolution_log2;
        ranged_ = true;
        if (s_) {
            const float a[3] = {mn.x, mn.y, mn.z}, b[3] = {mx.x, mx.y, mx.z};
            check(apbf_scene_set_position_range(s_->handle(), a, b, resolution_log2), "set_position_range");
        }
        return *this;
    }
    void apply() {
        if (s_) return s_->stage(APBF_STAGE_NEIGHBORHOOD, 0.0f);
        const apbf_scene_params& prm = context::get().params();
        const float a[3] = {ranged_ ? mn_.x : prm.hash_min[0], ranged_ ? mn_.y : prm.hash_min[1],
                            ranged_ ? mn_.z : prm.hash_min[2]};
        const float b[3] = {ranged_ ? mx_.x : prm.hash_max[0], ranged_ ? mx_.y : prm.hash_max[1],
                            ranged_ ? mx_.z : prm.hash_max[2]};
        apbf_particle_lists p = f_->lists();
        apbf_neighbor_lists n = nb_->lists();
        transfers* owner = transfers_of(f_);
        apbf_transfer_lists t{};
        if (owner) t = owner->lists();
        check(apbf_op_neighborhood(context::get().handle(), &p, owner ? &t : nullptr, &n, method_, scale_, a, b,
                                   ranged_ ? res_ : prm.hash_res_log2, list_stream()),
              "neighborhood");
    }

protected:
    int32_t method_ = 0;
    scene* s_ = nullptr;
    fluid* f_ = nullptr;
    neighbors* nb_ = nullptr;
    float scale_ = 1.0f;
    vec3 mn_{0, 0, 0}, mx_{1, 1, 1};
    uint32_t res_ = 4;
    bool ranged_ = false;
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
    neighborhood_binary_search& set_data(fluid* particles, neighbors* nb) {
        neighborhood_green::set_data(particles, nb);
        method_ = 1;
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
    neighborhood_brute_force& set_data(fluid* particles, neighbors* nb) {
        neighborhood_green::set_data(particles, nb);
        method_ = 2;
        return *this;
    }
};

/* spread_kernel_width.h:7-15 */
class spread_kernel_width {
public:
    spread_kernel_width& set_data(scene* s) {
        s_ = s;
        f_ = nullptr;
        return *this;
    }
    spread_kernel_width& set_data(fluid* f, neighbors* nb) {
        f_ = f;
        nb_ = nb;
        s_ = nullptr;
        return *this;
    }
    void apply() {
        if (s_) return s_->stage(APBF_STAGE_SPREAD, 0.0f);
        apbf_particle_lists p = f_->lists();
        apbf_neighbor_lists n = nb_->lists();
        check(apbf_op_spread_kernel_width(context::get().handle(), &p, &n, list_stream()), "spread_kernel_width");
    }

private:
    scene* s_ = nullptr;
    fluid* f_ = nullptr;
    neighbors* nb_ = nullptr;
};

/* box_collision.h:7-16: boxes as [min.xyz, pad] / [max.xyz, pad] per box
 * (the reference's gpu_list<16> box lists, host side here). */
class box_collision {
public:
    box_collision& set_data(scene* s) {
        s_ = s;
        f_ = nullptr;
        return *this;
    }
    box_collision& set_data(fluid* particles, const std::vector<float>* box_min4, const std::vector<float>* box_max4) {
        f_ = particles;
        mn_ = box_min4;
        mx_ = box_max4;
        s_ = nullptr;
        return *this;
    }
    void apply() {
        if (s_) return s_->stage(APBF_STAGE_BOX, 0.0f);
        apbf_particle_lists p = f_->lists();
        const uint32_t nb = (uint32_t)(mn_->size() / 4);
        check(apbf_op_box_collision(context::get().handle(), &p, nb, mn_->data(), mx_->data(), list_stream()),
              "box_collision");
    }

private:
    scene* s_ = nullptr;
    fluid* f_ = nullptr;
    const std::vector<float>*mn_ = nullptr, *mx_ = nullptr;
};

/* sphere_collision.h: set_sphere(center, radius, hollow), apply() */
class sphere_collision {
public:
    sphere_collision& set_data(scene* s) {
        s_ = s;
        f_ = nullptr;
        return *this;
    }
    sphere_collision& set_data(fluid* particles) {
        f_ = particles;
        s_ = nullptr;
        return *this;
    }
    sphere_collision& set_sphere(const vec3& center, float radius, bool hollow) {
        c_ = center;
        r_ = radius;
        hollow_ = hollow;
        if (s_) {
            const float c[3] = {center.x, center.y, center.z};
            check(apbf_scene_set_sphere(s_->handle(), c, radius, hollow ? 1 : 0, 1), "set_sphere");
        }
        return *this;
    }
    void apply() {
        if (s_) return s_->stage(APBF_STAGE_SPHERE, 0.0f);
        apbf_particle_lists p = f_->lists();
        const float c[3] = {c_.x, c_.y, c_.z};
        check(apbf_op_sphere_collision(context::get().handle(), &p, c, r_, hollow_ ? 1 : 0, list_stream()),
              "sphere_collision");
    }

private:
    scene* s_ = nullptr;
    fluid* f_ = nullptr;
    vec3 c_{0, 0, 0};
    float r_ = 0.0f;
    bool hollow_ = true;
};

/* incompressibility.h:7-18: one solver iteration per apply() */
class incompressibility {
public:
    incompressibility& set_data(scene* s) {
        s_ = s;
        f_ = nullptr;
        return *this;
    }
    incompressibility& set_data(fluid* f, neighbors* nb) {
        f_ = f;
        nb_ = nb;
        s_ = nullptr;
        return *this;
    }
    void apply() {
        if (s_) return s_->stage(APBF_STAGE_INCOMPRESSIBILITY, 0.0f);
        apbf_particle_lists p = f_->lists();
        apbf_neighbor_lists n = nb_->lists();
        check(apbf_op_incompressibility(context::get().handle(), &p, &n, list_stream()), "incompressibility");
    }

private:
    scene* s_ = nullptr;
    fluid* f_ = nullptr;
    neighbors* nb_ = nullptr;
};

/* update_transfers.h:7-16 */
class update_transfers {
public:
    update_transfers& set_data(scene* s) {
        s_ = s;
        f_ = nullptr;
        return *this;
    }
    update_transfers& set_data(fluid* f, neighbors* nb, transfers* t) {
        f_ = f;
        nb_ = nb;
        t_ = t;
        s_ = nullptr;
        return *this;
    }
    void apply() {
        if (s_) return s_->stage(APBF_STAGE_UPDATE_TRANSFERS, 0.0f);
        apbf_particle_lists p = f_->lists();
        apbf_transfer_lists t = t_->lists();
        apbf_neighbor_lists n = nb_->lists();
        check(apbf_op_update_transfers(context::get().handle(), &p, &t, &n, list_stream()), "update_transfers");
        f_->sync_lengths();
        t_->sync_lengths();
    }

private:
    scene* s_ = nullptr;
    fluid* f_ = nullptr;
    neighbors* nb_ = nullptr;
    transfers* t_ = nullptr;
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
    /* the time step recorded for the current step (the per-substep history
     * value that stepping through the past restores) */
    float current_dt() const { return dts_[history_index()]; }

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

/* save_particle_info.h:8-23. save_as_svg writes measurements_<id>.txt and, in
 * a planar scene, particles_<id>.svg (3D: the measurements only). The
 * reference's globals it reads (settings::color, settings::renderBoxes, the
 * measurement timers) come in through svg_options. */
struct svg_options {
    int32_t color = 0;          /* settings::color (settings.cpp:29) */
    bool render_boxes = true;   /* settings::renderBoxes (settings.cpp:15) */
    float step_ms = 0.0f, neighborhood_ms = 0.0f, solver_ms = 0.0f;
    const char* folder = "particle_svg";  /* SVG_FOLDER_NAME (cpu_gpu_shared_config.h:22) */
};
class save_particle_info {
public:
    save_particle_info& set_data(scene* s) {
        s_ = s;
        return *this;
    }
    void apply(const char* folder = "particle_data") {
        check(apbf_scene_save_particle_info(s_->handle(), folder), "save_particle_info");
    }
    void save_as_svg(uint32_t svg_id, const float view_box_min[2], const float view_box_max[2],
                     float render_scale = 1.0f, float max_expected_boundary_dist = 100.0f,
                     const svg_options& o = svg_options()) {
        check(apbf_scene_save_measurements(s_->handle(), o.folder, svg_id, o.step_ms, o.neighborhood_ms,
                                           o.solver_ms),
              "save_measurements");
        check(apbf_scene_save_svg(s_->handle(), o.folder, svg_id, view_box_min, view_box_max, render_scale,
                                  max_expected_boundary_dist, o.color, o.render_boxes ? 1 : 0),
              "save_svg");
    }

private:
    scene* s_ = nullptr;
};

/* uint_to_float_with_indexed_lower_bound (pool.cpp:77-80): kernel width from
 * the boundary distance. */
inline void kernel_width_from_boundary_distance(scene& s) { s.stage(APBF_STAGE_KERNEL_WIDTH_BD, 0.0f); }
inline void kernel_width_from_boundary_distance(fluid& f) {
    apbf_particle_lists p = f.lists();
    check(apbf_op_kernel_width_from_boundary_distance(context::get().handle(), &p, list_stream()),
          "kernel_width_from_boundary_distance");
}

}  // namespace apbf_pbd

#endif
