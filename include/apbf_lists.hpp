/* Host-side lists of include/apbf.h with the shape of the reference's list
 * containers: gpu_list<Stride> (source/gpu_list.h:11-57) over a shared
 * device buffer with a device-resident length (gpu_list_data.h:8-31), and the
 * list bundles of list_definitions.h:9-20 that the subsystems bind to.
 * Header-only over the C ABI; errors become apbf_pbd::error exceptions. */
#ifndef APBF_LISTS_HPP
#define APBF_LISTS_HPP

#include <algorithm>
#include <memory>
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

/* The stream every list operation is recorded on (the reference records all
 * of them into one command buffer, shader_provider.cpp:5-8); 0 = null stream. */
inline apbf_stream& list_stream() {
    static apbf_stream s = nullptr;
    return s;
}

/* gpu_list_data.h:8-31: a device buffer and its device length word. */
struct gpu_list_data {
    void* buffer = nullptr;
    uint32_t* length = nullptr;
    size_t capacity = 0;  // elements
    gpu_list_data(size_t cap, size_t stride) : capacity(cap) {
        check(apbf_device_alloc(std::max<size_t>(cap * stride, 4), &buffer), "apbf_device_alloc");
        void* l = nullptr;
        check(apbf_device_alloc(4, &l), "apbf_device_alloc");
        length = static_cast<uint32_t*>(l);
    }
    ~gpu_list_data() {
        apbf_device_free(buffer);
        apbf_device_free(length);
    }
    gpu_list_data(const gpu_list_data&) = delete;
    gpu_list_data& operator=(const gpu_list_data&) = delete;
};

/* gpu_list.h:11-57. Copies share the buffer (lazy copy); write() gives this
 * list a buffer of its own, sized to the requested length, when it is shared
 * or too small (gpu_list.h:172-177, 239-261). */
template <size_t Stride>
class gpu_list {
    static_assert(Stride % 4 == 0, "elements are whole 32-bit words");

public:
    gpu_list() = default;

    bool empty() const { return !d_; }
    uint32_t* length() const { return d_ ? d_->length : nullptr; }
    void* buffer() const { return d_ ? d_->buffer : nullptr; }
    size_t requested_length() const { return requested_; }
    gpu_list& request_length(size_t n) {
        requested_ = n;
        return *this;
    }
    /* gpu_list.h:81-89 */
    gpu_list& set_length(size_t n) {
        if (requested_ < n) requested_ = n;
        const uint32_t v = (uint32_t)n;
        check(apbf_copy(write().length(), &v, 4, list_stream()), "set_length");
        check(apbf_stream_synchronize(list_stream()), "set_length");  // the host word goes out of scope
        return *this;
    }
    /* gpu_list.h:91-96 */
    gpu_list& set_length(const uint32_t* d_len) {
        check(apbf_copy(write().length(), d_len, 4, list_stream()), "set_length");
        return *this;
    }
    gpu_list& write() {
        prepare(requested_, true);
        return *this;
    }
    /* host copy of the list (gpu_list.h:151-167; synchronises) */
    template <class T>
    std::vector<T> read(bool beyond_length = false) const {
        std::vector<T> out;
        if (empty()) return out;
        out.resize(d_->capacity * Stride / sizeof(T));
        check(apbf_copy(out.data(), d_->buffer, out.size() * sizeof(T), list_stream()), "read");
        uint32_t n = 0;
        check(apbf_copy(&n, d_->length, 4, list_stream()), "read");
        check(apbf_stream_synchronize(list_stream()), "read");
        if (!beyond_length) out.resize(std::min(out.size(), (size_t)n * Stride / sizeof(T)));
        return out;
    }
    /* gpu_list.h:126-135: this[i] = old[edit[i]], length = edit length */
    void apply_edit(const gpu_list<4>& edit) {
        std::shared_ptr<gpu_list_data> old = d_;
        const size_t need = std::max(requested_, edit.requested_length());
        d_.reset();
        prepare(need, false);
        check(apbf_copy_scattered_read(old ? old->buffer : nullptr, d_->buffer, static_cast<const uint32_t*>(edit.buffer()),
                                       edit.length(), (uint32_t)std::min(need, edit_capacity(edit)), (uint32_t)Stride,
                                       list_stream()),
              "apply_edit");
        check(apbf_copy(d_->length, edit.length(), 4, list_stream()), "apply_edit");
    }
    /* gpu_list.h:214-227: append, clamped to this list's buffer */
    gpu_list& operator+=(const gpu_list& rhs) {
        if (!d_) {
            d_ = rhs.d_;
            return *this;
        }
        if (!rhs.d_) return *this;
        write();
        check(apbf_append_list(d_->buffer, (uint32_t)d_->capacity, d_->length, rhs.d_->buffer, rhs.d_->length,
                               (uint32_t)Stride, d_->length, (uint32_t)rhs.d_->capacity, list_stream()),
              "append_list");
        return *this;
    }
    /* gpu_list.h:229-236 */
    gpu_list operator+(const gpu_list& rhs) const {
        gpu_list r = *this;
        r.requested_ = std::max(requested_, rhs.requested_);
        r += rhs;
        return r;
    }
    /* gpu_list.h:179-193 (Stride 4): ascending sort of the values as u32 */
    void sort(uint32_t value_upper_bound = 0xFFFFFFFFu) {
        static_assert(Stride == 4, "sort is defined for 4-byte elements");
        const size_t cap = std::max<size_t>(requested_, d_ ? d_->capacity : 0);
        gpu_list<4> idx, out_keys, out_idx;
        idx.request_length(cap).write();
        out_keys.request_length(cap).write();
        out_idx.request_length(cap).write();
        check(apbf_write_sequence(static_cast<uint32_t*>(idx.buffer()), length(), (uint32_t)cap, 0u, 1u, list_stream()),
              "sort");
        void* ws = nullptr;
        const size_t wsb = apbf_radix_sort_workspace_size((uint32_t)cap);
        check(apbf_device_alloc(wsb, &ws), "sort");
        const apbf_status st = apbf_radix_sort(static_cast<const uint32_t*>(buffer()), static_cast<const uint32_t*>(idx.buffer()),
                                               length(), (uint32_t)cap, value_upper_bound,
                                               static_cast<uint32_t*>(out_keys.buffer()),
                                               static_cast<uint32_t*>(out_idx.buffer()), ws, wsb, list_stream());
        check(apbf_stream_synchronize(list_stream()), "sort");
        apbf_device_free(ws);
        check(st, "sort");
        check(apbf_copy(out_keys.length(), length(), 4, list_stream()), "sort");
        d_ = out_keys.d_;
    }
    /* to_gpu_list of the reference's tests (test.cpp:13-24) */
    template <class T>
    static gpu_list from_host(const std::vector<T>& v, size_t requested = 0) {
        gpu_list l;
        const size_t n = v.size() * sizeof(T) / Stride;
        l.request_length(std::max(requested, n)).write();
        if (!v.empty()) check(apbf_copy(l.buffer(), v.data(), v.size() * sizeof(T), list_stream()), "from_host");
        l.set_length(n);
        return l;
    }
    size_t capacity() const { return d_ ? d_->capacity : 0; }

private:
    template <size_t>
    friend class gpu_list;
    static size_t edit_capacity(const gpu_list<4>& e) { return e.d_ ? e.d_->capacity : 0; }
    void prepare(size_t needed, bool keep_content) {
        const size_t have = d_ ? d_->capacity : 0;
        if (have == 0 && needed == 0) return;
        if (d_ && d_.use_count() == 1 && have >= needed) return;
        std::shared_ptr<gpu_list_data> old = d_;
        d_ = std::make_shared<gpu_list_data>(std::max(needed, have), Stride);
        if (old) {
            check(apbf_copy(d_->length, old->length, 4, list_stream()), "write");
            if (keep_content)
                check(apbf_copy(d_->buffer, old->buffer, std::min(have, d_->capacity) * Stride, list_stream()), "write");
        } else {
            const uint32_t zero = 0;
            check(apbf_copy(d_->length, &zero, 4, list_stream()), "write");
            check(apbf_stream_synchronize(list_stream()), "write");
        }
    }
    std::shared_ptr<gpu_list_data> d_;
    size_t requested_ = 0;
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

/* list_definitions.h:9-13: the hidden particles (position, velocity,
 * inverse_mass, radius, pos_backup, transferring) and the fluid attributes
 * (target_radius, kernel_width, boundariness, boundary_distance), stored with
 * fluid id == hidden slot. All members have the capacity of the particle
 * list; the particle count is the position list's length word. */
struct fluid {
    gpu_list<16> position, velocity, pos_backup;
    gpu_list<4> inverse_mass, radius, transferring, target_radius, kernel_width, boundariness, boundary_distance;
    uint32_t capacity;

    explicit fluid(uint32_t cap) : capacity(cap) {
        for_each([&](auto& l) { l.request_length(cap).write(); });
    }
    uint32_t* length() const { return position.length(); }
    apbf_particle_lists lists() {
        apbf_particle_lists p;
        p.position = static_cast<int32_t*>(position.write().buffer());
        p.velocity = static_cast<float*>(velocity.write().buffer());
        p.inverse_mass = static_cast<float*>(inverse_mass.write().buffer());
        p.radius = static_cast<float*>(radius.write().buffer());
        p.pos_backup = static_cast<int32_t*>(pos_backup.write().buffer());
        p.transferring = static_cast<uint32_t*>(transferring.write().buffer());
        p.target_radius = static_cast<float*>(target_radius.write().buffer());
        p.kernel_width = static_cast<float*>(kernel_width.write().buffer());
        p.boundariness = static_cast<float*>(boundariness.write().buffer());
        p.boundary_distance = static_cast<uint32_t*>(boundary_distance.write().buffer());
        p.d_length = position.length();
        p.capacity = capacity;
        return p;
    }
    /* the other members take the position list's length (uninterleaved_list
     * keeps its members at one length, uninterleaved_list.h:7-83) */
    void sync_lengths() {
        for_each([&](auto& l) {
            if (l.length() != position.length()) check(apbf_copy(l.length(), position.length(), 4, list_stream()), "length");
        });
    }
    void upload(const state& s) {
        const uint32_t n = s.size();
        auto put = [&](auto& l, const auto& v) {
            if (!v.empty()) check(apbf_copy(l.write().buffer(), v.data(), v.size() * sizeof(v[0]), list_stream()), "upload");
        };
        put(position, s.position);
        put(velocity, s.velocity);
        put(pos_backup, s.pos_backup);
        put(inverse_mass, s.inverse_mass);
        put(radius, s.radius);
        put(transferring, s.transferring);
        put(target_radius, s.target_radius);
        put(kernel_width, s.kernel_width);
        put(boundariness, s.boundariness);
        put(boundary_distance, s.boundary_distance);
        check(apbf_stream_synchronize(list_stream()), "upload");
        position.set_length(n);
        sync_lengths();
    }
    void download(state& s) const {
        s.position = position.read<int32_t>();
        s.velocity = velocity.read<float>();
        s.pos_backup = pos_backup.read<int32_t>();
        const size_t n = s.position.size() / 4;
        auto get = [&](const auto& l, auto& v) {
            using T = typename std::decay_t<decltype(v)>::value_type;
            v = l.template read<T>(true);
            v.resize(n);
        };
        get(inverse_mass, s.inverse_mass);
        get(radius, s.radius);
        get(transferring, s.transferring);
        get(target_radius, s.target_radius);
        get(kernel_width, s.kernel_width);
        get(boundariness, s.boundariness);
        get(boundary_distance, s.boundary_distance);
    }

private:
    template <class F>
    void for_each(F f) {
        f(position), f(velocity), f(pos_backup), f(inverse_mass), f(radius), f(transferring), f(target_radius),
            f(kernel_width), f(boundariness), f(boundary_distance);
    }
};

/* list_definitions.h:15-16: transfer source, target (particle slots) and time left. */
struct transfers {
    gpu_list<4> source, target, time_left;
    uint32_t capacity;
    explicit transfers(uint32_t cap) : capacity(cap) {
        source.request_length(cap).write();
        target.request_length(cap).write();
        time_left.request_length(cap).write();
    }
    apbf_transfer_lists lists() {
        apbf_transfer_lists t;
        t.source = static_cast<uint32_t*>(source.write().buffer());
        t.target = static_cast<uint32_t*>(target.write().buffer());
        t.time_left = static_cast<float*>(time_left.write().buffer());
        t.d_length = source.length();
        t.capacity = capacity;
        return t;
    }
    void sync_lengths() {
        check(apbf_copy(target.length(), source.length(), 4, list_stream()), "length");
        check(apbf_copy(time_left.length(), source.length(), 4, list_stream()), "length");
    }
    void upload(const state& s) {
        const size_t nt = s.t_source.size();
        if (nt) {
            check(apbf_copy(source.write().buffer(), s.t_source.data(), nt * 4, list_stream()), "upload");
            check(apbf_copy(target.write().buffer(), s.t_target.data(), nt * 4, list_stream()), "upload");
            check(apbf_copy(time_left.write().buffer(), s.t_time_left.data(), nt * 4, list_stream()), "upload");
            check(apbf_stream_synchronize(list_stream()), "upload");
        }
        source.set_length(nt);
        sync_lengths();
    }
    void download(state& s) const {
        s.t_source = source.read<uint32_t>();
        s.t_target = target.read<uint32_t>(true);
        s.t_time_left = time_left.read<float>(true);
        s.t_target.resize(s.t_source.size());
        s.t_time_left.resize(s.t_source.size());
    }
};

/* list_definitions.h:19 (neighbors = gpu_list<8> of uvec2 pairs), held as
 * CSR rows; pairs() gives the reference's layout. */
struct neighbors {
    gpu_list<4> row_offset, ids, rows;  // rows: one word, the particles the rows cover
    uint32_t particle_capacity, capacity;
    neighbors(uint32_t particles, uint32_t pair_capacity) : particle_capacity(particles), capacity(pair_capacity) {
        row_offset.request_length(particles + 1).write();
        ids.request_length(pair_capacity).write();
        rows.request_length(1).write();
        ids.set_length(size_t(0));
        rows.set_length(1);
        const uint32_t zero = 0;
        check(apbf_copy(rows.buffer(), &zero, 4, list_stream()), "neighbors");
        check(apbf_stream_synchronize(list_stream()), "neighbors");
    }
    apbf_neighbor_lists lists() {
        apbf_neighbor_lists n;
        n.row_offset = static_cast<uint32_t*>(row_offset.write().buffer());
        n.neighbors = static_cast<uint32_t*>(ids.write().buffer());
        n.d_length = ids.length();
        n.d_rows = static_cast<uint32_t*>(rows.write().buffer());
        n.capacity = capacity;
        return n;
    }
    /* the uvec2 pair list (apbf_neighbor_pairs) */
    gpu_list<8> pairs() {
        gpu_list<8> out;
        out.request_length(capacity).write();
        apbf_neighbor_lists n = lists();
        check(apbf_neighbor_pairs(n.row_offset, n.neighbors, n.d_rows, n.d_length, particle_capacity, capacity,
                                  static_cast<uint32_t*>(out.buffer()), list_stream()),
              "neighbor_pairs");
        out.set_length(ids.length());
        return out;
    }
};

}  // namespace apbf_pbd

#endif
