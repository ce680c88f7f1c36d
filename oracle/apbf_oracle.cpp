// Test infrastructure only: scalar CPU oracle of APBF's per-substep pipeline.
//
// Each function below restates one reference compute shader (cited by
// file:line) as a per-invocation body executed for gl_GlobalInvocationID.x =
// 0, 1, 2, ... in order, or one host routine of the reference (cited likewise).
// Sequential execution is a legal schedule of the reference, so atomics resolve
// in invocation order and the pair list, split list and merge claims are those
// of that schedule. Where the reference reads undefined data the choice made
// here is stated at the use site.
//
// Build: g++ -O2 -ffp-contract=off -fPIC -shared (see oracle/Makefile).
//
// ORACLE_MUTATION (default 0) selects one deliberately wrong variant of a
// single expression, numbered at its use site. Only tests/test_oracle_pins.py
// builds such variants (into a temporary directory), to show that each pin of
// the oracle fails on a plausible mistake. liboracle.so is always built with 0.
#include "apbf_oracle.h"
#include "oracle_math.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace om;

#ifndef ORACLE_MUTATION
#define ORACLE_MUTATION 0
#endif
#define MUT(k) (ORACLE_MUTATION == (k))

namespace {

// shaders/cpu_gpu_shared_config.h:2-17 (DIMENSIONS fixed to 3 for this build)
constexpr float POS_RES = 262144.0f;
constexpr float KW_RES = 262144.0f;
constexpr float INCOMP_RES = 262144.0f;
constexpr float KERNEL_SCALE = 4.0f;
constexpr float KW_PROPAGATION = 0.5f;
constexpr float PI_F = 3.14159265f;  // kernels.glsl:1
constexpr float INCOMP_EPSILON = 0.01f;  // incompressibility_2.comp:6
// initialize_split_particles.comp:26 writes the GLSL integer expression 1/0,
// which glslang constant-folds to INT_MAX; stored into a float buffer that is
// 2147483647 -> 2^31.
constexpr float SPLIT_INV_MASS = 2147483648.0f;

struct I4 { int32_t x, y, z, w; };
struct F4 { float x, y, z, w; };
struct U2 { uint32_t a, b; };
struct V3 { float x, y, z; };

inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 mul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 muls(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 divs(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
inline V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline float length(V3 a) { return sqrtf(dot(a, a)); }
inline V3 normalize(V3 a) { return muls(a, 1.0f / length(a)); }
inline float fract(float x) { return x - floorf(x); }
inline float sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
// ivec3 arithmetic wraps (GLSL), done in uint32 here.
inline int32_t iadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
inline int32_t isub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
inline V3 vec_of(int32_t x, int32_t y, int32_t z) { return v3((float)x, (float)y, (float)z); }
inline V3 pos_world(const I4& p) { return divs(vec_of(p.x, p.y, p.z), POS_RES); }
inline V3 idiff(const I4& a, const I4& b) {  // vec3(a.xyz - b.xyz)
    return vec_of(isub(a.x, b.x), isub(a.y, b.y), isub(a.z, b.z));
}

// ---------------------------------------------------------------- kernels.glsl
float poly6_kernel_height(V3 r, float h) {  // kernels.glsl:3-9
    float r2 = dot(r, r), h2 = h * h;
    if (r2 > h2) return 0.0f;
    return 315.0f * pow_f(h2 - r2, 3.0f) / (64.0f * PI_F * pow_f(h, 9.0f));
}
V3 spiky_kernel_gradient(V3 r, float h) {  // kernels.glsl:11-16
    float dist = length(r);
    if (dist > h || dist < 0.0001f) return v3(0, 0, 0);
    float s = -45.0f * pow_f(h - dist, 2.0f) / (PI_F * pow_f(h, 6.0f));
    return muls(divs(r, dist), s);
}
float cubic_kernel_height(V3 r, float h) {  // kernels.glsl:18-31
    float dist = length(r);
    if (dist > h) return 0.0f;
    float q = dist / h;
    float k = 8.0f / (PI_F * h * h * h);
    if (q <= 0.5f) {
        float q2 = q * q, q3 = q * q2;
        return k * (6.0f * q3 - 6.0f * q2 + 1.0f);
    }
    return k * (2.0f * pow_f(1.0f - q, 3.0f));
}
V3 cubic_kernel_gradient(V3 r, float h) {  // kernels.glsl:33-46
    float dist = length(r);
    if (dist > h || dist < 0.0001f) return v3(0, 0, 0);
    float q = dist / h;
    V3 gradq = muls(r, 1.0f / (dist * h));
    float l = 48.0f / (PI_F * h * h * h);
    if (q <= 0.5f) return muls(gradq, l * q * (3.0f * q - 2.0f));
    float factor = 1.0f - q;
    return muls(gradq, l * (-factor * factor));
}
// D is DIMENSIONS (cpu_gpu_shared_config.h:2), 2 or 3, as a float.
float cone_kernel_height(V3 r, float h, float D) {  // kernels.glsl:48-52
    float height = 3.0f / (PI_F * pow_f(h, D));
    return gmax(0.0f, (1.0f - length(r) / h) * height);
}
V3 cone_kernel_gradient(V3 r, float h, float D) {  // kernels.glsl:54-60
    float steep = 3.0f / (PI_F * pow_f(h, D + 1.0f));
    float dist = length(r);
    if (dist > h || dist < 0.0001f) return v3(0, 0, 0);
    return muls(neg(r), steep / dist);
}
float quadratic_spike_a(float h, float D) {  // kernels.glsl:64-68, 76-80
    return D == 3.0f ? 15.0f / (2.0f * PI_F * pow_f(h, 5.0f)) : 6.0f / (PI_F * pow_f(h, 4.0f));
}
float quadratic_spike_kernel_height(V3 r, float h, float D) {  // kernels.glsl:62-70
    float a = quadratic_spike_a(h, D);
    return a * pow_f(gmin(0.0f, length(r) - h), 2.0f);
}
V3 quadratic_spike_kernel_gradient(V3 r, float h, float D) {  // kernels.glsl:72-82
    float dist = length(r);
    if (dist > h || dist < 0.0001f) return v3(0, 0, 0);
    float a = quadratic_spike_a(h, D);
    return muls(r, -2.0f * a * gmax(0.0f, h - dist) / dist);
}

// Gaussian kernel constants derived from the kernel width. kernel_height /
// kernel_gradient (kernels.glsl:99-123) recompute them from the kernel width on
// every call; they depend on the width only, so the values are identical.
struct Gauss {
    float ivd;   // invDoubleVariance = pow(height, 2/D) * PI
    float norm;  // pow(invDoubleVarianceWithoutPi, D/2)
};
inline Gauss gauss_of_width(float kw, float D) {
    float height = 0.6f / pow_f(kw / 2.0f, D);       // kernels.glsl:104,117
    float ivdwp = pow_f(height, 2.0f / D);           // kernels.glsl:86
    return Gauss{ivdwp * PI_F, pow_f(ivdwp, D / 2.0f)};
}
inline float gauss_height(V3 r, const Gauss& g) {  // kernels.glsl:84-89
    return exp_f(-dot(r, r) * g.ivd) * g.norm;
}
inline V3 gauss_gradient(V3 r, const Gauss& g, float height_value) {  // kernels.glsl:91-97
    float dist = length(r);
    if (dist < 0.0001f) return v3(0, 0, 0);
    float s = -height_value * 2.0f * dist * g.ivd;
    return muls(normalize(r), s);
}

struct KernelEval {
    int height_id, gradient_id;
    float D;
    // returns W(r) and grad W(r) for kernel width kw with precomputed Gauss constants g
    void eval(V3 r, float kw, const Gauss& g, float* w, V3* grad, bool need_w, bool need_g) const {
        float gh = 0.0f;
        bool gh_done = false;
        if (need_w) {
            switch (height_id) {
                case 0: *w = cubic_kernel_height(r, kw); break;
                case 1: gh = gauss_height(r, g); gh_done = true; *w = gh; break;
                case 2: *w = poly6_kernel_height(r, kw); break;
                case 3: *w = cone_kernel_height(r, kw, D); break;
                case 4: *w = quadratic_spike_kernel_height(r, kw, D); break;
                default: *w = 0.0f;
            }
        }
        if (need_g) {
            switch (gradient_id) {
                case 0: *grad = cubic_kernel_gradient(r, kw); break;
                case 1: {
                    if (!gh_done) gh = gauss_height(r, g);
                    *grad = gauss_gradient(r, g, gh);
                    break;
                }
                case 2: *grad = spiky_kernel_gradient(r, kw); break;
                case 3: *grad = cone_kernel_gradient(r, kw, D); break;
                case 4: *grad = quadratic_spike_kernel_gradient(r, kw, D); break;
                default: *grad = v3(0, 0, 0);
            }
        }
    }
};

// ------------------------------------------------------------ list algorithms
// algorithms::sort (algorithms.cpp:59-91): stable LSD radix sort on 4-bit
// digits, one pass per digit while (upperBound >> offset) != 0.
void radix_sort(const uint32_t* keys, const uint32_t* vals, uint32_t n, uint32_t upper,
                std::vector<uint32_t>& ok, std::vector<uint32_t>& ov) {
    ok.assign(keys, keys + n);
    ov.assign(vals, vals + n);
    std::vector<uint32_t> tk(n), tv(n);
    for (uint32_t off = 0; off < 32u && (upper >> off) != 0u; off += 4u) {
        uint32_t cnt[17] = {0};
        for (uint32_t i = 0; i < n; ++i) cnt[((ok[i] >> off) & 15u) + 1]++;
        for (int b = 0; b < 16; ++b) cnt[b + 1] += cnt[b];
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t d = (ok[i] >> off) & 15u;
            tk[cnt[d]] = ok[i];
            tv[cnt[d]] = ov[i];
            cnt[d]++;
        }
        ok.swap(tk);
        ov.swap(tv);
    }
}

// calculate_position_hash.comp:23-36 / neighborhood_green.comp:35-47 for
// `dims` = DIMENSIONS. neighborhood_green.comp:36 zeroes the cell coordinates
// of the axes past DIMENSIONS (so in 2D the walk stays in z cell 0); the
// position hash never reads them, so zeroing them here changes no hash.
struct Grid {
    V3 mn, mx;
    uint32_t res;
    uint32_t dims;
    uint32_t cell_of(float p, float lo, float hi) const {
        return to_u32((p - lo) / (hi - lo) * (float)(1u << res));
    }
    void map(V3 p, uint32_t g[3]) const {
        g[0] = cell_of(p.x, mn.x, mx.x);
        g[1] = dims > 1u ? cell_of(p.y, mn.y, mx.y) : 0u;
        g[2] = dims > 2u ? cell_of(p.z, mn.z, mx.z) : 0u;
    }
    uint32_t hash(const uint32_t g[3]) const {
        uint32_t result = 0u;
        for (uint32_t i = 0; i < res; ++i)
            for (uint32_t d = 0; d < dims; ++d) result += ((g[d] >> i) & 1u) << (i * dims + d);
        return result;
    }
};

// calculate_position_code.comp:23-62: bit b of the 96-bit code is bit (b/3) of
// coordinate (b%3) of the position reinterpreted as uint.
uint32_t position_code(const I4& p, uint32_t section) {
    uint32_t u[3] = {(uint32_t)p.x, (uint32_t)p.y, (uint32_t)p.z};
    uint32_t r = 0u;
    for (uint32_t k = 0; k < 32u; ++k) {
        uint32_t b = section * 32u + k;
        if ((u[b % 3u] >> (b / 3u)) & 1u) r |= 1u << k;
    }
    return r;
}

// Index-list rebuild after a hidden-list edit: atomic_swap.comp:12-22 builds,
// per old hidden id, a linked list of the index-list positions that point to it
// (the most recently swapped position first, i.e. descending position order in
// the sequential schedule); generate_new_index_and_edit_list.comp:21-40 then
// walks, for each new hidden slot h' in ascending order, the list of its source
// slot and appends (h', position) until max_new entries exist.
uint32_t hidden_edit(const std::vector<uint32_t>& idx, uint32_t n_hidden, const std::vector<uint32_t>& edit,
                     uint32_t max_new, std::vector<uint32_t>& new_idx, std::vector<uint32_t>& new_edit) {
    std::vector<uint32_t> head(n_hidden, 0xFFFFFFFFu), next(idx.size(), 0xFFFFFFFFu);
    for (uint32_t p = 0; p < idx.size(); ++p) {  // atomic_swap, id order
        uint32_t b = idx[p];
        if (b >= n_hidden) continue;  // out-of-range entries point nowhere
        next[p] = head[b];
        head[b] = p;
    }
    new_idx.clear();
    new_edit.clear();
    for (uint32_t h = 0; h < edit.size(); ++h) {
        uint32_t it = head[edit[h]];
        while (it != 0xFFFFFFFFu) {
            if (new_idx.size() >= max_new) return (uint32_t)new_idx.size();
            new_idx.push_back(h);
            new_edit.push_back(it);
            it = next[it];
        }
    }
    return (uint32_t)new_idx.size();
}

// indexed_list::delete_these (indexed_list.h:126-138): flags, inclusive scan,
// find_value_changes.comp:17-27 -> edit list of the surviving hidden slots.
std::vector<uint32_t> delete_edit(const std::vector<uint32_t>& deleted, uint32_t n_hidden) {
    std::vector<uint32_t> flag(n_hidden, 1u);
    for (uint32_t d : deleted) flag[d] = 0u;  // scattered_write.comp:23
    std::vector<uint32_t> e;
    uint32_t run = 0;
    for (uint32_t id = 0; id < n_hidden; ++id) {  // prefix sum + find_value_changes
        uint32_t prev = run;
        run += flag[id];
        if (run != prev) e.push_back(id);
    }
    return e;
}

template <class T>
void gather(std::vector<T>& v, const std::vector<uint32_t>& edit) {  // copy_scattered_read.comp:17-27
    std::vector<T> out(edit.size());
    for (size_t i = 0; i < edit.size(); ++i) out[i] = v[edit[i]];
    v.swap(out);
}

// ---------------------------------------------------------------------- scene
struct Scene {
    orc_scene_params P;
    // hidden particles (list_definitions.h:9-10)
    std::vector<I4> position, pos_backup;
    std::vector<F4> velocity;
    std::vector<float> inverse_mass, radius;
    std::vector<uint32_t> transferring;
    // fluid (list_definitions.h:12-13), indexed by fluid id
    std::vector<uint32_t> fidx;
    std::vector<float> target_radius, kernel_width, boundariness;
    std::vector<uint32_t> boundary_distance;
    // hidden transfers (list_definitions.h:15-16)
    std::vector<uint32_t> t_source, t_target;
    std::vector<float> t_time_left;
    // neighbour pairs (fluid ids)
    std::vector<U2> pairs;
    // temporaries kept for inspection
    std::vector<float> lambda;
    std::vector<uint32_t> split_list;
    // intermediates of the last update_transfers: boundary distance after the
    // flood fill (find_split_and_merge_1), shortest neighbour distance and the
    // nearest neighbour (find_split_and_merge_2), per fluid id
    std::vector<uint32_t> fsm_bd_fill, fsm_min_dist, fsm_nearest;
    float last_dt = 1.0f;  // velocity_handling.h:17

    KernelEval kern() const { return KernelEval{P.s.mHeightKernelId, P.s.mGradientKernelId, dim()}; }
    // DIMENSIONS (cpu_gpu_shared_config.h:2): 2 or 3; 0 in the params means 3.
    uint32_t dims() const { return P.dimensions == 2u ? 2u : 3u; }
    float dim() const { return (float)dims(); }
    uint32_t H() const { return (uint32_t)position.size(); }
    uint32_t F() const { return (uint32_t)fidx.size(); }

    // A hidden-particle edit (hidden_data::apply_edit, indexed_list.h:60-64 and :289-308):
    // gather the hidden SoA, rebuild every owner index list. Owners in
    // registration order: transfer source list, transfer target list, fluid
    // particle list (pool.cpp:6-24 / waterdrop.cpp:6-22 with the drop released,
    // where mParticles holds the same indices as the fluid particle list).
    void hidden_particle_edit(const std::vector<uint32_t>& edit) {
        uint32_t nh = H();
        {
            std::vector<uint32_t> ni, ne;
            hidden_edit(t_source, nh, edit, P.cap_transfers, ni, ne);
            t_source = ni;
            gather(t_target, ne);
            gather(t_time_left, ne);
        }
        {
            std::vector<uint32_t> ni, ne;
            hidden_edit(t_target, nh, edit, P.cap_transfers, ni, ne);
            t_target = ni;
            gather(t_source, ne);
            gather(t_time_left, ne);
        }
        {
            std::vector<uint32_t> ni, ne;
            hidden_edit(fidx, nh, edit, P.cap_particles, ni, ne);
            fidx = ni;
            gather(target_radius, ne);
            gather(kernel_width, ne);
            gather(boundariness, ne);
            gather(boundary_distance, ne);
        }
        gather(position, edit);
        gather(velocity, edit);
        gather(inverse_mass, edit);
        gather(radius, edit);
        gather(pos_backup, edit);
        gather(transferring, edit);
    }

    // --------------------------------------------- velocity_handling::apply
    // velocity_handling.cpp:15-31 with infer_velocity.comp:25-26,
    // apply_acceleration.comp:23-24, apply_velocity.comp:24-25.
    void velocity_handling(float dt) {
        if (dt == 0.0f) {
            pos_backup = position;
            return;
        }
        float denom = last_dt * POS_RES;
        for (uint32_t id = 0; id < F(); ++id) {
            uint32_t idx = fidx[id];
            V3 d = idiff(position[idx], pos_backup[idx]);
            velocity[idx].x = d.x / denom;
            velocity[idx].y = d.y / denom;
            velocity[idx].z = d.z / denom;
        }
        pos_backup = position;
        V3 acc = v3(P.gravity[0] * dt, P.gravity[1] * dt, P.gravity[2] * dt);
        for (uint32_t id = 0; id < F(); ++id) {
            uint32_t idx = fidx[id];
            velocity[idx].x += acc.x;
            velocity[idx].y += acc.y;
            velocity[idx].z += acc.z;
        }
        float s = dt * POS_RES;
        for (uint32_t id = 0; id < F(); ++id) {
            uint32_t idx = fidx[id];
            I4& p = position[idx];
            const F4& v = velocity[idx];
            p.x = iadd(p.x, to_i32(v.x * s));
            p.y = iadd(p.y, to_i32(v.y * s));
            p.z = iadd(p.z, to_i32(v.z * s));
        }
        last_dt = dt;
    }

    // ---------------------------------------------- particle_transfer::apply
    // particle_transfer.cpp:10-28 with particle_transfer.comp:26-84.
    void particle_transfer(float dt) {
        std::vector<uint32_t> del_particles, del_transfers;
        for (uint32_t id = 0; id < (uint32_t)t_source.size(); ++id) {
            float ttl_and_type = t_time_left[id];
            float ttl = gmax(dt, fabsf(ttl_and_type));
            bool merge = ttl_and_type > 0.0f;
            uint32_t idS = t_source[id], idT = t_target[id];
            float rS = radius[idS], rT = radius[idT];
            float imS = inverse_mass[idS], imT = inverse_mass[idT];
            const float DIM = MUT(14) ? 3.0f : dim();  // mutation 14: 3D volumes in every build
            float volS = pow_f(rS, DIM), volT = pow_f(rT, DIM);
            float factorS = merge ? gmin(1.0f, dt / ttl) : dt * (volS - volT) / (2.0f * ttl * volS);
            float transfVol = factorS * volS;
            float normFactor = 1.0f / (imS + factorS * imT);
            inverse_mass[idS] = imS / (1.0f - factorS);
            inverse_mass[idT] = imS * imT * normFactor;
            radius[idS] = pow_f(volS - transfVol, 1.0f / DIM);
            radius[idT] = pow_f(volT + transfVol, 1.0f / DIM);
            t_time_left[id] = (ttl - dt) * (merge ? 1.0f : -1.0f);
            if (ttl == dt) {
                if (merge) del_particles.push_back(idS);
                else del_transfers.push_back(id);
                transferring[idS] = 0u;
                transferring[idT] = 0u;
            }
        }
        // deleteTransferList.delete_these(): compaction of the hidden transfer list
        {
            std::vector<uint32_t> e = delete_edit(del_transfers, (uint32_t)t_source.size());
            gather(t_source, e);
            gather(t_target, e);
            gather(t_time_left, e);
        }
        // deleteParticleList.delete_these(): compaction of the hidden particle list
        hidden_particle_edit(delete_edit(del_particles, H()));
    }

    // ---------------------- uint_to_float_with_indexed_lower_bound.comp:28-44
    static float move_towards_mult(float oldv, float newv, float step) {  // :20-26
        float dir = sign(newv - oldv);
        float result = oldv * (1.0f + step * dir);
        bool reached = (oldv < newv) != (result < newv);
        if (MUT(7)) reached = false;  // mutation 7: no snap at the crossing
        return reached ? newv : result;
    }
    void kernel_width_from_boundary_distance() {  // pool.cpp:77-80
        float factor = P.s.mTargetRadiusScaleFactor / POS_RES;
        for (uint32_t id = 0; id < F(); ++id) {
            uint32_t idx = fidx[id];
            float value = (float)boundary_distance[id] * factor;
            float lower = MUT(8) ? radius[idx] : radius[idx] * KERNEL_SCALE;  // mutation 8: bound without KERNEL_SCALE
            value = move_towards_mult(kernel_width[id], value, P.s.mKernelWidthAdaptionSpeed);
            kernel_width[id] = gmax(value, lower);
        }
    }

    // --------------------------------------------- neighborhood_green::apply
    Grid grid() const {
        return Grid{v3(P.hash_min[0], P.hash_min[1], P.hash_min[2]),
                    v3(P.hash_max[0], P.hash_max[1], P.hash_max[2]), P.hash_res_log2, dims()};
    }
    void neighborhood(float range_scale) {  // neighborhood_green.cpp:27-77
        Grid g = grid();
        uint32_t maxHash = 1u << (g.res * g.dims);
        uint32_t nh = H();
        std::vector<uint32_t> iota(nh), hashes(nh);
        for (uint32_t i = 0; i < nh; ++i) iota[i] = i;
        for (uint32_t id = 0; id < nh; ++id) {  // calculate_position_hash.comp:39-46
            uint32_t c[3];
            g.map(pos_world(position[id]), c);
            hashes[id] = g.hash(c);
        }
        std::vector<uint32_t> sorted_hash, sorted_index;
        radix_sort(hashes.data(), iota.data(), nh, maxHash, sorted_hash, sorted_index);
        hidden_particle_edit(sorted_index);  // :78
        // mParticles->sort() (:80, indexed_list.h:276-286): stable sort of the fluid
        // index list by value, applied to the fluid attributes.
        {
            std::vector<uint32_t> io(F()), sk, si;
            for (uint32_t i = 0; i < F(); ++i) io[i] = i;
            radix_sort(fidx.data(), io.data(), F(), P.cap_particles, sk, si);
            fidx = sk;
            gather(target_radius, si);
            gather(kernel_width, si);
            gather(boundariness, si);
            gather(boundary_distance, si);
        }
        std::vector<uint32_t> cs(maxHash, 0u), ce(maxHash, 0u);  // :82-86
        find_value_ranges(fidx.data(), sorted_hash.data(), F(), cs.data(), ce.data());
        pairs.clear();
        green(g, cs.data(), ce.data(), range_scale);
    }

    // find_value_ranges.comp:16-31. At id 0 the shader reads inIndexBuffer[-1];
    // the oracle treats id 0 as having no predecessor.
    static void find_value_ranges(const uint32_t* index, const uint32_t* values, uint32_t n,
                                  uint32_t* start, uint32_t* end) {
        for (uint32_t id = 0; id < n; ++id) {
            uint32_t cur = values[index[id]];
            bool has_prev = id > 0;
            uint32_t prev = has_prev ? values[index[id - 1]] : cur;
            if (id == 0 || cur != prev) start[cur] = id;
            if (has_prev && cur != prev) end[prev] = id;
            if (id == n - 1) end[cur] = id + 1u;
        }
    }

    // ------------------------------------- neighborhood_binary_search::apply
    // neighborhood_binary_search.cpp:22-75: three chained stable sorts by the
    // 32-bit sections of the 96-bit z-order code of the raw positions
    // (lexicographic by section 2, 1, 0), the particles permuted into that
    // order, then neighborhood_binary_search.comp per particle.
    int nb_method = 0;  // 0: neighborhood_green, 1: neighborhood_binary_search, 2: neighborhood_brute_force

    // neighborhood_brute_force.cpp:19-42 with neighborhood_brute_force.comp:31-50
    // and neighbor_add.glsl: every fluid particle against every other in fluid
    // order; no sort, no permutation; the list clamps at the pair capacity.
    void neighborhood_bf(float range_scale) {
        const uint32_t cap = P.cap_pairs;
        uint32_t len = 0;
        pairs.clear();
        for (uint32_t id = 0; id < F(); ++id) {
            const float range = kernel_width[id] * range_scale;
            const V3 pos = pos_world(position[fidx[id]]);
            for (uint32_t idN = 0; idN < F(); ++idN) {
                const V3 posN = pos_world(position[fidx[idN]]);
                if (id == idN || length(sub(pos, posN)) > range) continue;
                const uint32_t i = len++;
                if (i < cap) pairs.push_back(U2{id, idN});
                else len = cap;
            }
        }
    }

    void neighborhood_bs(float range_scale) {
        uint32_t nh = H();
        std::vector<uint32_t> idx(nh), keys(nh), sk, si;
        for (uint32_t i = 0; i < nh; ++i) idx[i] = i;
        for (uint32_t s = 0; s < 3u; ++s) {
            for (uint32_t j = 0; j < nh; ++j) keys[j] = position_code(position[idx[j]], s);
            radix_sort(keys.data(), idx.data(), nh, 0xFFFFFFFFu, sk, si);
            idx = si;
        }
        hidden_particle_edit(idx);
        {
            std::vector<uint32_t> io(F()), fk, fi;
            for (uint32_t i = 0; i < F(); ++i) io[i] = i;
            radix_sort(fidx.data(), io.data(), F(), P.cap_particles, fk, fi);
            fidx = fk;
            gather(target_radius, fi);
            gather(kernel_width, fi);
            gather(boundariness, fi);
            gather(boundary_distance, fi);
        }
        pairs.clear();
        binary_search_walk(range_scale);
    }

    struct C96 {
        uint32_t v[3];
    };
    static bool gt96(const C96& a, const C96& b) {  // greater96
        if (a.v[2] != b.v[2]) return a.v[2] > b.v[2];
        if (a.v[1] != b.v[1]) return a.v[1] > b.v[1];
        return a.v[0] > b.v[0];
    }
    static C96 plus96(const C96& a, const C96& b) {
        C96 r{{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]}};
        bool y = r.v[0] < a.v[0];
        bool z = r.v[1] < a.v[1] || (y && r.v[1] == 0xFFFFFFFFu);
        r.v[1] += y ? 1u : 0u;
        r.v[2] += z ? 1u : 0u;
        return r;
    }
    static C96 minus96(const C96& a, const C96& b) {
        bool y = a.v[0] < b.v[0];
        bool z = a.v[1] < b.v[1] || (y && a.v[1] == b.v[1]);
        return C96{{a.v[0] - b.v[0], a.v[1] - b.v[1] - (y ? 1u : 0u), a.v[2] - b.v[2] - (z ? 1u : 0u)}};
    }
    static C96 shl96(const C96& a, uint32_t d) {  // leftShift96 for 0 < d < 32
        return C96{{a.v[0] << d, (a.v[1] << d) | (a.v[0] >> (32u - d)), (a.v[2] << d) | (a.v[1] >> (32u - d))}};
    }
    static C96 and96(const C96& a, const C96& b) { return C96{{a.v[0] & b.v[0], a.v[1] & b.v[1], a.v[2] & b.v[2]}}; }
    static C96 or96(const C96& a, const C96& b) { return C96{{a.v[0] | b.v[0], a.v[1] | b.v[1], a.v[2] | b.v[2]}}; }
    static C96 not96(const C96& a) { return C96{{~a.v[0], ~a.v[1], ~a.v[2]}}; }

    // ceil(log2(x)) of the exact logarithm for positive finite x (the shader's
    // uint(ceil(log2(range * POS_RESOLUTION)))); 0 for x <= 1 and NaN, and
    // UINT_MAX for +inf (the saturating float-to-uint conversion).
    static uint32_t ceil_log2_u(float x) {
        if (!(x > 1.0f)) return 0u;
        if (x == INFINITY) return 0xFFFFFFFFu;
        int e;
        float m = std::frexp(x, &e);  // x = m * 2^e, m in [0.5, 1)
        return (uint32_t)(m == 0.5f ? e - 1 : e);
    }

    void binary_search_walk(float range_scale) {  // neighborhood_binary_search.comp:215-276
        const uint32_t n = F();
        const uint32_t cap = P.cap_pairs;
        uint32_t len = 0;
        std::vector<C96> codes(n);
        for (uint32_t id = 0; id < n; ++id)
            for (uint32_t s = 0; s < 3u; ++s) codes[id].v[s] = position_code(position[fidx[id]], s);
        const C96 ones{{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}};
        // The shader keeps its workgroup's 256 codes in shared memory; entries of
        // the last workgroup past the list end are never written there (their
        // invocations return first). They are taken as the largest code.
        auto block_code = [&](uint32_t k) { return k < n ? codes[k] : ones; };
        auto search = [&](const C96& code, uint32_t id) {  // binary_search, :148-190
            const uint32_t b0 = (id / 256u) * 256u;
            uint32_t lo = 0u, hi = n;
            if (!gt96(block_code(b0), code) && !gt96(code, block_code(b0 + 255u))) {
                lo = b0;
                hi = b0 + 256u;
            }
            uint32_t mid = (lo + hi) / 2u;
            for (int it = 0; it < 31; ++it) {
                const C96 c = (mid - b0 < 256u) ? block_code(mid) : codes[mid];
                if (gt96(code, c)) lo = mid + 1u;
                else hi = mid;
                mid = std::min((lo + hi) / 2u, n - 1u);
            }
            return lo;
        };
        const C96 xm{{011111111111u, 022222222222u, 04444444444u}};
        const C96 ym{{022222222222u, 04444444444u, 011111111111u}};
        const C96 zm{{04444444444u, 011111111111u, 022222222222u}};
        for (uint32_t id = 0; id < n; ++id) {
            const float range = kernel_width[id] * range_scale;
            const I4 ip = position[fidx[id]];
            const C96 code = codes[id];
            const uint32_t digits = ceil_log2_u(range * POS_RES) * 3u;
            C96 mask;
            mask.v[0] = (digits < 32u ? 1u << digits : 0u) - 1u;
            mask.v[1] = (digits < 64u ? 1u << (std::max(digits, 32u) - 32u) : 0u) - 1u;
            mask.v[2] = (digits < 96u ? 1u << (std::max(digits, 64u) - 64u) : 0u) - 1u;
            const C96 center{{code.v[0] - (code.v[0] & mask.v[0]), code.v[1] - (code.v[1] & mask.v[1]),
                              code.v[2] - (code.v[2] & mask.v[2])}};
            const C96 step = plus96(mask, C96{{1u, 0u, 0u}});
            C96 X[3], Y[3], Z[3];
            X[0] = and96(minus96(and96(center, xm), step), xm);
            X[2] = and96(plus96(or96(center, not96(xm)), step), xm);
            Y[0] = and96(minus96(and96(center, ym), shl96(step, 1u)), ym);
            Y[2] = and96(plus96(or96(center, not96(ym)), shl96(step, 1u)), ym);
            Z[0] = and96(minus96(and96(center, zm), shl96(step, 2u)), zm);
            Z[2] = and96(plus96(or96(center, not96(zm)), shl96(step, 2u)), zm);
            X[1] = and96(center, xm);
            Y[1] = and96(center, ym);
            Z[1] = and96(center, zm);
            const V3 pos = pos_world(ip);
            for (uint32_t c = 0; c < 27u; ++c) {
                const C96 cell = or96(or96(X[c % 3u], Y[(c / 3u) % 3u]), Z[c / 9u]);
                const C96 last = or96(cell, mask);
                for (uint32_t idN = search(cell, id); idN < n && !gt96(codes[idN], last); ++idN) {
                    const V3 posN = pos_world(position[fidx[idN]]);
                    if (id != idN && length(sub(pos, posN)) <= range) {
                        uint32_t i = len++;  // neighbor_add.glsl
                        if (i < cap) pairs.push_back(U2{id, idN});
                        else len = cap;
                    }
                }
            }
        }
    }

    // neighborhood_green.comp:49-87 with neighbor_add.glsl:11-25 (unsorted list)
    void green(const Grid& g, const uint32_t* cs, const uint32_t* ce, float range_scale) {
        uint32_t cap = P.cap_pairs;
        uint32_t len = 0;
        for (uint32_t id = 0; id < F(); ++id) {
            float range = kernel_width[id] * range_scale;
            uint32_t idx = fidx[id];
            V3 pos = pos_world(position[idx]);
            uint32_t gmn[3], gmx[3];
            g.map(sub(pos, v3(range, range, range)), gmn);
            g.map(add(pos, v3(range, range, range)), gmx);
            // The walk is bounded to 4 * 2^res cells per axis (wider walks only
            // revisit aliased cells; an infinite range would never finish).
            for (int d = 0; d < 3; ++d)
                if (gmx[d] >= gmn[d] && gmx[d] - gmn[d] >= (4u << g.res)) gmx[d] = gmn[d] + (4u << g.res) - 1u;
            uint32_t cell[3] = {gmn[0], gmn[1], gmn[2]};
            uint32_t h = g.hash(cell);
            uint32_t cellStart = cs[h], cellEnd = ce[h], idN = cellStart;
            for (;; idN++) {
                bool done = false;
                while (idN >= cellEnd) {
                    if (++cell[0] > gmx[0]) cell[0] = gmn[0];
                    if (cell[0] == gmn[0] && ++cell[1] > gmx[1]) cell[1] = gmn[1];
                    if (cell[1] == gmn[1] && cell[0] == gmn[0] && ++cell[2] > gmx[2]) { done = true; break; }
                    h = g.hash(cell);
                    cellStart = cs[h];
                    cellEnd = ce[h];
                    idN = cellStart;
                }
                if (done) break;
                uint32_t idxN = fidx[idN];
                V3 posN = pos_world(position[idxN]);
                if (id == idN || length(sub(pos, posN)) > range) continue;
                uint32_t i = len++;  // neighbor_add.glsl:12-23
                if (i < cap) pairs.push_back(U2{id, idN});
                else len = cap;
            }
        }
    }

    // --------------------------------------------- spread_kernel_width::apply
    // spread_kernel_width.cpp:12-26 with kernel_width_init.comp:17-31,
    // kernel_width.comp:22-56, uint_to_float_but_gradual.comp:18-36.
    void spread_kernel_width() {
        std::vector<U2> old = pairs;
        pairs.clear();
        std::vector<uint32_t> ukw(F());
        for (uint32_t id = 0; id < F(); ++id) {
            uint32_t idx = fidx[id];
            float tr = P.s.mBaseKernelWidthOnTargetRadius ? target_radius[id] : 0.0f;
            float okw = gmax(radius[idx], tr) * KERNEL_SCALE;
            ukw[id] = to_u32(okw * KW_RES);
        }
        for (uint32_t id = 0; id < (uint32_t)old.size(); ++id) {
            U2 n = old[id];
            uint32_t idx = fidx[n.a], idxN = fidx[n.b];
            float r = radius[idx];
            float tr = P.s.mBaseKernelWidthOnTargetRadius ? target_radius[n.a] : 0.0f;
            float oldKw = kernel_width[n.a];
            V3 diff = divs(idiff(position[idxN], position[idx]), POS_RES);
            float dist = length(diff);
            float okw = gmax(r, tr) * KERNEL_SCALE;
            float cutoff = gmax(okw, oldKw);
            float dfk = dist - okw;
            float influence = gmax(0.0f, 1.0f - gmax(0.0f, dfk / (okw * KW_PROPAGATION)));
            uint32_t v = to_u32(okw * influence * KW_RES);
            if (v > ukw[n.b]) ukw[n.b] = v;
            if (dist <= cutoff) pairs.push_back(n);
        }
        float factor = 1.0f / KW_RES;
        for (uint32_t id = 0; id < F(); ++id) {
            float value = (float)ukw[id] * factor;
            value = move_towards_mult(kernel_width[id], value, P.s.mKernelWidthAdaptionSpeed);
            kernel_width[id] = gmax(value, -INFINITY);
        }
    }

    // --------------------------------------------------- box_collision::apply
    // box_collision.cpp:12-24 with box_collision.comp:15-55.
    static V3 hash31(float p) {  // box_collision.comp:15-20 ("Hash without Sine", MIT)
        V3 p3 = v3(fract(p * 0.1031f), fract(p * 0.1030f), fract(p * 0.0973f));
        float d = dot(p3, v3(p3.y + 33.33f, p3.z + 33.33f, p3.x + 33.33f));
        p3 = v3(p3.x + d, p3.y + d, p3.z + d);
        return v3(fract((p3.x + p3.y) * p3.z), fract((p3.x + p3.z) * p3.y), fract((p3.y + p3.z) * p3.x));
    }
    static void smallest(V3 v, float m[3]) {  // box_collision.comp:22-28
        bool b1 = v.x <= v.y, b2 = v.x <= v.z, b3 = v.y <= v.z;
        m[0] = (b1 && b2) ? 1.0f : 0.0f;
        m[1] = (!b1 && b3) ? 1.0f : 0.0f;
        m[2] = (!b2 && !b3) ? 1.0f : 0.0f;
    }
    void box_collision() {
        for (uint32_t id = 0; id < F(); ++id) {
            uint32_t idx = fidx[id];
            V3 pos = pos_world(position[idx]);
            float r = radius[idx];
            for (uint32_t i = 0; i < P.n_boxes; ++i) {
                V3 jmin = muls(hash31((float)id * pos.x - pos.y - pos.z), 0.05f);
                V3 bmin = sub(sub(v3(P.box_min[i][0], P.box_min[i][1], P.box_min[i][2]), v3(r, r, r)), jmin);
                V3 jmax = muls(hash31((float)id * pos.y + pos.x + pos.z), 0.05f);
                V3 bmax = add(add(v3(P.box_max[i][0], P.box_max[i][1], P.box_max[i][2]), v3(r, r, r)), jmax);
                V3 toMin = sub(bmin, pos), toMax = sub(bmax, pos);
                float m[3];
                smallest(neg(toMin), m);
                toMin = mul(toMin, v3(m[0], m[1], m[2]));
                smallest(toMax, m);
                toMax = mul(toMax, v3(m[0], m[1], m[2]));
                float distToMin = -dot(toMin, v3(1, 1, 1));
                float distToMax = dot(toMax, v3(1, 1, 1));
                if (distToMin <= 0.0f || distToMax <= 0.0f) continue;
                pos = add(pos, distToMin < distToMax ? toMin : toMax);
            }
            I4& p = position[idx];
            p.x = to_i32(pos.x * POS_RES);
            p.y = to_i32(pos.y * POS_RES);
            p.z = to_i32(pos.z * POS_RES);
        }
    }

    // ----------------------------------------------- incompressibility::apply
    // incompressibility.cpp:12-45 with incompressibility_0..3.comp.
    struct Incomp { int32_t g[3]; uint32_t density, sq; };
    struct Com { int32_t x, y, z, w; };

    void density_pass(std::vector<Incomp>& inc, std::vector<Com>& com, std::vector<V3>* grads) const {
        KernelEval ke = kern();
        uint32_t nf = F();
        inc.assign(nf, Incomp{{0, 0, 0}, 0u, 0u});
        com.assign(nf, Com{0, 0, 0, 0});
        std::vector<Gauss> gk(nf);
        for (uint32_t id = 0; id < nf; ++id) gk[id] = gauss_of_width(kernel_width[id], dim());
        for (uint32_t id = 0; id < nf; ++id) {  // incompressibility_0.comp:33-45
            uint32_t idx = fidx[id];
            float w;
            V3 gdummy;
            ke.eval(v3(0, 0, 0), kernel_width[id], gk[id], &w, &gdummy, true, false);
            inc[id].density = to_u32(w / inverse_mass[idx] * INCOMP_RES);
        }
        if (grads) grads->assign(pairs.size(), v3(0, 0, 0));
        for (uint32_t id = 0; id < (uint32_t)pairs.size(); ++id) {  // incompressibility_1.comp:38-69
            U2 n = pairs[id];
            uint32_t idx = fidx[n.a], idxN = fidx[n.b];
            const I4& pos = position[idx];
            const I4& posN = position[idxN];
            float imN = inverse_mass[idxN];
            float kw = kernel_width[n.a];
            V3 diff = divs(idiff(posN, pos), POS_RES);
            float w;
            V3 grad;
            ke.eval(diff, kw, gk[n.a], &w, &grad, true, true);
            V3 wg = divs(grad, imN);
            Incomp& c = inc[n.a];
            c.density += to_u32(w / imN * INCOMP_RES);
            c.g[0] = iadd(c.g[0], to_i32(wg.x * INCOMP_RES));
            c.g[1] = iadd(c.g[1], to_i32(wg.y * INCOMP_RES));
            c.g[2] = iadd(c.g[2], to_i32(wg.z * INCOMP_RES));
            c.sq += to_u32(dot(grad, grad) / imN * INCOMP_RES);
            if (grads) (*grads)[id] = grad;
            if (P.s.mBoundarinessCalculationMethod == 1) {
                float s = imN * kw;
                V3 wd = divs(idiff(posN, pos), s);
                Com& m = com[n.a];
                m.x = iadd(m.x, to_i32(wd.x));
                m.y = iadd(m.y, to_i32(wd.y));
                m.z = iadd(m.z, to_i32(wd.z));
                m.w = iadd(m.w, to_i32(INCOMP_RES / imN));
            }
        }
    }

    static float move_towards_add(float oldv, float newv, float maxStep) {  // incompressibility_2.comp:37-41
        float step = newv - oldv;
        return oldv + gmin(maxStep, gmax(-maxStep, step));
    }

    // What an iteration's shifts do (tolerance mode and the warm-start
    // fallback of the library's solver loop; no counterpart in the reference).
    enum Action { APPLY = 0, SKIP = 1, RESTORE = 2 };

    // Mean over the fluid of max(0, rho / rho0 - 1) from the sums of
    // incompressibility_0/1 (the density-error definition of the library's
    // statistics), summed in id order in double precision.
    double mean_error_of(const std::vector<Incomp>& inc) const {
        double sum = 0.0;
        for (uint32_t id = 0; id < F(); ++id) {
            uint32_t idx = fidx[id];
            float ird = pow_f(2.0f * radius[idx], dim()) * inverse_mass[idx];
            float density = (float)inc[id].density / INCOMP_RES;
            sum += (double)gmax(0.0f, density * ird - 1.0f);
        }
        return F() ? sum / (double)F() : 0.0;
    }

    // One incompressibility::apply. Library extensions: `lam_override` replaces
    // the lambdas of incompressibility_2 (no boundariness update or filter:
    // the warm-start shift pass); `decide`, given the mean density error of the
    // positions the pass starts from, chooses what the shifts do (the
    // boundariness update and filter run either way). Returns that error.
    template <class Decide>
    double incompressibility_ext(const float* lam_override, Decide decide) {
        std::vector<Incomp> inc;
        std::vector<Com> com;
        std::vector<V3> grads;
        density_pass(inc, com, &grads);
        const double err = mean_error_of(inc);
        const int action = decide(err);
        const bool shifts = action == APPLY;
        if (action == RESTORE) position = saved_position;
        uint32_t nf = F();
        lambda.assign(nf, 0.0f);
        float lambda_scale_base = pow_f(2.0f * P.s.mSmallestTargetRadius, dim());
        for (uint32_t id = 0; id < nf; ++id) {  // incompressibility_2.comp:72-110
            uint32_t idx = fidx[id];
            float r = radius[idx];
            float im = inverse_mass[idx];
            float kw = kernel_width[id];
            float ird = pow_f(2.0f * r, MUT(15) ? 3.0f : dim()) * im;  // mutation 15: the 3D rest density
            float density = (float)inc[id].density / INCOMP_RES;
            V3 wgs = divs(vec_of(inc[id].g[0], inc[id].g[1], inc[id].g[2]), INCOMP_RES);
            float sgs = (float)inc[id].sq / INCOMP_RES;
            V3 selfGrad = muls(neg(wgs), ird);
            float wgss = dot(wgs, wgs);
            float selfGradLength = sqrtf(wgss) * ird;
            sgs += wgss * im;
            float under = 1.0f - density * ird;
            if (P.s.mUpdateBoundariness && !lam_override) {  // compute_boundariness, :43-69
                float sgl = selfGradLength * kw;
                float b;
                if (P.s.mBoundarinessCalculationMethod == 1) {
                    float totalMass = (float)com[id].w / INCOMP_RES + 1.0f / im;
                    float dev = length(vec_of(com[id].x, com[id].y, com[id].z)) / (POS_RES * totalMass);
                    b = dev * P.s.mBoundarinessSelfGradLengthFactor;
                } else {
                    sgl *= P.s.mBoundarinessSelfGradLengthFactor;
                    float up = gmax(0.0f, under) * P.s.mBoundarinessUnderpressureFactor;
                    b = sgl + up;
                }
                b = b >= 1.0f ? 1.0f : 0.0f;
                b = move_towards_add(boundariness[id], b, P.s.mBoundarinessAdaptionSpeed);
                boundariness[id] = gmin(1.0f, b);
            }
            float lam = under / (ird * ird * (sgs + INCOMP_EPSILON));
            lam /= lambda_scale_base / ird * im;
            if (lam_override) lam = lam_override[id];
            lambda[id] = lam;
            if (!shifts || lam >= 0.0f) continue;
            float s = lam * im * POS_RES;
            V3 d = muls(selfGrad, s);
            I4& p = position[idx];
            p.x = iadd(p.x, to_i32(d.x));
            p.y = iadd(p.y, to_i32(d.y));
            p.z = iadd(p.z, to_i32(d.z));
        }
        for (uint32_t id = 0; id < (uint32_t)pairs.size(); ++id) {  // incompressibility_3.comp:49-67
            U2 n = pairs[id];
            uint32_t idxN = fidx[n.b];
            float lam = lambda[n.a];
            V3 sg = grads[id];
            if (P.s.mBoundarinessCalculationMethod == 2 && !lam_override) {  // filter_boundariness, :41-46
                V3 empty = neg(normalize(vec_of(inc[n.a].g[0], inc[n.a].g[1], inc[n.a].g[2])));
                V3 nd = normalize(sg);
                if (dot(empty, nd) > 0.6f) boundariness[n.a] = 0.0f;
            }
            if (!shifts || lam >= 0.0f) continue;
            V3 d = muls(sg, lam * POS_RES);
            I4& p = position[idxN];
            p.x = iadd(p.x, to_i32(d.x));
            p.y = iadd(p.y, to_i32(d.y));
            p.z = iadd(p.z, to_i32(d.z));
        }
        return err;
    }

    std::vector<I4> saved_position;  // positions before the warm-start shift
    float tol = 0.0f;                // tolerance mode: iterate until the mean density error is <= tol
    int tol_max_iters = 0;
    std::vector<float> warm_lambda;  // predicted lambdas for the next update's warm-start shift
    uint64_t stat_iters = 0, stat_substeps = 0, stat_fallbacks = 0;

    void incompressibility() {
        incompressibility_ext(nullptr, [](double) { return (int)APPLY; });
    }

    void density_error(double* mean_err, double* max_err) const {
        std::vector<Incomp> inc;
        std::vector<Com> com;
        density_pass(inc, com, nullptr);
        double sum = 0.0, mx = 0.0;
        for (uint32_t id = 0; id < F(); ++id) {
            uint32_t idx = fidx[id];
            float ird = pow_f(2.0f * radius[idx], dim()) * inverse_mass[idx];
            float density = (float)inc[id].density / INCOMP_RES;
            float e = gmax(0.0f, density * ird - 1.0f);
            sum += (double)e;
            if ((double)e > mx) mx = (double)e;
        }
        *mean_err = F() ? sum / (double)F() : 0.0;
        *max_err = mx;
    }

    // ------------------------------------------------ update_transfers::apply
    // update_transfers.cpp:14-71 with find_split_and_merge_1..3.comp,
    // remove_impossible_splits.comp, initialize_split_particles.comp.
    void update_transfers() {
        uint32_t nf = F();
        std::vector<uint32_t> oldBd = boundary_distance;
        for (uint32_t id = 0; id < nf; ++id) boundary_distance[id] = 0xFFFFFFFFu;  // :37
        // :44 allocates and clears minNeighborDistList with the pair-list length but
        // indexes it by fluid id; sized by fluid id here.
        std::vector<uint32_t> minD(nf, 0xFFFFFFFFu);
        // nearestNeighborList (:34) is never initialised by the reference; a fluid id
        // without neighbours reads stale data. The oracle defines it as the id itself.
        std::vector<uint32_t> nearest(nf);
        for (uint32_t id = 0; id < nf; ++id) nearest[id] = id;
        for (uint32_t id = 0; id < (uint32_t)pairs.size(); ++id) {  // find_split_and_merge_1.comp:17-32
            U2 n = pairs[id];
            uint32_t idx = fidx[n.a], idxN = fidx[n.b];
            uint32_t bdN = MUT(1) ? oldBd[n.a] : oldBd[n.b];  // mutation 1: own old distance
            uint32_t dist = to_u32(length(idiff(position[idxN], position[idx])));
            uint32_t cand = bdN + dist;
            if (cand < boundary_distance[n.a]) boundary_distance[n.a] = cand;
            if (MUT(2) ? (cand < minD[n.a]) : (dist < minD[n.a])) minD[n.a] = MUT(2) ? cand : dist;  // mutation 2
        }
        fsm_bd_fill = boundary_distance;
        fsm_min_dist = minD;
        for (uint32_t id = 0; id < (uint32_t)pairs.size(); ++id) {  // find_split_and_merge_2.comp:17-33
            U2 n = pairs[id];
            uint32_t idx = fidx[n.a], idxN = fidx[n.b];
            uint32_t dist = to_u32(length(idiff(position[idxN], position[idx])));
            if (dist == minD[MUT(3) ? n.b : n.a]) nearest[n.a] = n.b;  // mutation 3: neighbour's minimum
        }
        fsm_nearest = nearest;
        uint32_t tlen = (uint32_t)t_source.size();
        uint32_t maxT = P.cap_transfers, maxS = P.cap_particles;
        split_list.clear();
        uint32_t slen = 0;
        float split_factor = pow_f(2.0f, 1.0f / (MUT(13) ? 3.0f : dim())) * 0.99f;  // mutation 13: the 3D factor
        for (uint32_t id = 0; id < nf; ++id) {  // find_split_and_merge_3.comp:42-117
            uint32_t idx = fidx[id];
            const I4& pos = position[idx];
            float r = radius[idx];
            float bdist = (float)boundary_distance[id] / POS_RES;
            float b = boundariness[id];
            uint32_t nnId = nearest[id];
            uint32_t nnIdx = fidx[nnId];
            float nnDist = length(idiff(position[nnIdx], pos)) / POS_RES;
            float nnR = radius[nnIdx];
            bool nnLarger = nnR > r || (nnR == r && nnIdx > idx);
            float tr;
            if (P.s.mUpdateTargetRadius) {
                if (P.s.mBaseKernelWidthOnBoundaryDistance) {
                    float f = P.s.mTargetRadiusScaleFactor;
                    tr = (f / (KERNEL_SCALE + KERNEL_SCALE * f)) * bdist;
                    tr = gmax(tr, P.s.mSmallestTargetRadius);
                } else {
                    tr = P.s.mSmallestTargetRadius +
                         (MUT(4) ? (bdist - P.s.mTargetRadiusOffset) * P.s.mTargetRadiusScaleFactor  // mutation 4: no clamp
                                 : gmax(0.0f, (bdist - P.s.mTargetRadiusOffset) * P.s.mTargetRadiusScaleFactor));
                }
                target_radius[id] = tr;
            } else {
                tr = target_radius[id];
            }
            b = b >= 1.0f ? 1.0f : 0.0f;
            boundary_distance[id] = MUT(5) ? to_u32((bdist * b + r * (1.0f - b)) * POS_RES)  // mutation 5: mix reversed
                                           : to_u32((bdist * (1.0f - b) + r * b) * POS_RES);  // mix()
            boundariness[id] = b;
            bool split = P.s.mSplit && (tr * split_factor <= r);
            bool merge = P.s.mMerge && (MUT(6) || nnDist < r) &&  // mutation 6: no distance condition
                         pow_f(r, dim()) + pow_f(nnR, dim()) <= pow_f(tr, dim());
            if (!merge && !split) continue;
            uint32_t s = (merge && nnLarger) ? nnIdx : idx;
            uint32_t t = (merge && nnLarger) ? idx : nnIdx;
            uint32_t old = transferring[s];
            transferring[s] = 1u;
            if (old == 1u) continue;
            if (merge) {
                uint32_t oldT = transferring[t];
                transferring[t] = 1u;
                if (oldT == 1u && !MUT(9)) { transferring[s] = 0u; continue; }  // mutation 9: target not claimed
                uint32_t tIdx = tlen++;
                if (tIdx >= maxT) {
                    tlen = maxT;
                    transferring[s] = 0u;
                    transferring[t] = 0u;
                    continue;
                }
                t_source.push_back(s);
                t_target.push_back(t);
                t_time_left.push_back(gmax(0.0001f, P.s.mMergeDuration));
            } else {
                uint32_t sIdx = slen++;
                if (sIdx >= maxS) {
                    slen = maxS;
                    transferring[s] = 0u;
                    continue;
                }
                split_list.push_back(idx);
            }
        }
        if (!P.s.mSplit) return;  // update_transfers.cpp:59 (settings::split)
        // remove_impossible_splits.comp:25-40
        uint32_t nsl = (uint32_t)split_list.size();
        uint32_t keep = std::min(std::min(nsl, maxT - (uint32_t)t_source.size()), P.cap_particles - H());
        for (uint32_t id = keep; id < nsl; ++id) transferring[split_list[id]] = 0u;
        split_list.resize(keep);
        // duplicate_these (indexed_list.h:141-152): edit = [0..H) ++ split list
        uint32_t nh = H();
        std::vector<uint32_t> edit(nh + keep);
        for (uint32_t i = 0; i < nh; ++i) edit[i] = i;
        for (uint32_t i = 0; i < keep; ++i) edit[nh + i] = split_list[i];
        hidden_particle_edit(edit);
        for (uint32_t k = 0; k < keep; ++k) {  // initialize_split_particles.comp:14-27
            uint32_t idx = nh + k;
            float r = radius[idx];
            position[idx].x = iadd(position[idx].x, to_i32(r * POS_RES * 0.1f));
            radius[idx] = 0.0f;
            inverse_mass[idx] = SPLIT_INV_MASS;
        }
        for (uint32_t k = 0; k < keep; ++k) {  // transfer lists += split / duplicates / -splitDuration
            if (t_source.size() >= P.cap_transfers) break;
            t_source.push_back(split_list[k]);
            t_target.push_back(nh + k);
            t_time_left.push_back(-P.split_duration);
        }
    }

    void update(float dt) {  // pool.cpp:67-106 / waterdrop.cpp:69-108
        const auto& s = P.s;
        velocity_handling(dt);
        if (!P.basic_pbf && (s.mMerge || s.mSplit)) particle_transfer(dt);
        if (!P.basic_pbf && s.mBaseKernelWidthOnBoundaryDistance) kernel_width_from_boundary_distance();
        const float scale = (P.basic_pbf || s.mBaseKernelWidthOnBoundaryDistance) ? 1.0f : 1.5f;
        if (nb_method == 1)
            neighborhood_bs(scale);
        else if (nb_method == 2)
            neighborhood_bf(scale);
        else
            neighborhood(scale);
        if (!P.basic_pbf && !s.mBaseKernelWidthOnBoundaryDistance) spread_kernel_width();
        solver_loop();
        if (!P.basic_pbf) update_transfers();
    }

    void colliders() {
        if (P.n_boxes) box_collision();  // scenes without boxes have no box stage
        sphere_collision();
    }

    // The solver loop of pool.cpp:92-95 (box collision + incompressibility,
    // solverIterations times), with the library's extensions: a tolerance mode
    // (up to tol_max_iters iterations; an iteration whose starting error is
    // <= tol applies no shifts and ends the loop) and the warm start (one shift
    // pass with given lambdas before the iterations; when the first plain
    // iteration starts from a larger error than the warm pass did, the
    // positions from before the warm shift are put back).
    void solver_loop() {
        const bool tolerance = tol > 0.0f;
        const int iters = tolerance ? tol_max_iters : P.solver_iterations;
        const bool warm = !warm_lambda.empty();
        if (iters <= 0) {
            warm_lambda.clear();
            return;
        }
        int applied = 0;
        bool done = false, fell_back = false;
        double e0 = 0.0;
        colliders();
        if (warm) {
            saved_position = position;
            e0 = incompressibility_ext(warm_lambda.data(), [](double) { return (int)APPLY; });
            if (!MUT(12)) colliders();
        }
        for (int k = 0; k < iters && !done; ++k) {
            incompressibility_ext(nullptr, [&](double e) {
                if (warm && k == 0 && (MUT(11) ? e < e0 : e > e0)) {
                    fell_back = true;
                    return (int)RESTORE;
                }
                if (tolerance && e <= (double)tol) {
                    done = true;
                    return MUT(10) ? (int)APPLY : (int)SKIP;
                }
                ++applied;
                return (int)APPLY;
            });
            if (k + 1 < iters && !done) colliders();
        }
        stat_iters += (tolerance || warm) ? (uint64_t)applied : (uint64_t)iters;
        stat_substeps += 1;
        stat_fallbacks += fell_back ? 1 : 0;
        warm_lambda.clear();
    }

    // sphere_collision.comp:42-60 (the collider of spherical_pool.cpp:29)
    bool sphere_on = false, sphere_hollow = false;
    float sphere_c[3] = {0, 0, 0}, sphere_r = 0.0f;
    static float hash11(float p) {  // sphere_collision.comp:24-30 ("Hash without Sine", MIT)
        p = fract(p * 0.1031f);
        p *= p + 33.33f;
        p *= p + p;
        return fract(p);
    }
    void sphere_collision() {
        if (!sphere_on) return;
        const V3 center = v3(sphere_c[0], sphere_c[1], sphere_c[2]);
        for (uint32_t id = 0; id < F(); ++id) {
            const uint32_t idx = fidx[id];
            V3 pos = pos_world(position[idx]);
            const float dir = sphere_hollow ? -1.0f : 1.0f;
            const float offset = dir * (radius[idx] + hash11((float)id * pos.y + pos.x + pos.z) * 0.05f);
            const V3 rel = sub(pos, center);
            const float d = length(rel);
            const float target = sphere_r + offset;
            if ((d < target) == sphere_hollow) continue;
            pos = add(center, muls(rel, target / d));
            I4& p = position[idx];
            p.x = to_i32(pos.x * POS_RES);
            p.y = to_i32(pos.y * POS_RES);
            p.z = to_i32(pos.z * POS_RES);
        }
    }
};

}  // namespace

// ============================================================== C interface
extern "C" {

void* orc_create(const orc_scene_params* p) {
    Scene* s = new Scene();
    s->P = *p;
    return s;
}
void orc_destroy(void* h) { delete static_cast<Scene*>(h); }

int orc_set_state(void* h, uint32_t nh, const int32_t* position, const float* velocity,
                  const float* inverse_mass, const float* radius, const int32_t* pos_backup,
                  const uint32_t* transferring, uint32_t nf, const uint32_t* fluid_index,
                  const float* target_radius, const float* kernel_width, const float* boundariness,
                  const uint32_t* boundary_distance, uint32_t nt, const uint32_t* t_source,
                  const uint32_t* t_target, const float* t_time_left, float last_delta_time) {
    Scene* s = static_cast<Scene*>(h);
    s->position.resize(nh);
    s->pos_backup.resize(nh);
    s->velocity.resize(nh);
    std::memcpy(s->position.data(), position, sizeof(I4) * nh);
    std::memcpy(s->pos_backup.data(), pos_backup, sizeof(I4) * nh);
    std::memcpy(s->velocity.data(), velocity, sizeof(F4) * nh);
    s->inverse_mass.assign(inverse_mass, inverse_mass + nh);
    s->radius.assign(radius, radius + nh);
    s->transferring.assign(transferring, transferring + nh);
    s->fidx.assign(fluid_index, fluid_index + nf);
    s->target_radius.assign(target_radius, target_radius + nf);
    s->kernel_width.assign(kernel_width, kernel_width + nf);
    s->boundariness.assign(boundariness, boundariness + nf);
    s->boundary_distance.assign(boundary_distance, boundary_distance + nf);
    s->t_source.assign(t_source, t_source + nt);
    s->t_target.assign(t_target, t_target + nt);
    s->t_time_left.assign(t_time_left, t_time_left + nt);
    s->last_dt = last_delta_time;
    s->pairs.clear();
    for (uint32_t i = 0; i < nf; ++i)
        if (fluid_index[i] >= nh) return -1;
    return 0;
}

void orc_get_counts(void* h, uint32_t* nh, uint32_t* nf, uint32_t* nt, uint32_t* np) {
    Scene* s = static_cast<Scene*>(h);
    *nh = s->H();
    *nf = s->F();
    *nt = (uint32_t)s->t_source.size();
    *np = (uint32_t)s->pairs.size();
}

void orc_get_state(void* h, int32_t* position, float* velocity, float* inverse_mass, float* radius,
                   int32_t* pos_backup, uint32_t* transferring, uint32_t* fluid_index,
                   float* target_radius, float* kernel_width, float* boundariness,
                   uint32_t* boundary_distance, uint32_t* t_source, uint32_t* t_target,
                   float* t_time_left, uint32_t* pairs, float* last_delta_time) {
    Scene* s = static_cast<Scene*>(h);
    uint32_t nh = s->H(), nf = s->F(), nt = (uint32_t)s->t_source.size();
    std::memcpy(position, s->position.data(), sizeof(I4) * nh);
    std::memcpy(velocity, s->velocity.data(), sizeof(F4) * nh);
    std::memcpy(pos_backup, s->pos_backup.data(), sizeof(I4) * nh);
    std::memcpy(inverse_mass, s->inverse_mass.data(), 4 * nh);
    std::memcpy(radius, s->radius.data(), 4 * nh);
    std::memcpy(transferring, s->transferring.data(), 4 * nh);
    std::memcpy(fluid_index, s->fidx.data(), 4 * nf);
    std::memcpy(target_radius, s->target_radius.data(), 4 * nf);
    std::memcpy(kernel_width, s->kernel_width.data(), 4 * nf);
    std::memcpy(boundariness, s->boundariness.data(), 4 * nf);
    std::memcpy(boundary_distance, s->boundary_distance.data(), 4 * nf);
    std::memcpy(t_source, s->t_source.data(), 4 * nt);
    std::memcpy(t_target, s->t_target.data(), 4 * nt);
    std::memcpy(t_time_left, s->t_time_left.data(), 4 * nt);
    std::memcpy(pairs, s->pairs.data(), 8 * s->pairs.size());
    *last_delta_time = s->last_dt;
}

void orc_get_lambda(void* h, float* lambda) {
    Scene* s = static_cast<Scene*>(h);
    std::memcpy(lambda, s->lambda.data(), 4 * s->lambda.size());
}
uint32_t orc_get_fsm(void* h, uint32_t* bd_fill, uint32_t* min_dist, uint32_t* nearest) {
    Scene* s = static_cast<Scene*>(h);
    const size_t n = s->fsm_nearest.size();
    if (bd_fill) std::memcpy(bd_fill, s->fsm_bd_fill.data(), 4 * n);
    if (min_dist) std::memcpy(min_dist, s->fsm_min_dist.data(), 4 * n);
    if (nearest) std::memcpy(nearest, s->fsm_nearest.data(), 4 * n);
    return (uint32_t)n;
}

uint32_t orc_get_splits(void* h, uint32_t* split_list) {
    Scene* s = static_cast<Scene*>(h);
    if (split_list) std::memcpy(split_list, s->split_list.data(), 4 * s->split_list.size());
    return (uint32_t)s->split_list.size();
}

void orc_set_neighborhood_method(void* h, int method) { static_cast<Scene*>(h)->nb_method = method; }
void orc_set_sphere(void* h, const float* center3, float radius, int hollow, int enable) {
    Scene* s = static_cast<Scene*>(h);
    s->sphere_on = enable != 0;
    s->sphere_hollow = hollow != 0;
    for (int d = 0; d < 3; ++d) s->sphere_c[d] = enable ? center3[d] : 0.0f;
    s->sphere_r = enable ? radius : 0.0f;
}

void orc_stage(void* h, int stage, float dt) {
    Scene* s = static_cast<Scene*>(h);
    const auto& st = s->P.s;
    switch (stage) {
        case ORC_STAGE_VELOCITY: s->velocity_handling(dt); break;
        case ORC_STAGE_PARTICLE_TRANSFER: s->particle_transfer(dt); break;
        case ORC_STAGE_KERNEL_WIDTH_BD: s->kernel_width_from_boundary_distance(); break;
        case ORC_STAGE_NEIGHBORHOOD:
        {
            const float scale = (s->P.basic_pbf || st.mBaseKernelWidthOnBoundaryDistance) ? 1.0f : 1.5f;
            if (s->nb_method == 1)
                s->neighborhood_bs(scale);
            else if (s->nb_method == 2)
                s->neighborhood_bf(scale);
            else
                s->neighborhood(scale);
            break;
        }
        case ORC_STAGE_SPREAD: s->spread_kernel_width(); break;
        case ORC_STAGE_BOX: s->box_collision(); break;
        case ORC_STAGE_SPHERE: s->sphere_collision(); break;
        case ORC_STAGE_INCOMPRESSIBILITY: s->incompressibility(); break;
        case ORC_STAGE_UPDATE_TRANSFERS: s->update_transfers(); break;
        default: break;
    }
}

void orc_update(void* h, float dt) { static_cast<Scene*>(h)->update(dt); }
void orc_set_solver_tolerance(void* h, float tol, int max_iters) {
    Scene* s = static_cast<Scene*>(h);
    s->tol = tol > 0.0f ? tol : 0.0f;
    s->tol_max_iters = tol > 0.0f ? max_iters : 0;
}
void orc_set_warm_lambdas(void* h, const float* lam, uint32_t n) {
    Scene* s = static_cast<Scene*>(h);
    s->warm_lambda.assign(lam, lam + n);
}
void orc_solver_stats(void* h, uint64_t* iters, uint64_t* substeps, uint64_t* fallbacks) {
    Scene* s = static_cast<Scene*>(h);
    *iters = s->stat_iters;
    *substeps = s->stat_substeps;
    *fallbacks = s->stat_fallbacks;
}

void orc_density_error(void* h, double* mean_err, double* max_err) {
    static_cast<Scene*>(h)->density_error(mean_err, max_err);
}

void orc_stable_radix_sort(const uint32_t* keys, const uint32_t* values, uint32_t n,
                           uint32_t upper_bound, uint32_t* out_keys, uint32_t* out_values) {
    std::vector<uint32_t> k, v;
    radix_sort(keys, values, n, upper_bound, k, v);
    std::memcpy(out_keys, k.data(), 4 * n);
    std::memcpy(out_values, v.data(), 4 * n);
}

void orc_prefix_sum(uint32_t* values, uint32_t n) {  // prefix_sum_apply_on_block_level + spread
    uint32_t run = 0;
    for (uint32_t i = 0; i < n; ++i) { run += values[i]; values[i] = run; }
}

void orc_position_hash(const int32_t* position, uint32_t n, const float* mn, const float* mx,
                       uint32_t res, uint32_t dims, uint32_t* out) {
    Grid g{v3(mn[0], mn[1], mn[2]), v3(mx[0], mx[1], mx[2]), res, dims};
    const I4* p = reinterpret_cast<const I4*>(position);
    for (uint32_t id = 0; id < n; ++id) {
        uint32_t c[3];
        g.map(pos_world(p[id]), c);
        out[id] = g.hash(c);
    }
}

void orc_position_code(const int32_t* position, uint32_t n, uint32_t section, uint32_t* out) {
    const I4* p = reinterpret_cast<const I4*>(position);
    for (uint32_t id = 0; id < n; ++id) out[id] = position_code(p[id], section);
}

void orc_find_value_ranges(const uint32_t* sorted_keys, uint32_t n, uint32_t n_cells,
                           uint32_t* start, uint32_t* end) {
    std::vector<uint32_t> io(n);
    for (uint32_t i = 0; i < n; ++i) io[i] = i;
    for (uint32_t c = 0; c < n_cells; ++c) start[c] = end[c] = 0u;
    Scene::find_value_ranges(io.data(), sorted_keys, n, start, end);
}

uint32_t orc_neighborhood_green(const int32_t* position, const float* range, uint32_t n,
                                const uint32_t* cs, const uint32_t* ce, float range_scale,
                                const float* mn, const float* mx, uint32_t res, uint32_t dims,
                                uint32_t* pairs, uint32_t cap) {
    Scene s;
    std::memset(&s.P, 0, sizeof(s.P));
    s.P.cap_pairs = cap;
    s.position.resize(n);
    std::memcpy(s.position.data(), position, sizeof(I4) * n);
    s.kernel_width.assign(range, range + n);
    s.fidx.resize(n);
    for (uint32_t i = 0; i < n; ++i) s.fidx[i] = i;
    Grid g{v3(mn[0], mn[1], mn[2]), v3(mx[0], mx[1], mx[2]), res, dims};
    s.green(g, cs, ce, range_scale);
    std::memcpy(pairs, s.pairs.data(), 8 * s.pairs.size());
    return (uint32_t)s.pairs.size();
}

uint32_t orc_neighborhood_brute_force(const int32_t* position, const float* range, uint32_t n,
                                      float range_scale, uint32_t* pairs, uint32_t cap) {
    // neighborhood_brute_force.comp:31-50
    const I4* p = reinterpret_cast<const I4*>(position);
    uint32_t len = 0;
    for (uint32_t id = 0; id < n; ++id) {
        float r = range[id] * range_scale;
        V3 pos = pos_world(p[id]);
        for (uint32_t idN = 0; idN < n; ++idN) {
            V3 posN = pos_world(p[idN]);
            if (id == idN || length(sub(pos, posN)) > r) continue;
            if (len < cap) { pairs[2 * len] = id; pairs[2 * len + 1] = idN; ++len; }
        }
    }
    return len;
}

uint32_t orc_apply_hidden_edit(const uint32_t* index_list, uint32_t n_index, uint32_t n_hidden,
                               const uint32_t* edit, uint32_t n_edit, uint32_t max_new,
                               uint32_t* new_index, uint32_t* new_edit) {
    std::vector<uint32_t> idx(index_list, index_list + n_index), e(edit, edit + n_edit), ni, ne;
    uint32_t n = hidden_edit(idx, n_hidden, e, max_new, ni, ne);
    std::memcpy(new_index, ni.data(), 4 * n);
    std::memcpy(new_edit, ne.data(), 4 * n);
    return n;
}

uint32_t orc_delete_edit_list(const uint32_t* deleted, uint32_t n_deleted, uint32_t n_hidden,
                              uint32_t* edit) {
    std::vector<uint32_t> d(deleted, deleted + n_deleted);
    std::vector<uint32_t> e = delete_edit(d, n_hidden);
    std::memcpy(edit, e.data(), 4 * e.size());
    return (uint32_t)e.size();
}

float orc_exp(float x) { return exp_f(x); }
float orc_pow(float x, float y) { return pow_f(x, y); }

void orc_kernel(int kernel_id, const float* r3, float kw, uint32_t dims, float* height, float* grad3) {
    const float D = dims == 2u ? 2.0f : 3.0f;
    KernelEval ke{kernel_id, kernel_id, D};
    Gauss g = gauss_of_width(kw, D);
    V3 gr;
    ke.eval(v3(r3[0], r3[1], r3[2]), kw, g, height, &gr, true, true);
    grad3[0] = gr.x;
    grad3[1] = gr.y;
    grad3[2] = gr.z;
}

}  // extern "C"
