// Launch wrappers of the plain-substep kernels (definitions in plain_kernels.hip).
#pragma once
#include "scene_kernels.h"

namespace apbf {

// One buffer set of the particle SoA (hidden particles + fluid attributes,
// list_definitions.h:9-13) indexed by particle slot.
struct ParticleArrays {
    int4* pos;
    float4* vel;
    float* inv_mass;
    float* radius;
    int4* backup;
    uint32_t* transferring;
    float* target_radius;
    float* kernel_width;
    float* boundariness;
    uint32_t* boundary_distance;
    float* lam_prev;     // total lambda of the previous substep (warm-start feature); may be null
    uint32_t* prev_old;  // optional: receives the source slot of every gathered particle
    uint32_t* sid;       // scene index of every particle in a packed batch; null for a single scene
};

void launch_integrate(int4* pos, int4* backup, float4* vel, const uint32_t* counts, uint32_t cap,
                      float last_dt, float dt, const float acc[3], hipStream_t s);
void launch_copy_int4(int4* dst, const int4* src, const uint32_t* counts, uint32_t cap, hipStream_t s);
void launch_hash_iota(const int4* pos, const uint32_t* d_len, uint32_t cap, const GridDesc& g,
                      uint32_t* keys, uint32_t* vals, hipStream_t s, const SegmentTable* seg = nullptr);
// seg_start[sid[i]] = i where the (cell-sorted) scene index changes.
void launch_segment_starts(const uint32_t* sid, const uint32_t* counts, uint32_t cap, uint32_t* seg_start,
                           hipStream_t s);
void launch_position_code(const int4* pos, const uint32_t* d_len, uint32_t cap, uint32_t section,
                          uint32_t* out, hipStream_t s);
void launch_permute(const uint32_t* perm, const uint32_t* counts, uint32_t cap,
                    const ParticleArrays& src, const ParticleArrays& dst, hipStream_t s);
void launch_ranges(const uint32_t* keys, const uint32_t* d_len, uint32_t cap, uint32_t* start,
                   uint32_t* end, hipStream_t s);
// Row-group neighbour walk (COUNT / FILL passes), optionally fused with the
// kernel-width spreading and its short-list filter.
// mode 0 count, 1 fill (exact, clamped; after a mode-2 walk `gate` is the
// count block, which selects all rows, the rows longer than `stride`, or
// none), 2 single pass into fixed-stride slots (see walk.hip).
void launch_walk(int mode, bool spread, const int4* pos, const float* kw, const float* radius, const float* trad,
                 int on_target, const uint32_t* cs, const uint32_t* ce, uint32_t* counts, uint32_t cap,
                 float range_scale, const GridDesc& g, uint32_t pcap, uint32_t* cnt_raw, uint32_t* cnt_f,
                 const uint32_t* row_raw, const uint32_t* row_f, uint32_t* col, uint32_t* ukw, hipStream_t s,
                 int lanes = 8, uint32_t stride = 0, const uint32_t* gate = nullptr,
                 const SegmentTable* seg = nullptr, unsigned long long* cand = nullptr);
void launch_walk_decide(uint32_t* counts, uint32_t pcap, hipStream_t s);
void launch_brute(bool fill, const int4* pos, const float* kw, const uint32_t* counts, uint32_t cap,
                  float range_scale, uint32_t pcap, uint32_t* cnt, const uint32_t* row, uint32_t* col,
                  hipStream_t s);
void launch_pack_rows(const uint32_t* counts, uint32_t cap, const uint32_t* row, const uint32_t* slots,
                      uint32_t pcap, uint32_t* col, hipStream_t s);
void launch_fix_clamped_rows(const uint32_t* counts, uint32_t cap, uint32_t pcap, uint32_t* row_f, hipStream_t s);
void launch_pairs_total(uint32_t* counts, const uint32_t* row, uint32_t pair_cap, hipStream_t s);
void launch_set_n_plus_1(uint32_t* counts, hipStream_t s);
void launch_kernel_consts(const float* kw, const float* radius, const float* inv_mass, const uint32_t* counts,
                          uint32_t cap, int hid, int gid, float dim, float4* gc,
                          hipStream_t s);
void launch_sphere(int4* pos, const float* radius, const uint32_t* counts, uint32_t cap, const BoxSet& c,
                   hipStream_t s);
void launch_box(int4* pos, const float* radius, const uint32_t* counts, uint32_t cap,
                const BoxSet& boxes, hipStream_t s, const SegmentTable* seg = nullptr);
void launch_count_particle_substeps(const uint32_t* counts, unsigned long long* acc, hipStream_t s);
void launch_density_error_finish(const double* part_sum, const float* part_max, uint32_t blocks,
                                 uint32_t* counts, double* stats, hipStream_t s);

// rows.hip: kernels over CSR rows with a group of lanes per row
void launch_solve_rows(const int4* pos, const float* inv_mass, const float* radius, const float* kw,
                       const float4* gc, const uint32_t* counts, uint32_t cap, const uint32_t* row,
                       const uint32_t* col, uint32_t pcap, const SolverArgs& a, float* boundariness,
                       float* lambda, int4* acc, hipStream_t s, int lanes = 16, bool long_rows = false,
                       const uint32_t* long_list = nullptr, const uint32_t* long_list2 = nullptr);
void launch_long_rows(uint32_t* counts, uint32_t cap, const uint32_t* row, uint32_t pcap, int lanes, uint32_t* list,
                      uint32_t* list2, hipStream_t s);
void launch_features_rows(const int4* pos, const float4* vel, const float* inv_mass, const float* radius,
                          const float* kw, const float4* gc, const float* boundariness, const float* lam_prev,
                          const uint32_t* counts, uint32_t cap, const uint32_t* row, const uint32_t* col,
                          uint32_t pcap, const SolverArgs& a, float dt, float* feat, hipStream_t s);
void launch_solve_apply(int4* pos, int4* acc, const float* radius, const uint32_t* counts, uint32_t cap,
                        const BoxSet& boxes, hipStream_t s, const uint32_t* action = nullptr,
                        const int4* saved = nullptr, const SegmentTable* seg = nullptr);
void launch_solve_begin(uint32_t* counts, double* tol_err, uint32_t slots, hipStream_t s);
void launch_solve_decide(uint32_t* counts, const double* tol_err, uint32_t k, float tol, bool warm, hipStream_t s);
void launch_solve_stats(const uint32_t* counts, unsigned long long* stats, uint32_t fixed_iterations, hipStream_t s);
uint32_t density_rows_blocks(uint32_t cap);
void launch_density_rows(const int4* pos, const float* inv_mass, const float* radius, const float* kw,
                         const float4* gc, const uint32_t* counts, uint32_t cap, const uint32_t* row,
                         const uint32_t* col, uint32_t pcap, int hid, int gid, float dim, double* part_sum,
                         float* part_max, hipStream_t s);
void launch_fsm12_density_rows(const int4* pos, const float* inv_mass, const float* kw, const float4* gc,
                               const uint32_t* counts, uint32_t cap, const uint32_t* row, const uint32_t* col,
                               uint32_t pcap, const uint32_t* old_bd, uint32_t* bd, uint32_t* nearest, int hid,
                               int gid, float dim, double* part_sum, float* part_max, hipStream_t s);
void launch_fsm12_rows(const int4* pos, const uint32_t* counts, uint32_t cap, const uint32_t* row,
                       const uint32_t* col, uint32_t pcap, const uint32_t* old_bd, uint32_t* bd, uint32_t* nearest,
                       hipStream_t s);
void launch_spread_rows(bool fill, const int4* pos, const float* radius, const float* trad, const float* kw_old,
                        const uint32_t* counts, uint32_t cap, const uint32_t* row, const uint32_t* col, uint32_t pcap,
                        int on_target, uint32_t* ukw, uint32_t* cnt2, const uint32_t* row2, uint32_t* col2,
                        hipStream_t s);

}  // namespace apbf
