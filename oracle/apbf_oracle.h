/* Test infrastructure only: C interface of the scalar CPU oracle.
 *
 * The oracle is a sequential restatement of APBF's per-substep particle
 * pipeline (reference: the shaders/ compute shaders and the host orchestration in
 * source/pool.cpp, waterdrop.cpp, neighborhood_green.cpp, incompressibility.cpp,
 * spread_kernel_width.cpp, update_transfers.cpp, particle_transfer.cpp,
 * velocity_handling.cpp, box_collision.cpp, indexed_list.h, gpu_list.h,
 * algorithms.cpp). Every compute-shader invocation is executed in ascending
 * gl_GlobalInvocationID order, which is one legal schedule of the reference;
 * atomics therefore resolve in id order.
 *
 * Only tests/, the smoke check and the CPU-baseline leg of the benchmark load
 * this library; the product never does.
 *
 * Pinning: list algorithms against the reference's test.cpp vectors; hash,
 * codes, ranges and neighbour sets against closed forms and brute force;
 * exp/pow against libm; the substep stages by invariants and closed forms
 * (tests/test_oracle_invariants.py); the split/merge find (flood fill,
 * nearest neighbour, target radius, decay, predicates, maximal claims) and
 * the kernel width from boundary distance by brute force over the pair list
 * and per-particle closed forms, each shown to fail on a planted mistake
 * (tests/test_oracle_pins.py). Not pinnable without a reference run: the
 * rounding of GLSL built-ins in the real driver and which contended
 * candidate its scheduler lets win.
 */
#ifndef APBF_ORACLE_H
#define APBF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field order and meaning of shaders/cpu_gpu_shared_config.h:74-94. */
typedef struct orc_settings {
    int32_t mHeightKernelId;
    int32_t mGradientKernelId;
    int32_t mMerge;
    int32_t mSplit;
    int32_t mBaseKernelWidthOnTargetRadius;
    int32_t mBaseKernelWidthOnBoundaryDistance;
    int32_t mUpdateTargetRadius;
    int32_t mUpdateBoundariness;
    int32_t mNeighborListSorted;
    int32_t mBoundarinessCalculationMethod;
    float mBoundarinessAdaptionSpeed;
    float mKernelWidthAdaptionSpeed;
    float mBoundarinessSelfGradLengthFactor;
    float mBoundarinessUnderpressureFactor;
    float mMergeDuration;
    float mSmallestTargetRadius;
    float mTargetRadiusOffset;
    float mTargetRadiusScaleFactor;
} orc_settings;

#define ORC_MAX_BOXES 16

/* Host-side knobs of source/settings.cpp:5-31 that are not in the UBO, plus the
 * scene wiring of pool.cpp / waterdrop.cpp. */
typedef struct orc_scene_params {
    orc_settings s;
    int32_t basic_pbf;             /* settings::basicPbf */
    int32_t solver_iterations;     /* settings::solverIterations */
    float split_duration;          /* settings::splitDuration */
    float gravity[3];              /* velocity_handling::set_acceleration */
    float hash_min[3];             /* neighborhood_green::set_position_range */
    float hash_max[3];
    uint32_t hash_res_log2;
    uint32_t n_boxes;              /* user_controlled_boxes box_min / box_max lists */
    float box_min[ORC_MAX_BOXES][4];
    float box_max[ORC_MAX_BOXES][4];
    uint32_t cap_particles;        /* hidden particle / fluid requested length */
    uint32_t cap_transfers;        /* MAX_TRANSFERS */
    uint32_t cap_pairs;            /* neighbour list requested length */
    uint32_t dimensions;           /* DIMENSIONS (cpu_gpu_shared_config.h:2): 2 or 3; 0 means 3 */
} orc_scene_params;

void* orc_create(const orc_scene_params* p);
void orc_destroy(void* h);

/* State exchange. Particle arrays are indexed by hidden slot; fluid arrays by
 * fluid id; `fluid_index` maps fluid id -> hidden slot. position/pos_backup/
 * velocity hold 4 components per particle (w kept as given). */
int orc_set_state(void* h, uint32_t n_hidden, const int32_t* position, const float* velocity,
                  const float* inverse_mass, const float* radius, const int32_t* pos_backup,
                  const uint32_t* transferring, uint32_t n_fluid, const uint32_t* fluid_index,
                  const float* target_radius, const float* kernel_width, const float* boundariness,
                  const uint32_t* boundary_distance, uint32_t n_transfers, const uint32_t* t_source,
                  const uint32_t* t_target, const float* t_time_left, float last_delta_time);
void orc_get_counts(void* h, uint32_t* n_hidden, uint32_t* n_fluid, uint32_t* n_transfers,
                    uint32_t* n_pairs);
void orc_get_state(void* h, int32_t* position, float* velocity, float* inverse_mass, float* radius,
                   int32_t* pos_backup, uint32_t* transferring, uint32_t* fluid_index,
                   float* target_radius, float* kernel_width, float* boundariness,
                   uint32_t* boundary_distance, uint32_t* t_source, uint32_t* t_target,
                   float* t_time_left, uint32_t* pairs, float* last_delta_time);
/* Extra per-fluid-id outputs of the last solver iteration (lambda) and of the
 * last update_transfers pass (nearest neighbour, split list). */
void orc_get_lambda(void* h, float* lambda);
uint32_t orc_get_splits(void* h, uint32_t* split_list);

/* Stages of one substep, in the order of pool.cpp:67-106. */
enum {
    ORC_STAGE_VELOCITY = 0,         /* velocity_handling::apply */
    ORC_STAGE_PARTICLE_TRANSFER = 1,/* particle_transfer::apply */
    ORC_STAGE_KERNEL_WIDTH_BD = 2,  /* uint_to_float_with_indexed_lower_bound (pool.cpp:77-80) */
    ORC_STAGE_NEIGHBORHOOD = 3,     /* neighborhood_green::apply */
    ORC_STAGE_SPREAD = 4,           /* spread_kernel_width::apply */
    ORC_STAGE_BOX = 5,              /* box_collision::apply */
    ORC_STAGE_INCOMPRESSIBILITY = 6,/* incompressibility::apply */
    ORC_STAGE_UPDATE_TRANSFERS = 7, /* update_transfers::apply */
    ORC_STAGE_SPHERE = 8            /* sphere_collision::apply */
};
/* 0: neighborhood_green (default), 1: neighborhood_binary_search,
 * 2: neighborhood_brute_force. */
void orc_set_neighborhood_method(void* scene, int method);
/* One sphere collider (sphere_collision::set_sphere); enable = 0 removes it. */
void orc_set_sphere(void* scene, const float* center3, float radius, int hollow, int enable);
void orc_stage(void* h, int stage, float dt);
/* Intermediates of the last update_transfers stage, per fluid id: boundary
 * distance after the flood fill (find_split_and_merge_1.comp:17-32), shortest
 * neighbour distance and nearest neighbour (find_split_and_merge_2.comp:17-33).
 * Any pointer may be NULL; returns the fluid count they cover. */
uint32_t orc_get_fsm(void* h, uint32_t* bd_fill, uint32_t* min_dist, uint32_t* nearest);
/* One full substep (pool::update / waterdrop::update with the time machine off). */
void orc_update(void* h, float dt);
/* Library extensions of the solver loop (no counterpart in the reference):
 * tolerance mode (tol > 0: up to max_iters iterations, stopping at the first
 * whose starting mean density error is <= tol; tol <= 0: solverIterations),
 * the warm-start shift with the given lambdas (per fluid id at solver time,
 * used by the next update only) judged by the first plain iteration, and the
 * sums of applied iterations, substeps and undone warm starts. */
void orc_set_solver_tolerance(void* h, float tol, int max_iters);
void orc_set_warm_lambdas(void* h, const float* lambda, uint32_t n);
void orc_solver_stats(void* h, uint64_t* iters, uint64_t* substeps, uint64_t* fallbacks);
/* Mean and maximum per-particle compression error max(0, rho/rho0 - 1) of the
 * current positions, from a density pass that leaves the state untouched. */
void orc_density_error(void* h, double* mean_err, double* max_err);

/* Unit entry points for the list algorithms and single kernels. */
void orc_stable_radix_sort(const uint32_t* keys, const uint32_t* values, uint32_t n,
                           uint32_t upper_bound, uint32_t* out_keys, uint32_t* out_values);
void orc_prefix_sum(uint32_t* values, uint32_t n);
void orc_position_hash(const int32_t* position, uint32_t n, const float* min3, const float* max3,
                       uint32_t res_log2, uint32_t dimensions, uint32_t* out_hash);
void orc_position_code(const int32_t* position, uint32_t n, uint32_t section, uint32_t* out_code);
void orc_find_value_ranges(const uint32_t* sorted_keys, uint32_t n, uint32_t n_cells,
                           uint32_t* start, uint32_t* end);
/* Green neighbour search (identity index list); returns the pair count
 * (clamped to cap) and writes the pairs in sequential order. */
uint32_t orc_neighborhood_green(const int32_t* position, const float* range, uint32_t n,
                                const uint32_t* cell_start, const uint32_t* cell_end,
                                float range_scale, const float* min3, const float* max3,
                                uint32_t res_log2, uint32_t dimensions, uint32_t* pairs, uint32_t cap);
uint32_t orc_neighborhood_brute_force(const int32_t* position, const float* range, uint32_t n,
                                      float range_scale, uint32_t* pairs, uint32_t cap);
/* Index-list rebuild after a hidden-list edit (indexed_list.h:289-308); returns
 * the new length and writes the new index list and the owner edit list. */
uint32_t orc_apply_hidden_edit(const uint32_t* index_list, uint32_t n_index, uint32_t n_hidden,
                               const uint32_t* edit, uint32_t n_edit, uint32_t max_new,
                               uint32_t* new_index, uint32_t* new_edit);
/* Stable compaction edit list of delete_these (indexed_list.h:126-138). */
uint32_t orc_delete_edit_list(const uint32_t* deleted, uint32_t n_deleted, uint32_t n_hidden,
                              uint32_t* edit);
/* Scalar math helpers (pinned against libm in the tests). */
float orc_exp(float x);
float orc_pow(float x, float y);
void orc_kernel(int kernel_id, const float* r3, float kernel_width, uint32_t dimensions, float* height,
                float* grad3);

#ifdef __cplusplus
}
#endif
#endif
