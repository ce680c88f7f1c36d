/* apbf.h — C ABI of the MI355X (gfx950) implementation of APBF's per-substep
 * particle pipeline.
 *
 * The reference (cg-tuwien/APBF) drives its GLSL compute shaders through the
 * static `shader_provider` class (source/shader_provider.h:9-76), which takes
 * device buffer handles plus separate 4-byte device buffers for every list
 * length, and through the list containers and subsystem `apply()` calls above
 * it (source/gpu_list.h, indexed_list.h, algorithms.h, neighborhood_green.h,
 * incompressibility.h, spread_kernel_width.h, update_transfers.h,
 * particle_transfer.h, velocity_handling.h, box_collision.h, pool.h).
 * Every entry point below replaces one of those interfaces (cited per
 * function). Conventions, following the reference:
 *  - pointers are device pointers unless the name says `host`;
 *  - list lengths live in device memory (`d_len`, u32) so nothing reads back
 *    to the host inside a substep; `capacity` is the allocated element count;
 *  - positions are int32 fixed point, 2^18 units per world unit, 4 components
 *    per particle (x, y, z, w; w is carried, never interpreted);
 *  - every call is enqueued on `stream` (0 = the null stream) and returns
 *    immediately; no call allocates device memory unless documented;
 *  - every call returns an apbf_status and never throws.
 */
#ifndef APBF_H
#define APBF_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t apbf_status;
#define APBF_OK 0
#define APBF_ERR_INVALID_ARGUMENT -1
#define APBF_ERR_HIP -2
#define APBF_ERR_CAPACITY -3
#define APBF_ERR_NO_DEVICE -4
#define APBF_ERR_WORKSPACE -5

typedef void* apbf_stream; /* a hipStream_t */

/* Uniform-block settings, field order and meaning of
 * shaders/cpu_gpu_shared_config.h:74-94 (defaults in source/settings.cpp:5-31). */
typedef struct apbf_settings {
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
} apbf_settings;

#define APBF_MAX_BOXES 16

/* Scene wiring of source/pool.cpp:6-56 and source/waterdrop.cpp:6-58 plus the
 * host-only knobs of source/settings.cpp (basicPbf, solverIterations,
 * splitDuration). Capacities are the reference's requested lengths. */
typedef struct apbf_scene_params {
    apbf_settings s;
    int32_t basic_pbf;
    int32_t solver_iterations;
    float split_duration;
    float gravity[3];
    float hash_min[3];
    float hash_max[3];
    uint32_t hash_res_log2;
    uint32_t n_boxes;
    float box_min[APBF_MAX_BOXES][4];
    float box_max[APBF_MAX_BOXES][4];
    uint32_t cap_particles;
    uint32_t cap_transfers;
    uint32_t cap_pairs;
    /* DIMENSIONS (cpu_gpu_shared_config.h:2): 3, or 2 for the reference's
     * default planar build (one particle layer in z, 2D z-curve hash,
     * 2D kernel normalisations, rest density and split/merge volumes);
     * 0 means 3. */
    uint32_t dimensions;
} apbf_scene_params;

/* ------------------------------------------------------------------ scene */
/* A scene owns the particle state of one simulation on the current device:
 * the hidden particle SoA, the fluid attributes, the transfer list and the
 * neighbour lists (list_definitions.h:9-20), plus all temporaries of a substep
 * (allocated here, once; the substep itself allocates nothing).
 * Replaces the scene classes pool / waterdrop (pool.h:14-56, waterdrop.h). */
apbf_status apbf_scene_create(const apbf_scene_params* params, void** out_scene);
apbf_status apbf_scene_destroy(void* scene);
/* Later calls on the scene run on `stream`; work already queued on the old
 * stream completes before them (an event, no host synchronisation). */
apbf_status apbf_scene_set_stream(void* scene, apbf_stream stream);

/* Upload a state from host arrays (layout as the reference lists: position,
 * velocity and pos_backup 4 x 32 bit per particle). The fluid list must cover
 * every hidden particle once (fluid_index a permutation); it is stored with the
 * particles reordered so that fluid id == hidden slot. */
apbf_status apbf_scene_set_state(void* scene, uint32_t n, const int32_t* host_position,
                                 const float* host_velocity, const float* host_inverse_mass,
                                 const float* host_radius, const int32_t* host_pos_backup,
                                 const uint32_t* host_transferring, const uint32_t* host_fluid_index,
                                 const float* host_target_radius, const float* host_kernel_width,
                                 const float* host_boundariness,
                                 const uint32_t* host_boundary_distance, uint32_t n_transfers,
                                 const uint32_t* host_t_source, const uint32_t* host_t_target,
                                 const float* host_t_time_left, float last_delta_time);
/* Device-resident counts, copied to the host (synchronises the stream). */
apbf_status apbf_scene_get_counts(void* scene, uint32_t* n_particles, uint32_t* n_transfers,
                                  uint32_t* n_pairs);
/* Download the state (synchronises). `host_pairs` receives the neighbour pairs
 * (fluid id, neighbour fluid id) ordered by first id, then by the order of the
 * reference's cell walk (neighborhood_green.comp:69-86). Any pointer may be NULL. */
apbf_status apbf_scene_get_state(void* scene, int32_t* host_position, float* host_velocity,
                                 float* host_inverse_mass, float* host_radius,
                                 int32_t* host_pos_backup, uint32_t* host_transferring,
                                 float* host_target_radius, float* host_kernel_width,
                                 float* host_boundariness, uint32_t* host_boundary_distance,
                                 uint32_t* host_t_source, uint32_t* host_t_target,
                                 float* host_t_time_left, uint32_t* host_pairs,
                                 float* host_last_delta_time);
apbf_status apbf_scene_get_lambda(void* scene, float* host_lambda);

/* Stages of one substep in the order of pool.cpp:67-106 (same numbering as the
 * subsystem apply() calls they replace). */
#define APBF_STAGE_VELOCITY 0          /* velocity_handling::apply (velocity_handling.cpp:15-31) */
#define APBF_STAGE_PARTICLE_TRANSFER 1 /* particle_transfer::apply (particle_transfer.cpp:10-28) */
#define APBF_STAGE_KERNEL_WIDTH_BD 2   /* uint_to_float_with_indexed_lower_bound (pool.cpp:77-80) */
#define APBF_STAGE_NEIGHBORHOOD 3      /* neighborhood_green::apply (neighborhood_green.cpp:27-77) */
#define APBF_STAGE_SPREAD 4            /* spread_kernel_width::apply (spread_kernel_width.cpp:12-26) */
#define APBF_STAGE_BOX 5               /* box_collision::apply (box_collision.cpp:12-24) */
#define APBF_STAGE_INCOMPRESSIBILITY 6 /* incompressibility::apply (incompressibility.cpp:12-45) */
#define APBF_STAGE_UPDATE_TRANSFERS 7  /* update_transfers::apply (update_transfers.cpp:14-71) */
#define APBF_STAGE_SPHERE 8            /* sphere_collision::apply (sphere_collision.comp:42-60) */
apbf_status apbf_scene_stage(void* scene, int32_t stage, float dt);
/* One substep (pool::update / waterdrop::update, pool.cpp:67-106). The solver
 * loop runs box collision + incompressibility solverIterations times. */
apbf_status apbf_scene_update(void* scene, float dt);
/* Mean and maximum of max(0, rho_i/rho0_i - 1) over the current positions (a
 * density pass like incompressibility_0/1 that leaves the state untouched).
 * Needs a neighbour list that refers to the current particle slots (after the
 * neighbourhood stage and before particle_transfer / update_transfers edit the
 * particle list); returns APBF_ERR_INVALID_ARGUMENT otherwise. Synchronises. */
apbf_status apbf_scene_density_error(void* scene, double* mean_err, double* max_err);
/* Density-error statistics accumulated inside apbf_scene_update: when enabled,
 * every substep ends its solver loop with a density pass (as above) whose mean
 * and max are added to device-resident accumulators, without synchronising.
 * apbf_scene_error_stats returns the mean over substeps of the per-substep mean
 * error max(0, rho/rho0 - 1), the maximum error, the last substep's mean, the
 * mean over substeps of the mean absolute deviation |rho/rho0 - 1| and the
 * substep count, and clears them when `reset` is non-zero (synchronises). */
apbf_status apbf_scene_set_error_measurement(void* scene, int32_t enable);
apbf_status apbf_scene_error_stats(void* scene, double* mean_of_means, double* max_err,
                                   double* last_mean, double* mean_abs_of_means, uint32_t* substeps,
                                   int32_t reset);
/* With error measurement on, the solver also records the mean density error
 * of the positions each iteration starts from (its own density sums, no
 * extra pass): out[k] = mean over substeps of that error before iteration k,
 * for k < max_iters (iterations beyond the solver's count read 0). The error
 * after the last iteration is apbf_scene_error_stats. Synchronises; `reset`
 * clears the record (the substep count is cleared by apbf_scene_error_stats). */
apbf_status apbf_scene_iteration_errors(void* scene, double* out, uint32_t max_iters, int32_t reset);
/* Tolerance mode of the solver loop (pool.cpp:92-95 runs a fixed
 * solverIterations; no counterpart in the reference): with tolerance > 0 each
 * substep runs up to max_iterations iterations (box collision +
 * incompressibility), and an iteration whose starting positions have a mean
 * density error max(0, rho/rho0 - 1) <= tolerance applies no shifts and ends
 * the solve (its boundariness update still happens, as in incompressibility_2);
 * tolerance <= 0 returns to the fixed count. The decisions are device words,
 * so substeps stay graph-capturable. With the learned warm start (mode 2) the
 * plain solver also judges the warm shift: when the error after it exceeds
 * the error before it, the positions from before it are put back and the
 * plain iterations continue from there. Drops a captured graph. */
apbf_status apbf_scene_set_solver_tolerance(void* scene, float tolerance, int32_t max_iterations);
/* Sums over substeps of the solver iterations whose shifts were applied, the
 * substeps, and the warm starts that were undone; cleared when `reset` is
 * non-zero. Synchronises. */
apbf_status apbf_scene_solver_stats(void* scene, uint64_t* iterations, uint64_t* substeps, uint64_t* fallbacks,
                                    int32_t reset);
/* Sums over substeps of the number of particles the solver ran on (the unit of
 * the particle-substeps/s metric) and of the neighbour pairs, kept on the
 * device; cleared when `reset` is non-zero. Synchronises. */
apbf_status apbf_scene_particle_substeps(void* scene, uint64_t* total, uint64_t* pairs_total,
                                         int32_t reset);
/* Time every launch of one kernel with HIP events on the scene stream:
 * kernel_id 0 solver row pass, 1 solver apply pass, 2 neighbour count walk, 3 neighbour fill
 * walk, 4 kernel-width spreading, 5 boundary-distance flood fill, 6 particle
 * permutation; -1 stops timing. At most max_launches launches are recorded.
 * apbf_scene_kernel_time returns the summed duration and the launch count of
 * the recorded launches and restarts the recording. Both synchronise. */
apbf_status apbf_scene_time_kernel(void* scene, int32_t kernel_id, uint32_t max_launches);
apbf_status apbf_scene_kernel_time(void* scene, double* total_ms, uint32_t* launches);
/* The same for several kernels at once (bit k of kernel_mask: kernel id k);
 * apbf_scene_kernel_time_of returns one kernel's share of the recording
 * (kernel_id -1: all) without restarting it. Both synchronise. */
apbf_status apbf_scene_time_kernels(void* scene, uint32_t kernel_mask, uint32_t max_launches);
apbf_status apbf_scene_kernel_time_of(void* scene, int32_t kernel_id, double* total_ms, uint32_t* launches);
/* Sums over substeps, kept on the device: out[0] particles the solver ran on,
 * out[1] its neighbour pairs, out[2] the neighbour walk's pairs before the
 * kernel-width spreading filtered them (equal to out[1] without spreading),
 * out[3] particles created by splits, out[4] particles deleted when merges
 * finished. Cleared when `reset` is non-zero (with
 * apbf_scene_particle_substeps' reset too). Synchronises. */
apbf_status apbf_scene_substep_totals(void* scene, uint64_t* out, uint32_t n_out, int32_t reset);
/* Copies the first n_words words of the scene's device-resident count block:
 * 0 particles, 2 neighbour pairs, 3 transfers, 4 splits, 6 overflow flags,
 * 7 neighbour-walk pairs before the capacity clamp, 14 particles covered by
 * the neighbour rows, 17 largest pre-clamp pair count seen, 22 whether rows
 * outgrew their slots in the last single-pass walk, 27 the longest of them,
 * 28-29 (low, high word) the candidates counted since
 * apbf_scene_count_walk_candidates enabled counting. Synchronises. */
apbf_status apbf_scene_diagnostics(void* scene, uint32_t* host_counts, uint32_t n_words);
/* Diagnostic with no counterpart in the reference: while enabled, every
 * neighbourhood build also runs the counting walk, which adds the number of
 * candidates the walk tests (the particles of the visited, unculled cells)
 * to diagnostics words 28-29. Enabling or disabling clears the sum. Costs
 * one more walk per substep; eager updates only (a graph captured before
 * keeps its launches). */
apbf_status apbf_scene_count_walk_candidates(void* scene, int32_t enable);
/* Overflow flags raised by capacity clamps since the last call (bit 0 pairs,
 * bit 1 transfers, bit 2 particles); cleared by the call. Synchronises. */
apbf_status apbf_scene_overflow_flags(void* scene, uint32_t* flags);

/* ------------------------------------------------------ list algorithms */
/* Workspace sizes (bytes) for the primitives below. */
size_t apbf_prefix_sum_workspace_size(uint32_t capacity);
size_t apbf_radix_sort_workspace_size(uint32_t capacity);

/* Inclusive prefix sum in place over d_len[0] values (algorithms::prefix_sum,
 * algorithms.cpp:93-118). */
apbf_status apbf_prefix_sum(uint32_t* values, const uint32_t* d_len, uint32_t capacity,
                            void* workspace, size_t workspace_bytes, apbf_stream stream);

/* Stable LSD radix sort of (key, value) pairs by key (algorithms::sort,
 * algorithms.cpp:59-91). Digits are processed while upper_bound >> offset != 0,
 * as in the reference; keys must be < upper_bound for a full sort. Inputs are
 * left unchanged; outputs must not alias inputs. */
apbf_status apbf_radix_sort(const uint32_t* keys, const uint32_t* values, const uint32_t* d_len,
                            uint32_t capacity, uint32_t upper_bound, uint32_t* out_keys,
                            uint32_t* out_values, void* workspace, size_t workspace_bytes,
                            apbf_stream stream);

/* The sort of apbf_radix_sort without a workspace argument and without any
 * allocation: no counterpart in the reference. For callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Its workspace is a fixed device array
 * inside the library, one per GPU, held until the process ends
 * (apbf_radix_sort_workspace_size(APBF_RADIX_SORT_NOALLOC_MAX) bytes, about
 * 8.5 MiB). capacity above APBF_RADIX_SORT_NOALLOC_MAX returns
 * APBF_ERR_CAPACITY; *d_len must not exceed capacity. Calls that overlap in
 * time on one GPU (other streams or threads) share that array and race;
 * calls ordered on one stream are safe. Results equal apbf_radix_sort's. */
#define APBF_RADIX_SORT_NOALLOC_MAX (1u << 20)
apbf_status apbf_radix_sort_noalloc(const uint32_t* keys, const uint32_t* values, const uint32_t* d_len,
                                    uint32_t capacity, uint32_t upper_bound, uint32_t* out_keys,
                                    uint32_t* out_values, apbf_stream stream);

/* z-order hash of each position in the grid [min, max] with 2^res cells per
 * axis (shader_provider::calculate_position_hash, shader_provider.cpp:805;
 * calculate_position_hash.comp:23-46) over the first `dimensions` (2 or 3,
 * the reference's compile-time DIMENSIONS) axes; res * dimensions <= 30. */
apbf_status apbf_calculate_position_hash(const int32_t* position, const uint32_t* d_len,
                                         uint32_t capacity, const float* host_min3,
                                         const float* host_max3, uint32_t res_log2,
                                         uint32_t dimensions, uint32_t* out_hash, apbf_stream stream);

/* One 32-bit section (0..2) of the 96-bit Morton code of the raw positions
 * (shader_provider::calculate_position_code, shader_provider.cpp:782;
 * calculate_position_code.comp:23-72). */
apbf_status apbf_calculate_position_code(const int32_t* position, const uint32_t* d_len,
                                         uint32_t capacity, uint32_t section, uint32_t* out_code,
                                         apbf_stream stream);

/* Cell start/end tables over sorted keys (clears both tables of n_cells entries,
 * then shader_provider::find_value_ranges, shader_provider.cpp:221;
 * find_value_ranges.comp:16-31). */
apbf_status apbf_find_value_ranges(const uint32_t* sorted_keys, const uint32_t* d_len,
                                   uint32_t capacity, uint32_t n_cells, uint32_t* cell_start,
                                   uint32_t* cell_end, apbf_stream stream);

/* Green hash-grid neighbour search over particles sorted by cell
 * (shader_provider::neighborhood_green, shader_provider.cpp:659;
 * neighborhood_green.comp:49-87). Output is a CSR list: row_offset[i] ..
 * row_offset[i+1] holds the neighbours of particle i in the reference's cell-
 * walk order; pairs beyond pair_capacity are dropped (the reference's clamp,
 * neighbor_add.glsl:22-23) and *d_n_pairs receives the clamped total.
 * With dimensions 2 the cells are those of the x-y grid (the walk stays in
 * z cell 0, neighborhood_green.comp:36); the distance test is 3D as in the
 * reference. `workspace` >= apbf_neighborhood_green_workspace_size(capacity). */
size_t apbf_neighborhood_green_workspace_size(uint32_t capacity);
apbf_status apbf_neighborhood_green(const int32_t* position, const float* range,
                                    const uint32_t* d_len, uint32_t capacity, float range_scale,
                                    const float* host_min3, const float* host_max3,
                                    uint32_t res_log2, uint32_t dimensions, const uint32_t* cell_start,
                                    const uint32_t* cell_end, uint32_t* row_offset,
                                    uint32_t* neighbors, uint32_t pair_capacity,
                                    uint32_t* d_n_pairs, void* workspace, size_t workspace_bytes,
                                    apbf_stream stream);

/* Brute-force neighbour search (shader_provider::neighborhood_brute_force;
 * neighborhood_brute_force.comp:31-50): every particle against every other in
 * index order, no sort. Same CSR output, clamp and workspace as
 * apbf_neighborhood_green; with capacity 0 (an empty list) position and range
 * may be null. */
apbf_status apbf_neighborhood_brute_force(const int32_t* position, const float* range,
                                          const uint32_t* d_len, uint32_t capacity, float range_scale,
                                          uint32_t* row_offset, uint32_t* neighbors, uint32_t pair_capacity,
                                          uint32_t* d_n_pairs, void* workspace, size_t workspace_bytes,
                                          apbf_stream stream);

/* Gather by edit list: dst[i] = src[edit[i]] for i < d_edit_len, element size
 * `stride` bytes (a multiple of 4) (gpu_list::apply_edit, gpu_list.h:126-135;
 * copy_scattered_read.comp:17-27). */
apbf_status apbf_copy_scattered_read(const void* src, void* dst, const uint32_t* edit,
                                     const uint32_t* d_edit_len, uint32_t capacity, uint32_t stride,
                                     apbf_stream stream);

/* Index-list rebuild after a hidden-list edit (indexed_list::apply_hidden_edit,
 * indexed_list.h:289-308; atomic_swap.comp, generate_new_index_and_edit_list.comp).
 * For each new hidden slot h (ascending) and each index-list position p with
 * index[p] == edit[h] (descending p), appends h to new_index and p to new_edit,
 * up to max_new entries; *d_new_len receives the length. Deterministic.
 * `workspace` >= apbf_hidden_edit_workspace_size(index_capacity, hidden_capacity, edit_capacity). */
size_t apbf_hidden_edit_workspace_size(uint32_t index_capacity, uint32_t hidden_capacity,
                                       uint32_t edit_capacity);
apbf_status apbf_apply_hidden_edit(const uint32_t* index_list, const uint32_t* d_index_len,
                                   uint32_t index_capacity, uint32_t hidden_capacity,
                                   const uint32_t* edit, const uint32_t* d_edit_len,
                                   uint32_t edit_capacity, uint32_t max_new, uint32_t* new_index,
                                   uint32_t* new_edit, uint32_t* d_new_len, void* workspace,
                                   size_t workspace_bytes, apbf_stream stream);

/* Edit list of indexed_list::delete_these (indexed_list.h:126-138): the hidden
 * slots [0, d_hidden_len) not named in `deleted`, ascending; *d_edit_len receives
 * the count. `workspace` >= apbf_delete_edit_list_workspace_size(hidden_capacity). */
size_t apbf_delete_edit_list_workspace_size(uint32_t hidden_capacity);
apbf_status apbf_delete_edit_list(const uint32_t* deleted, const uint32_t* d_deleted_len,
                                  uint32_t deleted_capacity, const uint32_t* d_hidden_len,
                                  uint32_t hidden_capacity, uint32_t* edit, uint32_t* d_edit_len,
                                  void* workspace, size_t workspace_bytes, apbf_stream stream);

/* gpu_list::operator+= (gpu_list.h:214-227; shader_provider::append_list,
 * shader_provider.cpp:49-72, append_list.comp:21-25): appends the
 * *d_append_len elements of `appending` (stride bytes each, a multiple of 4)
 * after the *d_target_len elements of `target`; *d_new_len = min(target
 * length + appended length, target_capacity), and elements that would land
 * beyond the capacity are dropped. `append_capacity` bounds the launch (the
 * appended list's allocated length); appending may be NULL when it is 0.
 * d_new_len may equal d_target_len. */
apbf_status apbf_append_list(void* target, uint32_t target_capacity, const uint32_t* d_target_len,
                             const void* appending, const uint32_t* d_append_len, uint32_t stride,
                             uint32_t* d_new_len, uint32_t append_capacity, apbf_stream stream);
/* shader_provider::write_increasing_sequence (shader_provider.cpp:198-218,
 * write_increasing_sequence.comp:23-33), the core of
 * indexed_list::increase_length (indexed_list.h:194-202): writes
 * min + 0, 1, ... into target for len = min(target_capacity, sequence_length,
 * value_upper_bound - min) elements, with min = *d_sequence_min;
 * *d_new_sequence_min = min + len and *d_new_target_len = len (the two length
 * results the reference returns). */
apbf_status apbf_write_increasing_sequence(uint32_t* target, uint32_t target_capacity,
                                           const uint32_t* d_sequence_min, uint32_t value_upper_bound,
                                           uint32_t sequence_length, uint32_t* d_new_sequence_min,
                                           uint32_t* d_new_target_len, apbf_stream stream);
/* shader_provider::write_increasing_sequence_from_to (shader_provider.cpp:178-196,
 * write_increasing_sequence_from_to.comp:19-27), used by
 * indexed_list::duplicate_these (indexed_list.h:141-152): out[i] = from + i
 * for from + i < to (at most out_capacity values), *d_out_len = to - from. */
apbf_status apbf_write_increasing_sequence_from_to(uint32_t* out, uint32_t out_capacity, uint32_t* d_out_len,
                                                   const uint32_t* d_from, const uint32_t* d_to, apbf_stream stream);
/* shader_provider::write_sequence (write_sequence.comp:15-20): out[i] =
 * start + i * step for i < *d_len (fills with step 0, iota with step 1). */
apbf_status apbf_write_sequence(uint32_t* out, const uint32_t* d_len, uint32_t capacity, uint32_t start, uint32_t step,
                                apbf_stream stream);
/* Device memory of the host-side lists (the buffers gpu_list_data draws from
 * its pool, gpu_list_data.cpp:6-45), a stream-ordered copy in any direction
 * (algorithms::copy_bytes, algorithms.cpp:4-36) and a stream
 * synchronisation (the reference's end_recording + wait). */
apbf_status apbf_device_alloc(size_t bytes, void** out);
apbf_status apbf_device_free(void* ptr);
apbf_status apbf_copy(void* dst, const void* src, size_t bytes, apbf_stream stream);
apbf_status apbf_stream_synchronize(apbf_stream stream);

/* Scene settings the reference's subsystems take through setters:
 * velocity_handling::set_acceleration (velocity_handling.h:12),
 * neighborhood_green::set_range_scale (neighborhood_green.h:12; a negative
 * scale restores the default of pool.cpp:83: 1.5 with spreading, else 1.0)
 * and neighborhood_green::set_position_range (neighborhood_green.h:13). Call
 * them between substeps; they drop a captured graph. */
apbf_status apbf_scene_set_gravity(void* scene, const float* gravity3);
/* sphere_collision::set_sphere (spherical_pool.cpp:29): one sphere collider;
 * hollow != 0 keeps the particles inside it, 0 outside. It runs after the
 * boxes at the start of every solver iteration (alone in the reference's
 * spherical_pool, spherical_pool.cpp:86-89) and as APBF_STAGE_SPHERE.
 * enable = 0 removes it. Drops a captured graph. */
apbf_status apbf_scene_set_sphere(void* scene, const float* center3, float radius, int32_t hollow, int32_t enable);
/* The neighbour search of the scene's neighbourhood stage, the compile-time
 * NEIGHBORHOOD_TYPE of the reference (cpu_gpu_shared_config.h:54-60):
 * 0 neighborhood_green (hash grid, default), 1 neighborhood_binary_search
 * (z-order codes of the raw positions, neighborhood_binary_search.cpp:22-75),
 * 2 neighborhood_brute_force (every pair tested, no sort;
 * neighborhood_brute_force.cpp:19-42). Every method writes each particle's
 * neighbours as one row in walk order: the reference's list with
 * mNeighborListSorted set (neighbor_add.glsl:5-18 and
 * linked_list_to_neighbor_list.comp), and one ordering of its unsorted
 * (atomic-order) list. */
apbf_status apbf_scene_set_neighborhood_method(void* scene, int32_t method);
apbf_status apbf_scene_set_range_scale(void* scene, float scale);
apbf_status apbf_scene_set_position_range(void* scene, const float* min3, const float* max3, uint32_t res_log2);

/* -------------------------------------------- stage operators over lists */
/* The subsystem apply() calls of the reference run on the lists they were
 * bound to with set_data (velocity_handling.h:10-12, particle_transfer.h:11-12,
 * neighborhood_green.h:11-14, spread_kernel_width.h:10-11, box_collision.h:10-11,
 * incompressibility.h:11-12, update_transfers.h:11-12). The operators below do
 * the same on caller-owned device arrays: the particle lists of
 * list_definitions.h:9-13 (hidden particles and fluid attributes, stored with
 * fluid id == hidden slot, as the library's scenes keep them), the transfer
 * lists (list_definitions.h:15-16, endpoints are particle slots) and the
 * neighbour list (list_definitions.h:19) as CSR rows: row_offset[i] ..
 * row_offset[i + 1] index the neighbours of particle i for i < *d_rows (the
 * particles present when the list was built), *d_length pairs in all
 * (apbf_neighbor_pairs gives the reference's uvec2 layout). Lengths are
 * device words, read and written by the operators on the stream. */
typedef struct apbf_particle_lists {
    int32_t* position;           /* 4 x int32 per particle, fixed point 2^-18 */
    float* velocity;             /* 4 x f32 */
    float* inverse_mass;
    float* radius;
    int32_t* pos_backup;         /* 4 x int32 */
    uint32_t* transferring;
    float* target_radius;
    float* kernel_width;
    float* boundariness;
    uint32_t* boundary_distance; /* fixed point 2^-18 */
    uint32_t* d_length;          /* particles */
    uint32_t capacity;           /* must equal the context's particle capacity */
} apbf_particle_lists;
typedef struct apbf_transfer_lists {
    uint32_t* source;
    uint32_t* target;
    float* time_left;            /* > 0 merge, <= 0 split (particle_transfer.comp:37-84) */
    uint32_t* d_length;
    uint32_t capacity;           /* must equal the context's transfer capacity */
} apbf_transfer_lists;
typedef struct apbf_neighbor_lists {
    uint32_t* row_offset;        /* particle capacity + 1 entries */
    uint32_t* neighbors;         /* neighbour ids, `capacity` entries */
    uint32_t* d_length;          /* pairs */
    uint32_t* d_rows;            /* particles the rows cover */
    uint32_t capacity;           /* must equal the context's pair capacity */
} apbf_neighbor_lists;
/* The temporaries of every stage for the capacities and settings of
 * `params` (cap_particles, cap_transfers, cap_pairs; the apbf_settings
 * uniform block; the hash cell count 2^(3 hash_res_log2)). Allocates. One
 * context serves one stream at a time. */
typedef struct apbf_context apbf_context;
apbf_status apbf_context_create(const apbf_scene_params* params, apbf_context** out_context);
apbf_status apbf_context_destroy(apbf_context* context);
/* velocity_handling::apply (velocity_handling.cpp:15-31; shader_provider
 * infer_velocity / apply_acceleration / apply_velocity,
 * shader_provider.cpp:1288-1354): *last_delta_time is the member
 * mLastDeltaTime (velocity_handling.h:17), read and updated. */
apbf_status apbf_op_velocity_handling(apbf_context* context, const apbf_particle_lists* particles,
                                      float* last_delta_time, float delta_time, const float* acceleration3,
                                      apbf_stream stream);
/* particle_transfer::apply (particle_transfer.cpp:10-28; shader_provider::
 * particle_transfer, shader_provider.cpp:1247): mass and volume moves, then
 * the finished transfers' particles are deleted (indexed_list::delete_these)
 * and both lists compacted. */
apbf_status apbf_op_particle_transfer(apbf_context* context, const apbf_particle_lists* particles,
                                      const apbf_transfer_lists* transfers, float delta_time, apbf_stream stream);
/* uint_to_float_with_indexed_lower_bound (pool.cpp:77-80,
 * shader_provider.cpp:305-330; settings from the context). */
apbf_status apbf_op_kernel_width_from_boundary_distance(apbf_context* context, const apbf_particle_lists* particles,
                                                        apbf_stream stream);
/* neighborhood_green::apply (neighborhood_green.cpp:27-77), or with method 1
 * neighborhood_binary_search::apply (neighborhood_binary_search.cpp:22-75),
 * method 2 neighborhood_brute_force::apply: the particles are reordered by
 * cell (the hidden-list edit of neighborhood_green.cpp:50), the transfer
 * endpoints follow them (indexed_list owner rebuild), and the neighbour list
 * is rebuilt with range kernel_width * range_scale. res_log2 must give the
 * context's cell count. transfers may be NULL (no transfers). */
apbf_status apbf_op_neighborhood(apbf_context* context, const apbf_particle_lists* particles,
                                 const apbf_transfer_lists* transfers, const apbf_neighbor_lists* neighbors,
                                 int32_t method, float range_scale, const float* min3, const float* max3,
                                 uint32_t res_log2, apbf_stream stream);
/* spread_kernel_width::apply (spread_kernel_width.cpp:12-26; kernel_width_init,
 * kernel_width, uint_to_float_but_gradual, shader_provider.cpp:826-875,284-304):
 * the neighbour list is replaced by its kept pairs. */
apbf_status apbf_op_spread_kernel_width(apbf_context* context, const apbf_particle_lists* particles,
                                        const apbf_neighbor_lists* neighbors, apbf_stream stream);
/* box_collision::apply (box_collision.cpp:12-24; shader_provider::box_collision,
 * shader_provider.cpp:559) with boxes [box_min4[4i..], box_max4[4i..]]. */
apbf_status apbf_op_box_collision(apbf_context* context, const apbf_particle_lists* particles, uint32_t n_boxes,
                                  const float* host_box_min4, const float* host_box_max4, apbf_stream stream);
/* sphere_collision::apply (sphere_collision.comp:42-60). */
apbf_status apbf_op_sphere_collision(apbf_context* context, const apbf_particle_lists* particles,
                                     const float* host_center3, float radius, int32_t hollow, apbf_stream stream);
/* incompressibility::apply (incompressibility.cpp:12-45): one Jacobi
 * iteration, incompressibility_0..3 (shader_provider.cpp:970-1093). */
apbf_status apbf_op_incompressibility(apbf_context* context, const apbf_particle_lists* particles,
                                      const apbf_neighbor_lists* neighbors, apbf_stream stream);
/* update_transfers::apply (update_transfers.cpp:14-71): find_split_and_merge_1..3
 * (shader_provider.cpp:1094-1192), remove_impossible_splits and
 * initialize_split_particles (:1193-1246), duplicate_these, appends. */
apbf_status apbf_op_update_transfers(apbf_context* context, const apbf_particle_lists* particles,
                                     const apbf_transfer_lists* transfers, const apbf_neighbor_lists* neighbors,
                                     apbf_stream stream);
/* The neighbour list in the reference's layout: out_pairs[2k], out_pairs[2k+1]
 * = (particle, neighbour) of pair k, in row order (list_definitions.h:19,
 * neighbor_add.glsl:20-23). rows_capacity bounds the launch. */
apbf_status apbf_neighbor_pairs(const uint32_t* row_offset, const uint32_t* neighbors, const uint32_t* d_rows,
                                const uint32_t* d_n_pairs, uint32_t rows_capacity, uint32_t pair_capacity,
                                uint32_t* out_pairs, apbf_stream stream);
/* The reverse: CSR rows over `rows` particles from *d_n_pairs uvec2 pairs in
 * any order (an atomic-order list of neighborhood_green.comp); the pairs of
 * one particle keep their relative order. First ids must be < rows.
 * row_offset holds rows + 1 entries. */
size_t apbf_neighbor_rows_from_pairs_workspace_size(uint32_t pair_capacity, uint32_t rows);
apbf_status apbf_neighbor_rows_from_pairs(const uint32_t* pairs, const uint32_t* d_n_pairs, uint32_t pair_capacity,
                                          uint32_t rows, uint32_t* row_offset, uint32_t* neighbors, void* workspace,
                                          size_t workspace_bytes, apbf_stream stream);

/* ---------------------------------------------------------- packed batches */
/* Many independent scenes with the same settings (apbf_settings, basicPbf,
 * solverIterations, splitDuration, gravity) stepped as one particle list:
 * every stage is one launch over all of them. No counterpart in the
 * reference, which runs one scene per process (main.cpp:62-68). Each
 * particle carries its scene's index; the hash key holds that index above
 * the cell hash of the scene's own grid, so one stable sort orders the list
 * by scene, then by cell, and the neighbour walk stays inside the scene's
 * block of cells; the box collision takes the scene's boxes and the
 * particle's list position inside its scene. Every scene's particles are
 * ordered, searched, solved, split and merged as if the scene ran alone
 * (bit for bit as long as no capacity clamps; the capacities of `params` are
 * shared by all scenes). Neighbour search: the hash grid only; no sphere. */
typedef struct apbf_segment_desc {
    float hash_min[3];
    float hash_max[3];
    uint32_t hash_res_log2;
    uint32_t n_boxes;
    float box_min[APBF_MAX_BOXES][4];
    float box_max[APBF_MAX_BOXES][4];
} apbf_segment_desc;
/* params' hash domain and boxes are ignored (they come from the segments). */
apbf_status apbf_scene_create_packed(const apbf_scene_params* params, uint32_t n_segments,
                                     const apbf_segment_desc* segments, void** out_scene);
/* Scene index of every particle, in the order of the state given to
 * apbf_scene_set_state (call after it); reading gives the current order. */
apbf_status apbf_scene_set_segment_ids(void* scene, const uint32_t* host_ids);
apbf_status apbf_scene_get_segment_ids(void* scene, uint32_t* host_ids);

/* ------------------------------------------ graphs, batches, launch shapes */
/* No counterpart in the reference, which records one command buffer with a
 * barrier after every dispatch (shader_provider.cpp:20-47). */
/* Captures two substeps of dt into a HIP graph (the scene's buffer roles
 * repeat every two substeps); apbf_scene_replay launches it `times` times,
 * advancing the scene by 2 * times substeps with results identical to
 * apbf_scene_update. The scene gets a stream of its own if it had none.
 * Changing the launch shape drops the graph. */
apbf_status apbf_scene_capture(void* scene, float dt);
apbf_status apbf_scene_replay(void* scene, uint32_t times);
/* Many independent scenes stepped together, each on a stream of its own:
 * apbf_batch_update enqueues one substep of every scene; apbf_batch_capture
 * captures two substeps of every scene (apbf_scene_capture) and
 * apbf_batch_replay launches all of those graphs `times` times. */
typedef struct apbf_batch apbf_batch;
apbf_status apbf_batch_create(void* const* scenes, uint32_t n_scenes, apbf_batch** out_batch);
apbf_status apbf_batch_destroy(apbf_batch* batch);
apbf_status apbf_batch_update(apbf_batch* batch, float dt);
apbf_status apbf_batch_capture(apbf_batch* batch, float dt);
apbf_status apbf_batch_replay(apbf_batch* batch, uint32_t times);
apbf_status apbf_batch_synchronize(apbf_batch* batch);
/* Launch shapes: lanes per particle in the neighbour walk (4, 8, 16, or 64:
 * a whole wavefront per particle over a flattened stream of the box's cells)
 * and per row in the solver (8, 16, 32); every shape computes identical
 * results. Until a shape is set or tuned, the walk takes 4 lanes when a
 * recent substep's mean neighbour row is below 62 entries and whole
 * wavefronts otherwise, and the solver 8 lanes for such short rows and 16
 * otherwise; 0 for either returns it to that choice. The recent
 * substep's counts reach the host by a copy that is never waited for; a
 * graph captured by apbf_scene_capture keeps the choices of the state it
 * was captured from.
 * apbf_scene_autotune times `substeps` substeps under each shape, every
 * shape from the same saved state (twice, keeping the faster time), and keeps
 * the fastest; it then times the scene's hash sort alone under every sort
 * shape (below) on the keys of one substep from the same state and keeps the
 * fastest. The scene and its statistics are put back afterwards, except the
 * neighbour list, which is empty until the next substep. */
apbf_status apbf_scene_set_launch_shape(void* scene, int32_t walk_lanes, int32_t solve_lanes);
/* The radix sort of the scene's neighbourhood builds: digit width 4 (the
 * reference's, algorithms.cpp:61-64), 8 (the default) or 11 bits, and 4, 8 or
 * 16 values per thread of a 256-thread tile. Every shape gives the same
 * stable order. */
apbf_status apbf_scene_set_sort_shape(void* scene, int32_t digit_bits, int32_t items_per_thread);
apbf_status apbf_scene_sort_shape(void* scene, int32_t* digit_bits, int32_t* items_per_thread);
/* The solver's row groups per shift-table flush in LDS: 2 for long rows, 4
 * for short ones. mode -1 (the default) picks by the mean row length of a
 * recent substep, read without synchronisation; 0 and 1 pin the short- and
 * long-row variants. No counterpart in the reference; both variants compute
 * identical results. */
apbf_status apbf_scene_set_long_rows(void* scene, int32_t mode);
apbf_status apbf_scene_autotune(void* scene, float dt, uint32_t substeps, int32_t* best_walk_lanes,
                                int32_t* best_solve_lanes, float* best_ms_per_substep);

/* ------------------------------------------------- learned lambda warm start */
/* No counterpart in the reference (it recomputes lambda from zero every
 * iteration and discards it, incompressibility.cpp:22-25): a 16-64-64-1 ReLU
 * network, BF16 on the matrix cores with FP32 accumulation, that predicts
 * from 16 per-particle features the total lambda the plain Jacobi solve
 * applies in a substep. Parameters are a flat FP32 vector of
 * APBF_MLP_PARAMS values (W1[64][16], b1[64], W2[64][64], b2[64], w3[64],
 * b3) plus APBF_MLP_NORM normalisation values (feature means[16], feature
 * scales[16], output scale). */
#define APBF_MLP_FEATURES 16
#define APBF_MLP_PARAMS 5313
#define APBF_MLP_NORM 33
typedef struct apbf_mlp apbf_mlp;
apbf_status apbf_mlp_create(apbf_mlp** out_model);
apbf_status apbf_mlp_destroy(apbf_mlp* model);
apbf_status apbf_mlp_set_params(apbf_mlp* model, const float* host_params, const float* host_norm);
apbf_status apbf_mlp_get_params(apbf_mlp* model, float* host_params, float* host_norm);
/* out[s] = prediction for features[s * 16 .. +16), s < *d_len (device arrays). */
apbf_status apbf_mlp_forward(apbf_mlp* model, const float* features, const uint32_t* d_len, uint32_t capacity,
                             float* out, apbf_stream stream);
/* `steps` Adam steps (lr, betas 0.9/0.999) on 0.5 * mean((y - target/out_scale)^2)
 * over batches perm[(k * batch) % n ..] (device arrays, n samples); the loss of
 * each step goes to host_losses[steps] when not NULL (synchronises). */
apbf_status apbf_mlp_train(apbf_mlp* model, const float* features, const float* targets, const uint32_t* perm,
                           uint32_t n, uint32_t batch, uint32_t steps, float lr, float* host_losses,
                           apbf_stream stream);
/* Gradient of the same loss over all n samples, without an update. */
apbf_status apbf_mlp_gradient(apbf_mlp* model, const float* features, const float* targets, const uint32_t* perm,
                              uint32_t n, float* host_grad, float* host_loss, apbf_stream stream);

/* Scene hooks. mode 0: plain solver. mode 1 (harvest): before the solver
 * loop the scene writes the features of every particle, and after it the
 * total lambda applied (sum over iterations of min(lambda, 0)); read them
 * with apbf_scene_harvest. mode 2 (warm start): the features go through
 * `model`, and one shift pass with the predicted lambdas runs before the
 * solver's iterations. The model must outlive its use by the scene. */
apbf_status apbf_scene_set_warm_start(void* scene, apbf_mlp* model, int32_t mode);
/* The lambdas the network predicted in the last warm-started substep, one
 * per particle of that substep's solver (the particles its neighbour rows
 * cover, apbf_scene_get_counts' count before split particles were added).
 * Synchronises. */
apbf_status apbf_scene_get_warm_lambdas(void* scene, float* host_lambda);
/* Copies the features (n x 16) and targets (n) of the last harvested substep
 * to device buffers of `capacity` rows; *count receives n. Synchronises. */
apbf_status apbf_scene_harvest(void* scene, float* features, float* targets, uint32_t capacity, uint32_t* count);

/* ------------------------------------------------------ state snapshots */
/* The keyframes of the reference's time machine (time_machine.h:46-143): a
 * device-resident copy of a scene's particle, fluid and transfer lists, count
 * block and last time step (not the neighbour list, which every substep
 * rebuilds; a restored scene reports no pairs until its next neighbourhood
 * stage). create allocates and saves; save overwrites; restore copies back
 * into the scene. Copies are stream-ordered on the scene's stream (no host
 * synchronisation). A snapshot fits only the scene (capacities) it came from. */
typedef struct apbf_snapshot apbf_snapshot;
apbf_status apbf_snapshot_create(void* scene, apbf_snapshot** out_snapshot);
apbf_status apbf_snapshot_save(apbf_snapshot* snapshot, void* scene);
apbf_status apbf_snapshot_restore(apbf_snapshot* snapshot, void* scene);
apbf_status apbf_snapshot_destroy(apbf_snapshot* snapshot);

/* --------------------------------------------------------- particle dumps */
/* save_particle_info::apply (save_particle_info.cpp:21-128): writes
 * centerDist.txt, radius.txt, neighborCount.txt, kernelWidth.txt,
 * targetRadius.txt, boundaryDistance.txt and data.csv of the scene's fluid into
 * `folder` (created if missing; the reference uses "particle_data"), in the
 * reference's formats. Synchronises. */
apbf_status apbf_scene_save_particle_info(void* scene, const char* folder);
/* save_particle_info::save_additional_info (save_particle_info.cpp:271-280):
 * folder/measurements_<id>.txt with the particle and pair counts, the given
 * timings and the scene's neighbour search (the reference's 3D SVG dump writes
 * only this file). Synchronises. */
apbf_status apbf_scene_save_measurements(void* scene, const char* folder, uint32_t id, float step_ms,
                                         float neighborhood_ms, float solver_ms);
/* save_particle_info::save_as_svg (save_particle_info.cpp:130-254, boxes_to_svg
 * :256-270), the drawing: folder/particles_<svg_id>.svg of a planar scene
 * (dimensions 2) in the view box [view_min2, view_max2], with the particle
 * circles scaled by render_scale (settings::particleRenderScale), coloured by
 * `color` (settings::color: 0 boundariness, 1 boundary-distance gradient up to
 * 0.8 max_expected_boundary_distance, 2 transfer pairs, 3 boundariness plus
 * kernel-width circles) and the scene's boxes when render_boxes
 * (settings::renderBoxes). Numbers as std::format "{}" prints them. The
 * reference's save_as_svg also writes measurements_<id>.txt first
 * (apbf_scene_save_measurements); in 3D it draws nothing, and neither does
 * this call. Synchronises. */
apbf_status apbf_scene_save_svg(void* scene, const char* folder, uint32_t svg_id, const float* view_min2,
                                const float* view_max2, float render_scale, float max_expected_boundary_distance,
                                int32_t color, int32_t render_boxes);

/* ------------------------------------------------------------ utilities */
/* Library version string and the device it was built for ("gfx950"). */
const char* apbf_version(void);
/* Number of HIP devices visible (0 without a GPU); never fails. */
int32_t apbf_device_count(void);
/* Self-check of the device float->int conversions (truncation toward zero,
 * saturating, NaN -> 0; the GLSL conversions the reference relies on, e.g.
 * incompressibility_1.comp:52-58): runs all 2^32 float bit patterns through
 * the hardware conversion the kernels use and through its definition.
 * *mismatches = patterns where they differ (*first_bad the smallest);
 * sample_i32 / sample_u32 (host, ceil(2^32 / stride) entries each) receive the
 * device results of every stride-th pattern. stride >= 1024. No counterpart
 * in the reference. */
apbf_status apbf_check_conversions(uint32_t stride, uint64_t* mismatches, uint32_t* first_bad,
                                   int32_t* sample_i32, uint32_t* sample_u32);
/* Self-check of the device square root and reciprocal the Gaussian gradient
 * uses (IEEE sqrt and 1/x, kernels.glsl:91-97): every float bit pattern of
 * their argument ranges through the short device forms and through the full
 * IEEE forms; and of the solver's quotients by inverse masses (IEEE x / b,
 * incompressibility_1.comp:53-59) through a reciprocal and one correction:
 * 512 divisors, each with every dividend significand. mismatches3[f] /
 * first_bad3[f] for f = 0 quotient, 1 square root, 2 reciprocal. No
 * counterpart in the reference. */
apbf_status apbf_check_math(uint64_t* mismatches3, uint32_t* first_bad3);

#ifdef __cplusplus
}
#endif
#endif
