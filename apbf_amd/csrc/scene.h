// Scene: device buffers and subsystem methods of one simulation.
#pragma once
#include <vector>

#include "mlp.h"
#include "plain_kernels.h"

namespace apbf {

// Kernels that can be timed with HIP events (apbf_scene_time_kernel).
enum KernelId : int {
    K_SOLVE_ROWS = 0,      // k_solve_rows: density, lambda, boundariness, shifts (one iteration)
    K_SOLVE_APPLY = 1,      // k_solve_apply: apply the shifts + next box collision
    K_WALK_COUNT = 2,  // k_walk_rows<count>: neighbour counting walk (with spreading when fused)
    K_WALK_FILL = 3,   // k_walk_rows<fill>: neighbour writing walk (with spreading when fused)
    K_SPREAD = 4,       // k_spread_rows<count>: kernel-width spreading (counting pass)
    K_FSM12 = 5,        // k_fsm12_rows: boundary-distance flood fill + nearest neighbour
    K_PERMUTE = 6,      // k_permute: particle SoA gather by the sort permutation
    K_COUNT_IDS = 7
};

constexpr int MAX_ITER_RECORD = 64;  // solver iterations with a recorded density error

struct Scene {
    apbf_scene_params P{};
    hipStream_t stream = nullptr;
    int device = 0;
    uint32_t cap = 0, tcap = 0, pcap = 0, n_cells = 0;
    GridDesc grid{};
    BoxSet boxes{}, no_boxes{};
    // A packed batch of independent scenes (segments): per-particle scene
    // index, per-scene grids and boxes, per-scene block of 2^seg_shift cells.
    uint32_t n_seg = 1, seg_shift = 0;
    GridDesc* d_grids = nullptr;
    BoxSet* d_boxes = nullptr;
    uint32_t* seg_start = nullptr;
    uint32_t* sid[2] = {};
    bool packed() const { return n_seg > 1 || sid[0] != nullptr; }
    // DIMENSIONS (cpu_gpu_shared_config.h:2): 2 or 3 (init maps 0 to 3)
    uint32_t dims() const { return P.dimensions; }
    float dim() const { return (float)P.dimensions; }
    SegmentTable seg_table() const {
        return packed() ? SegmentTable{d_grids, d_boxes, seg_start, sid[cur], seg_shift}
                        : SegmentTable{nullptr, nullptr, nullptr, nullptr, 0u};
    }
    apbf_status init_segments(uint32_t n, const apbf_segment_desc* segs);
    float last_dt = 1.0f;  // velocity_handling.h:17
    float range_scale_override = -1.0f;  // neighborhood_green::set_range_scale; < 0: the mode's default
    int cur = 0;           // current particle buffer set
    int tcur = 0;          // current transfer buffer set
    bool pairs_valid = false;    // CSR rows refer to the current particle slots
    bool measure_error = false;  // accumulate the density error after every solver loop
    double* err_stats = nullptr; // [sum of means, max, last mean, last max, sum of abs means, last abs mean]
    // sums over substeps: [0] the solver's particle count, [1] its neighbour
    // pairs, [2] the neighbour walk's pairs before the spreading filter,
    // [3] particles created by splits, [4] particles deleted by finished merges
    unsigned long long* particle_substeps = nullptr;
    double* iter_err = nullptr;  // per solver iteration: sum over substeps of the mean error it started from
    // learned lambda warm start (mode 0 off, 1 harvest, 2 warm start)
    int ws_mode = 0;
    apbf_mlp* ws_model = nullptr;
    float* feat = nullptr;     // cap x 16 features of the current substep
    float* lam_acc = nullptr;  // per particle: sum of min(lambda, 0) over this substep's iterations
    float* lam_hat = nullptr;  // predicted lambdas
    // tolerance mode and the warm-start fallback (solver_loop)
    float tol = 0.0f;          // > 0: iterate until the mean density error is <= tol
    int tol_max_iters = 0;     // iteration limit in tolerance mode
    double* tol_err = nullptr; // per substep: error sums of the warm pass (slot 0) and iteration k (slot 1 + k)
    int4* pos_save = nullptr;  // positions before the warm-start shift
    unsigned long long* solve_stats = nullptr;  // [iterations applied, substeps, warm-start fallbacks] summed

    // Optional HIP-event timing of kernels (bit KernelId of timed_mask) on the
    // scene stream; event pair k belongs to kernel event_ids[k].
    uint32_t timed_mask = 0;
    std::vector<hipEvent_t> events;
    std::vector<int> event_ids;
    size_t events_used = 0;
    bool tbegin(int id) {
        if (!(timed_mask & (1u << id)) || events_used + 2 > events.size()) return false;
        event_ids[events_used / 2] = id;
        (void)hipEventRecord(events[events_used++], stream);
        return true;
    }
    void tend(bool on) {
        if (on) (void)hipEventRecord(events[events_used++], stream);
    }

    // particle SoA, two buffer sets (gathers are out of place)
    int4* pos[2] = {};
    float4* vel[2] = {};
    float* inv_mass[2] = {};
    float* radius[2] = {};
    int4* backup[2] = {};
    uint32_t* transferring[2] = {};
    float* target_radius[2] = {};
    float* kernel_width[2] = {};
    float* boundariness[2] = {};
    uint32_t* boundary_distance[2] = {};
    float* lam_prev[2] = {};  // total lambda applied to the particle in the previous substep
    // hidden transfer list (list_definitions.h:15-16), two buffer sets
    uint32_t* t_source[2] = {};
    uint32_t* t_target[2] = {};
    float* t_time_left[2] = {};

    uint32_t* counts = nullptr;  // CountSlot block
    // Pinned host copy of the count block, refreshed without synchronisation
    // after every substep (so it lags by a substep or two). It only steers
    // the choice between launch variants that compute the same results.
    uint32_t* h_counts = nullptr;
    hipEvent_t h_counts_ready = nullptr;  // recorded after each copy into h_counts
    uint32_t seen_n = 0, seen_pairs = 0;   // particles and pairs of the last completed copy
    void refresh_seen_counts();
    bool long_rows = false;  // mean neighbour row of a recent substep above 62 entries
    int long_rows_mode = -1;  // -1: long_rows from h_counts; 0 / 1: pinned
    // neighbourhood
    uint32_t *keys = nullptr, *vals = nullptr, *skeys = nullptr, *perm = nullptr;
    uint32_t *cell_start = nullptr, *cell_end = nullptr;
    uint32_t *cnt = nullptr, *row = nullptr, *col = nullptr;
    uint32_t* col_tmp = nullptr;
    // solver
    float4* gc = nullptr;  // per particle: Gaussian constants, rest-density factor, own density term
    uint32_t* long_list = nullptr;   // rows the solver takes with 32 lanes per row (counts[C_LONG])
    uint32_t* long_list2 = nullptr;  // rows the solver takes with a wavefront per row (counts[C_LONG2])
    float* lambda = nullptr;
    int4* shift = nullptr;  // per-particle shift accumulator of the solver (x, y, z, unused)
    // adaptive scratch
    uint32_t* ukw = nullptr;
    uint32_t *u0 = nullptr, *u1 = nullptr, *u2 = nullptr, *u3 = nullptr, *u4 = nullptr;
    uint32_t* small = nullptr;
    // split/merge claim resolution (per particle) and transfer compaction scratch
    uint32_t *a_type = nullptr, *a_s = nullptr, *a_t = nullptr, *a_state = nullptr, *a_owner = nullptr;
    uint32_t* a_owner2 = nullptr;  // second owner buffer of the claim rounds
    uint32_t *a_flag = nullptr, *a_scan = nullptr, *a_list = nullptr, *a_t0 = nullptr;
    uint32_t *t_flag = nullptr, *t_scan = nullptr;
    int claim_rounds = 6;
    // launch shapes picked by the autotuner (all compute identical results)
    int walk_lanes = 8;    // lanes per particle in the neighbour walk: 4, 8, 16 or 64 (a whole wave)
    bool walk_auto = true;  // walk lanes from a recent mean row length (4 below 62 entries, else 64)
    int solve_lanes = 16;  // lanes per row in the solver: 8, 16 or 32
    bool solve_auto = true;  // solver lanes from a recent mean row length (8 below 62 entries, else 16)
    SortShape sort_shape = SORT_DEFAULT;  // digit width and tile of the hash sorts
    bool exact_walks = false;  // pair build by the two exact walks only (no fixed-stride slots)
    bool count_candidates = false;  // diagnostic: a counting walk adds its tested candidates to C_CANDIDATES
    // captured substeps (two per graph: the buffer roles repeat after two)
    hipStream_t own_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    float graph_dt = 0.0f;
    // density error partials
    uint32_t err_blocks = 0;
    double* part_sum = nullptr;  // 2 x err_blocks: compression-only sums, then absolute sums
    float* part_max = nullptr;
    void* ws_mem = nullptr;
    size_t ws_bytes = 0;

    Scene() = default;
    ~Scene();
    apbf_status init(const apbf_scene_params& p);
    ParticleArrays arrays(int set);

    apbf_status set_state(uint32_t n, const int32_t* h_pos, const float* h_vel, const float* h_im,
                          const float* h_r, const int32_t* h_backup, const uint32_t* h_tr,
                          const uint32_t* h_fidx, const float* h_trad, const float* h_kw,
                          const float* h_b, const uint32_t* h_bd, uint32_t nt, const uint32_t* h_ts,
                          const uint32_t* h_tt, const float* h_ttl, float last_delta_time);
    apbf_status get_state(int32_t* h_pos, float* h_vel, float* h_im, float* h_r, int32_t* h_backup,
                          uint32_t* h_tr, float* h_trad, float* h_kw, float* h_b, uint32_t* h_bd,
                          uint32_t* h_ts, uint32_t* h_tt, float* h_ttl, uint32_t* h_pairs,
                          float* h_ldt);

    // subsystems (reference L4 apply() calls)
    void velocity_handling(float dt);
    void particle_transfer(float dt);
    void kernel_width_from_bd();
    void neighborhood_green(float range_scale, bool spread);
    void neighborhood_bs(float range_scale);  // neighborhood_binary_search (bsearch.hip)
    void neighborhood_bf(float range_scale);  // neighborhood_brute_force
    // 0: neighborhood_green, 1: neighborhood_binary_search, 2: neighborhood_brute_force
    int nb_method = 0;
    void spread_kernel_width();
    void box_collision();
    void incompressibility();
    void update_transfers(bool with_density_error = false);
    void update(float dt);
    apbf_status stage(int st, float dt);
    apbf_status density_error(double* mean_err, double* max_err);

    // helpers
    void build_pairs(float range_scale, bool spread);
    void spread_begin();
    void spread_reset_if_fallback();
    void spread_end();
    void prepare_solver();
    SolverArgs solver_args() const;
    void incompressibility_iteration(const BoxSet& next_boxes, int it, const float* lam_override = nullptr,
                                     bool decided = false);
    void sphere_collision();
    void colliders();
    void solver_loop();
    void remap_transfers_after_permute();
    void accumulate_density_error();
};

apbf_status scene_switch_stream(Scene* s, hipStream_t next);  // graph.hip: stream-ordered switch
void scene_drop_graph(Scene* s);     // graph.hip: drops a captured graph
void scene_release_graph(Scene* s);  // graph.hip: graph and the scene's own stream

}  // namespace apbf
