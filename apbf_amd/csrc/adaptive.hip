// Adaptive part of a substep: kernel width from boundary distance, kernel-
// width spreading, split/merge detection with claim resolution, split start,
// particle transfer and compaction, plus state upload/download.
//
// The reference resolves merge/split claims with atomicExchange races
// (find_split_and_merge_3.comp:91-116) and appends in atomic order. Here the
// outcome is the one of the sequential schedule (invocations in ascending id
// order), computed in parallel: a proposal wins iff none of its particles is
// already transferring and no lower-id winning proposal touches them. Rounds
// of (owner = atomicMin over undecided-or-winning proposals per particle;
// decide) settle at least the lowest undecided proposal each round; after a
// fixed number of rounds a one-thread pass settles any rest in id order, and
// the same pass replays the whole sequence when merges overflow the transfer
// capacity (where a failed claim releases particles, as in :99-104).
#include "scene.h"

#include <algorithm>
#include <vector>

namespace apbf {


enum : uint32_t { CL_NONE = 0, CL_UNDECIDED = 1, CL_WIN = 2, CL_LOSE = 3 };
enum : uint32_t { P_NONE = 0, P_MERGE = 1, P_SPLIT = 2 };

// uint_to_float_with_indexed_lower_bound.comp:28-44 (pool.cpp:77-80)
__global__ __launch_bounds__(BLOCK) void k_kw_from_bd(const uint32_t* __restrict__ bd,
                                                      float* __restrict__ kw,
                                                      const float* __restrict__ radius,
                                                      const uint32_t* __restrict__ counts, float factor,
                                                      float step) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) {
        float value = (float)bd[i] * factor;
        float lower = radius[i] * KERNEL_SCALE;
        value = move_towards_mult(kw[i], value, step);
        kw[i] = fmax_g(value, lower);
    }
}

// particle_transfer.comp:26-84 per transfer; marks finished merges' sources
// (deleted particles) and finished splits (deleted transfers).
__global__ __launch_bounds__(BLOCK) void k_particle_transfer(const uint32_t* __restrict__ src,
                                                             const uint32_t* __restrict__ tgt,
                                                             float* __restrict__ ttl_arr,
                                                             const uint32_t* __restrict__ counts,
                                                             float* __restrict__ radius,
                                                             float* __restrict__ inv_mass,
                                                             uint32_t* __restrict__ transferring,
                                                             uint32_t* __restrict__ del_p,
                                                             uint32_t* __restrict__ del_t, float dt, float dim) {
    const uint32_t nt = counts[C_TRANSFERS];
    GRID_STRIDE(t, nt) {
        float tt = ttl_arr[t];
        float ttl = fmax_g(dt, __builtin_fabsf(tt));
        bool merge = tt > 0.0f;
        uint32_t s = src[t], g = tgt[t];
        float rS = radius[s], rT = radius[g];
        float imS = inv_mass[s], imT = inv_mass[g];
        float volS = pow_f(rS, dim), volT = pow_f(rT, dim);
        float f = merge ? fmin_g(1.0f, dt / ttl) : dt * (volS - volT) / (2.0f * ttl * volS);
        float tv = f * volS;
        float nf = 1.0f / (imS + f * imT);
        inv_mass[s] = imS / (1.0f - f);
        inv_mass[g] = imS * imT * nf;
        radius[s] = pow_f(volS - tv, 1.0f / dim);
        radius[g] = pow_f(volT + tv, 1.0f / dim);
        ttl_arr[t] = (ttl - dt) * (merge ? 1.0f : -1.0f);
        uint32_t dt_flag = 0;
        if (ttl == dt) {
            if (merge) del_p[s] = 1u;
            else dt_flag = 1u;
            transferring[s] = 0u;
            transferring[g] = 0u;
        }
        del_t[t] = dt_flag;
    }
}

__global__ __launch_bounds__(BLOCK) void k_invert_flags(const uint32_t* __restrict__ in,
                                                        uint32_t* __restrict__ out,
                                                        const uint32_t* __restrict__ counts) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) out[i] = in[i] ? 0u : 1u;
}

// Transfers survive if not a finished split and both particles survive.
__global__ __launch_bounds__(BLOCK) void k_transfer_keep(const uint32_t* __restrict__ src,
                                                         const uint32_t* __restrict__ tgt,
                                                         const uint32_t* __restrict__ del_t,
                                                         const uint32_t* __restrict__ del_p,
                                                         const uint32_t* __restrict__ counts,
                                                         uint32_t* __restrict__ keep) {
    const uint32_t nt = counts[C_TRANSFERS];
    GRID_STRIDE(t, nt) keep[t] = (!del_t[t] && !del_p[src[t]] && !del_p[tgt[t]]) ? 1u : 0u;
}

// Stable compaction of the particle SoA (indexed_list::delete_these,
// indexed_list.h:126-138): kept particle i moves to slot newid[i].
__global__ __launch_bounds__(BLOCK) void k_compact_particles(const uint32_t* __restrict__ del_p,
                                                             const uint32_t* __restrict__ newid,
                                                             const uint32_t* __restrict__ counts,
                                                             ParticleArrays s, ParticleArrays d) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) {
        if (del_p[i]) continue;
        uint32_t o = newid[i];
        d.pos[o] = s.pos[i];
        d.vel[o] = s.vel[i];
        d.inv_mass[o] = s.inv_mass[i];
        d.radius[o] = s.radius[i];
        d.backup[o] = s.backup[i];
        d.transferring[o] = s.transferring[i];
        d.target_radius[o] = s.target_radius[i];
        d.kernel_width[o] = s.kernel_width[i];
        d.boundariness[o] = s.boundariness[i];
        d.boundary_distance[o] = s.boundary_distance[i];
        if (s.lam_prev) d.lam_prev[o] = s.lam_prev[i];
        if (s.sid) d.sid[o] = s.sid[i];
    }
}

__global__ __launch_bounds__(BLOCK) void k_compact_transfers(const uint32_t* __restrict__ keep,
                                                             const uint32_t* __restrict__ pos_t,
                                                             const uint32_t* __restrict__ newid,
                                                             const uint32_t* __restrict__ counts,
                                                             const uint32_t* __restrict__ src,
                                                             const uint32_t* __restrict__ tgt,
                                                             const float* __restrict__ ttl,
                                                             uint32_t* __restrict__ dsrc,
                                                             uint32_t* __restrict__ dtgt,
                                                             float* __restrict__ dttl) {
    const uint32_t nt = counts[C_TRANSFERS];
    GRID_STRIDE(t, nt) {
        if (!keep[t]) continue;
        uint32_t o = pos_t[t];
        dsrc[o] = newid[src[t]];
        dtgt[o] = newid[tgt[t]];
        dttl[o] = ttl[t];
    }
}

__global__ void k_commit_compaction(uint32_t* __restrict__ counts, unsigned long long* __restrict__ totals) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    totals[4] += (unsigned long long)(counts[C_N] - counts[C_KEEP]);  // particles deleted
    counts[C_N] = counts[C_KEEP];
    counts[C_TRANSFERS] = counts[C_TKEEP];
}

// Transfer endpoints follow a particle permutation: inv[perm[k]] = k.
__global__ __launch_bounds__(BLOCK) void k_invert_perm(const uint32_t* __restrict__ perm,
                                                       const uint32_t* __restrict__ counts,
                                                       uint32_t* __restrict__ inv) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(k, n) inv[perm[k]] = k;
}
__global__ __launch_bounds__(BLOCK) void k_remap_transfers(uint32_t* __restrict__ src,
                                                           uint32_t* __restrict__ tgt,
                                                           const uint32_t* __restrict__ inv,
                                                           const uint32_t* __restrict__ counts) {
    const uint32_t nt = counts[C_TRANSFERS];
    GRID_STRIDE(t, nt) {
        src[t] = inv[src[t]];
        tgt[t] = inv[tgt[t]];
    }
}

// kernel_width_init.comp:17-31 (when `gate` is set: only if *gate != 0)
__global__ __launch_bounds__(BLOCK) void k_kw_init(const float* __restrict__ radius,
                                                   const float* __restrict__ trad,
                                                   const uint32_t* __restrict__ counts, int on_target,
                                                   uint32_t* __restrict__ ukw, const uint32_t* __restrict__ gate) {
    if (gate && *gate == 0u) return;
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) {
        float tr = on_target ? trad[i] : 0.0f;
        ukw[i] = to_u32(fmax_g(radius[i], tr) * KERNEL_SCALE * KW_RES);
    }
}

// uint_to_float_but_gradual.comp:27-36 with the default lower bound -inf.
__global__ __launch_bounds__(BLOCK) void k_kw_gradual(const uint32_t* __restrict__ ukw,
                                                      float* __restrict__ kw,
                                                      const uint32_t* __restrict__ counts, float step) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) {
        float value = (float)ukw[i] * (1.0f / KW_RES);
        value = move_towards_mult(kw[i], value, step);
        kw[i] = fmax_g(value, -__builtin_inff());
    }
}

__global__ void k_set_pairs_from_row(uint32_t* __restrict__ counts, const uint32_t* __restrict__ row) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counts[C_PAIRS] = row[counts[C_N]];
}

struct Fsm3Args {
    int update_tr, base_bd, split, merge;
    float smallest, offset, scale, merge_duration, split_factor, dim;
};

// find_split_and_merge_3.comp:44-90: target radius, boundary-distance decay,
// boundariness snap and the split/merge predicates. Emits a proposal; claims
// are resolved separately.
__global__ __launch_bounds__(BLOCK) void k_fsm3(const int4* __restrict__ pos,
                                                const float* __restrict__ radius,
                                                float* __restrict__ boundariness,
                                                uint32_t* __restrict__ bd, float* __restrict__ trad,
                                                const uint32_t* __restrict__ nearest,
                                                const uint32_t* __restrict__ transferring,
                                                const uint32_t* __restrict__ counts, Fsm3Args a,
                                                uint32_t* __restrict__ ptype, uint32_t* __restrict__ ps,
                                                uint32_t* __restrict__ pt, uint32_t* __restrict__ state,
                                                uint32_t* __restrict__ flag) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) {
        const int4 p = pos[i];
        const float r = radius[i];
        const float bdist = (float)bd[i] / POS_RES;
        float b = boundariness[i];
        const uint32_t nn = nearest[i];
        const int4 q = pos[nn];
        const float nnDist = length(vec_i(isub(q.x, p.x), isub(q.y, p.y), isub(q.z, p.z))) / POS_RES;
        const float nnR = radius[nn];
        const bool nnLarger = nnR > r || (nnR == r && nn > i);
        float tr;
        if (a.update_tr) {
            if (a.base_bd) {
                tr = (a.scale / (KERNEL_SCALE + KERNEL_SCALE * a.scale)) * bdist;
                tr = fmax_g(tr, a.smallest);
            } else {
                tr = a.smallest + fmax_g(0.0f, (bdist - a.offset) * a.scale);
            }
            trad[i] = tr;
        } else {
            tr = trad[i];
        }
        b = b >= 1.0f ? 1.0f : 0.0f;
        bd[i] = to_u32((bdist * (1.0f - b) + r * b) * POS_RES);
        boundariness[i] = b;
        const bool split = a.split && (tr * a.split_factor <= r);
        const bool merge = a.merge && nnDist < r && pow_f(r, a.dim) + pow_f(nnR, a.dim) <= pow_f(tr, a.dim);
        uint32_t type = merge ? P_MERGE : (split ? P_SPLIT : P_NONE);
        uint32_t s = (merge && nnLarger) ? nn : i;
        uint32_t t = (merge && nnLarger) ? i : nn;
        // A claim on a transferring particle, or a merge with itself, fails
        // without leaving a trace (the source claim is released at :95).
        bool cand = type != P_NONE && transferring[s] == 0u &&
                    (type == P_SPLIT || (transferring[t] == 0u && s != t));
        ptype[i] = type;
        ps[i] = s;
        pt[i] = t;
        state[i] = cand ? CL_UNDECIDED : CL_NONE;
        flag[i] = cand ? 1u : 0u;
    }
}

__global__ __launch_bounds__(BLOCK) void k_scatter_list(const uint32_t* __restrict__ flag,
                                                        const uint32_t* __restrict__ scan,
                                                        const uint32_t* __restrict__ counts,
                                                        uint32_t* __restrict__ list) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) if (flag[i]) list[scan[i]] = i;
}

__global__ __launch_bounds__(BLOCK) void k_claim_reset(const uint32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ ps,
                                                       const uint32_t* __restrict__ pt,
                                                       const uint32_t* __restrict__ ptype,
                                                       uint32_t* __restrict__ owner) {
    const uint32_t np = counts[C_PROPOSALS];
    GRID_STRIDE(k, np) {
        uint32_t i = list[k];
        owner[ps[i]] = 0xFFFFFFFFu;
        if (ptype[i] == P_MERGE) owner[pt[i]] = 0xFFFFFFFFu;
    }
}
__global__ __launch_bounds__(BLOCK) void k_claim_owner(const uint32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ ps,
                                                       const uint32_t* __restrict__ pt,
                                                       const uint32_t* __restrict__ ptype,
                                                       const uint32_t* __restrict__ state,
                                                       uint32_t* __restrict__ owner) {
    const uint32_t np = counts[C_PROPOSALS];
    GRID_STRIDE(k, np) {
        uint32_t i = list[k];
        uint32_t st = state[i];
        if (st != CL_UNDECIDED && st != CL_WIN) continue;
        atomicMin(&owner[ps[i]], i);
        if (ptype[i] == P_MERGE) atomicMin(&owner[pt[i]], i);
    }
}
// owner_next (optional): the other owner buffer, reset here for the next
// round as k_claim_reset would (no kernel of this round reads it).
__global__ __launch_bounds__(BLOCK) void k_claim_decide(const uint32_t* __restrict__ list,
                                                        const uint32_t* __restrict__ counts,
                                                        const uint32_t* __restrict__ ps,
                                                        const uint32_t* __restrict__ pt,
                                                        const uint32_t* __restrict__ ptype,
                                                        uint32_t* __restrict__ state,
                                                        const uint32_t* __restrict__ owner,
                                                        uint32_t* __restrict__ owner_next) {
    const uint32_t np = counts[C_PROPOSALS];
    GRID_STRIDE(k, np) {
        uint32_t i = list[k];
        bool merge = ptype[i] == P_MERGE;
        if (owner_next) {
            owner_next[ps[i]] = 0xFFFFFFFFu;
            if (merge) owner_next[pt[i]] = 0xFFFFFFFFu;
        }
        if (state[i] != CL_UNDECIDED) continue;
        uint32_t os = owner[ps[i]];
        uint32_t ot = merge ? owner[pt[i]] : i;
        if (os == i && ot == i) {
            state[i] = CL_WIN;
        } else if ((os < i && __atomic_load_n(&state[os], __ATOMIC_RELAXED) == CL_WIN) ||
                   (merge && ot < i && __atomic_load_n(&state[ot], __ATOMIC_RELAXED) == CL_WIN)) {
            state[i] = CL_LOSE;
        }
    }
}

// Apply the winners' claims to the transferring flags and count what is left.
__global__ __launch_bounds__(BLOCK) void k_claim_apply(const uint32_t* __restrict__ list,
                                                       uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ ps,
                                                       const uint32_t* __restrict__ pt,
                                                       const uint32_t* __restrict__ ptype,
                                                       const uint32_t* __restrict__ state,
                                                       uint32_t* __restrict__ transferring) {
    const uint32_t np = counts[C_PROPOSALS];
    GRID_STRIDE(k, np) {
        uint32_t i = list[k];
        uint32_t st = state[i];
        if (st == CL_WIN) {
            transferring[ps[i]] = 1u;
            if (ptype[i] == P_MERGE) {
                transferring[pt[i]] = 1u;
                atomicAdd(&counts[C_MERGES], 1u);
            }
        } else if (st == CL_UNDECIDED) {
            atomicAdd(&counts[C_UNDECIDED], 1u);
        }
    }
}

// One thread: settle undecided proposals in id order, and, if the winning
// merges do not fit the transfer list, replay the whole sequence with the
// capacity clamp of find_split_and_merge_3.comp:98-104.
__global__ void k_claim_sequential(const uint32_t* __restrict__ list, uint32_t* __restrict__ counts,
                                   const uint32_t* __restrict__ ps, const uint32_t* __restrict__ pt,
                                   const uint32_t* __restrict__ ptype, uint32_t* __restrict__ state,
                                   uint32_t* __restrict__ transferring, const uint32_t* __restrict__ t0,
                                   uint32_t tcap) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t np = counts[C_PROPOSALS];
    const uint32_t tlen0 = counts[C_TRANSFERS];
    if (counts[C_UNDECIDED] != 0u) {
        uint32_t merges = counts[C_MERGES];
        for (uint32_t k = 0; k < np; ++k) {
            uint32_t i = list[k];
            if (state[i] != CL_UNDECIDED) continue;
            uint32_t s = ps[i], t = pt[i];
            if (transferring[s]) { state[i] = CL_LOSE; continue; }
            if (ptype[i] == P_MERGE) {
                if (transferring[t]) { state[i] = CL_LOSE; continue; }
                transferring[t] = 1u;
                ++merges;
            }
            transferring[s] = 1u;
            state[i] = CL_WIN;
        }
        counts[C_MERGES] = merges;
    }
    if (tlen0 + counts[C_MERGES] <= tcap) return;
    // Replay with the capacity clamp.
    counts[C_OVERFLOW] |= 2u;
    for (uint32_t k = 0; k < np; ++k) {
        uint32_t i = list[k];
        transferring[ps[i]] = t0[ps[i]];
        if (ptype[i] == P_MERGE) transferring[pt[i]] = t0[pt[i]];
    }
    uint32_t tlen = tlen0, merges = 0;
    for (uint32_t k = 0; k < np; ++k) {
        uint32_t i = list[k];
        uint32_t s = ps[i], t = pt[i];
        state[i] = CL_LOSE;
        if (transferring[s]) continue;
        if (ptype[i] == P_MERGE) {
            if (transferring[t]) continue;
            if (tlen >= tcap) continue;
            ++tlen;
            ++merges;
            transferring[t] = 1u;
        }
        transferring[s] = 1u;
        state[i] = CL_WIN;
    }
    counts[C_MERGES] = merges;
}

__global__ __launch_bounds__(BLOCK) void k_claim_flags(const uint32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ ptype,
                                                       const uint32_t* __restrict__ state,
                                                       uint32_t* __restrict__ mflag,
                                                       uint32_t* __restrict__ sflag) {
    const uint32_t np = counts[C_PROPOSALS];
    GRID_STRIDE(k, np) {
        uint32_t i = list[k];
        bool win = state[i] == CL_WIN;
        mflag[k] = (win && ptype[i] == P_MERGE) ? 1u : 0u;
        sflag[k] = (win && ptype[i] == P_SPLIT) ? 1u : 0u;
    }
}

// Append winning merges to the transfer list and winning splits to the split
// list, both in id order (find_split_and_merge_3.comp:98-116).
__global__ __launch_bounds__(BLOCK) void k_claim_emit(const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ counts,
                                                      const uint32_t* __restrict__ ps,
                                                      const uint32_t* __restrict__ pt,
                                                      const uint32_t* __restrict__ mflag,
                                                      const uint32_t* __restrict__ mscan,
                                                      const uint32_t* __restrict__ sflag,
                                                      const uint32_t* __restrict__ sscan,
                                                      uint32_t* __restrict__ tsrc, uint32_t* __restrict__ ttgt,
                                                      float* __restrict__ tttl, float merge_ttl,
                                                      uint32_t* __restrict__ split_list) {
    const uint32_t np = counts[C_PROPOSALS];
    const uint32_t tlen = counts[C_TRANSFERS];
    GRID_STRIDE(k, np) {
        uint32_t i = list[k];
        if (mflag[k]) {
            uint32_t o = tlen + mscan[k];
            tsrc[o] = ps[i];
            ttgt[o] = pt[i];
            tttl[o] = merge_ttl;
        }
        if (sflag[k]) split_list[sscan[k]] = i;
    }
}

__global__ void k_claim_counts(uint32_t* __restrict__ counts) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    counts[C_TRANSFERS] += counts[C_MERGES];
    counts[C_SPLITS] = counts[C_TMP0];
}

// remove_impossible_splits.comp:25-40, indexed_list::duplicate_these
// (indexed_list.h:141-152: the duplicate of split k goes to slot n + k with a
// copy of every particle and fluid attribute), initialize_split_particles.comp:
// 14-27, and the transfer-list appends of update_transfers.cpp:66-68.
__global__ __launch_bounds__(BLOCK) void k_split_finish(const uint32_t* __restrict__ split_list,
                                                        const uint32_t* __restrict__ counts,
                                                        uint32_t cap, uint32_t tcap, ParticleArrays a,
                                                        uint32_t* __restrict__ tsrc,
                                                        uint32_t* __restrict__ ttgt,
                                                        float* __restrict__ tttl, float split_ttl) {
    const uint32_t ns = counts[C_SPLITS];
    const uint32_t n = counts[C_N];
    const uint32_t tlen = counts[C_TRANSFERS];
    const uint32_t keep = min(min(ns, tcap - tlen), cap - n);
    GRID_STRIDE(k, ns) {
        uint32_t s = split_list[k];
        if (k >= keep) {
            a.transferring[s] = 0u;
            continue;
        }
        uint32_t d = n + k;
        int4 p = a.pos[s];
        float r = a.radius[s];
        a.vel[d] = a.vel[s];
        a.backup[d] = a.backup[s];
        a.transferring[d] = a.transferring[s];
        a.target_radius[d] = a.target_radius[s];
        a.kernel_width[d] = a.kernel_width[s];
        a.boundariness[d] = a.boundariness[s];
        a.boundary_distance[d] = a.boundary_distance[s];
        if (a.lam_prev) a.lam_prev[d] = a.lam_prev[s];
        if (a.sid) a.sid[d] = a.sid[s];
        p.x = iadd(p.x, to_i32(r * POS_RES * 0.1f));
        a.pos[d] = p;
        a.radius[d] = 0.0f;
        a.inv_mass[d] = SPLIT_INV_MASS;
        tsrc[tlen + k] = s;
        ttgt[tlen + k] = d;
        tttl[tlen + k] = split_ttl;
    }
}
__global__ void k_split_counts(uint32_t* __restrict__ counts, uint32_t cap, uint32_t tcap,
                               unsigned long long* __restrict__ totals) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t ns = counts[C_SPLITS], n = counts[C_N], tlen = counts[C_TRANSFERS];
    uint32_t keep = min(min(ns, tcap - tlen), cap - n);
    totals[3] += keep;  // particles created
    counts[C_SPLITS] = keep;
    counts[C_N] = n + keep;
    counts[C_TRANSFERS] = tlen + keep;
}
__global__ void k_zero_claim_counts(uint32_t* __restrict__ counts) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    counts[C_UNDECIDED] = 0;
    counts[C_MERGES] = 0;
}
__global__ __launch_bounds__(BLOCK) void k_copy_u32_n(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                                      const uint32_t* __restrict__ counts) {
    const uint32_t n = counts[C_N];
    GRID_STRIDE(i, n) dst[i] = src[i];
}
__global__ void k_set_slot_from(uint32_t* __restrict__ counts, uint32_t slot, const uint32_t* __restrict__ src) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counts[slot] = *src;
}

#define LAUNCH(kern, grid, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, s, __VA_ARGS__)
#define LAUNCH1(kern, ...) hipLaunchKernelGGL(kern, dim3(1), dim3(64), 0, s, __VA_ARGS__)

// ------------------------------------------------------------- subsystems
void Scene::kernel_width_from_bd() {  // pool.cpp:77-80
    hipStream_t s = stream;
    float factor = P.s.mTargetRadiusScaleFactor / POS_RES;
    LAUNCH(k_kw_from_bd, stride_grid(cap), boundary_distance[cur], kernel_width[cur], radius[cur], counts,
           factor, P.s.mKernelWidthAdaptionSpeed);
}

void Scene::remap_transfers_after_permute() {
    hipStream_t s = stream;
    LAUNCH(k_invert_perm, stride_grid(cap), perm, counts, u4);
    LAUNCH(k_remap_transfers, stride_grid(tcap), t_source[tcur], t_target[tcur], u4, counts);
}

// particle_transfer.cpp:10-28
void Scene::particle_transfer(float dt) {
    hipStream_t s = stream;
    uint32_t* del_p = u1;
    fill_u32_dev(del_p, 0u, counts + C_N, cap, s);
    LAUNCH(k_particle_transfer, stride_grid(tcap), t_source[tcur], t_target[tcur], t_time_left[tcur], counts,
           radius[cur], inv_mass[cur], transferring[cur], del_p, t_flag, dt, dim());
    // particle compaction map
    LAUNCH(k_invert_flags, stride_grid(cap), del_p, u2, counts);
    scan_u32(u2, u3, counts + C_N, cap, false, counts + C_KEEP, ws_mem, s);
    // transfer compaction (finished splits, and transfers whose particles went away)
    LAUNCH(k_transfer_keep, stride_grid(tcap), t_source[tcur], t_target[tcur], t_flag, del_p, counts, t_scan);
    scan_u32(t_scan, t_flag, counts + C_TRANSFERS, tcap, false, counts + C_TKEEP, ws_mem, s);
    LAUNCH(k_compact_transfers, stride_grid(tcap), t_scan, t_flag, u3, counts, t_source[tcur],
           t_target[tcur], t_time_left[tcur], t_source[1 - tcur], t_target[1 - tcur], t_time_left[1 - tcur]);
    tcur = 1 - tcur;
    LAUNCH(k_compact_particles, stride_grid(cap), del_p, u3, counts, arrays(cur), arrays(1 - cur));
    cur = 1 - cur;
    LAUNCH1(k_commit_compaction, counts, particle_substeps);
    pairs_valid = false;
}

// spread_kernel_width.cpp:12-26 on the CSR list built with range scale 1.5.
void Scene::spread_kernel_width() {
    hipStream_t s = stream;
    int on_t = P.s.mBaseKernelWidthOnTargetRadius;
    LAUNCH(k_kw_init, stride_grid(cap), radius[cur], target_radius[cur], counts, on_t, ukw,
           (const uint32_t*)nullptr);
    launch_set_n_plus_1(counts, s);
    bool tk = tbegin(K_SPREAD);
    launch_spread_rows(false, pos[cur], radius[cur], target_radius[cur], kernel_width[cur], counts, cap, row, col, pcap,
                       on_t, ukw, cnt, nullptr, nullptr, s);
    tend(tk);
    scan_u32(cnt, u0, counts + C_N_PLUS_1, cap + 1, false, nullptr, ws_mem, s);
    launch_spread_rows(true, pos[cur], radius[cur], target_radius[cur], kernel_width[cur], counts, cap, row, col, pcap,
                       on_t, ukw, nullptr, u0, col_tmp, s);
    std::swap(row, u0);
    std::swap(col, col_tmp);
    LAUNCH1(k_set_pairs_from_row, counts, row);
    LAUNCH(k_kw_gradual, stride_grid(cap), ukw, kernel_width[cur], counts, P.s.mKernelWidthAdaptionSpeed);
}

// The spreading fused into the pair build: kernel_width_init.comp:17-31 before
// the walk, uint_to_float_but_gradual.comp:27-36 after it.
void Scene::spread_begin() {
    hipStream_t s = stream;
    LAUNCH(k_kw_init, stride_grid(cap), radius[cur], target_radius[cur], counts, P.s.mBaseKernelWidthOnTargetRadius,
           ukw, (const uint32_t*)nullptr);
}
void Scene::spread_reset_if_fallback() {  // the exact walk spreads again from the initial widths
    hipStream_t s = stream;
    LAUNCH(k_kw_init, stride_grid(cap), radius[cur], target_radius[cur], counts, P.s.mBaseKernelWidthOnTargetRadius,
           ukw, counts + C_FALLBACK);
}
void Scene::spread_end() {
    hipStream_t s = stream;
    launch_fix_clamped_rows(counts, cap, pcap, row, s);
    LAUNCH1(k_set_pairs_from_row, counts, row);
    LAUNCH(k_kw_gradual, stride_grid(cap), ukw, kernel_width[cur], counts, P.s.mKernelWidthAdaptionSpeed);
}

// update_transfers.cpp:14-71
// With `with_density_error` the substep's density-error measurement rides on
// the first pass over the rows (k_fsm12_density_rows).
void Scene::update_transfers(bool with_density_error) {
    hipStream_t s = stream;
    uint32_t* old_bd = u0;
    uint32_t* nearest = u1;
    LAUNCH(k_copy_u32_n, stride_grid(cap), old_bd, boundary_distance[cur], counts);
    bool tk = tbegin(K_FSM12);
    if (with_density_error) {
        // the solver loop of this substep computed gc from the same widths,
        // radii and masses (it moves positions only)
        if (P.solver_iterations <= 0)
            launch_kernel_consts(kernel_width[cur], radius[cur], inv_mass[cur], counts, cap, P.s.mHeightKernelId,
                                 P.s.mGradientKernelId, dim(), gc, s);
        launch_fsm12_density_rows(pos[cur], inv_mass[cur], kernel_width[cur], gc, counts, cap, row, col, pcap,
                                  old_bd, boundary_distance[cur], nearest, P.s.mHeightKernelId,
                                  P.s.mGradientKernelId, dim(), part_sum, part_max, s);
        launch_density_error_finish(part_sum, part_max, err_blocks, counts, err_stats, s);
    } else {
        launch_fsm12_rows(pos[cur], counts, cap, row, col, pcap, old_bd, boundary_distance[cur], nearest, s);
    }
    tend(tk);
    Fsm3Args fa;
    fa.update_tr = P.s.mUpdateTargetRadius;
    fa.base_bd = P.s.mBaseKernelWidthOnBoundaryDistance;
    fa.split = P.s.mSplit;
    fa.merge = P.s.mMerge;
    fa.smallest = P.s.mSmallestTargetRadius;
    fa.offset = P.s.mTargetRadiusOffset;
    fa.scale = P.s.mTargetRadiusScaleFactor;
    fa.merge_duration = P.s.mMergeDuration;
    fa.split_factor = pow_f(2.0f, 1.0f / dim()) * 0.99f;
    fa.dim = dim();
    LAUNCH(k_fsm3, stride_grid(cap), pos[cur], radius[cur], boundariness[cur], boundary_distance[cur],
           target_radius[cur], nearest, transferring[cur], counts, fa, a_type, a_s, a_t, a_state, a_flag);
    // candidate list in id order
    scan_u32(a_flag, a_scan, counts + C_N, cap, false, counts + C_PROPOSALS, ws_mem, s);
    LAUNCH(k_scatter_list, stride_grid(cap), a_flag, a_scan, counts, a_list);
    LAUNCH(k_copy_u32_n, stride_grid(cap), a_t0, transferring[cur], counts);
    // owner buffers alternate between rounds, so each round's decide pass
    // resets the next round's buffer
    uint32_t* own[2] = {a_owner, a_owner2};
    LAUNCH(k_claim_reset, stride_grid(cap), a_list, counts, a_s, a_t, a_type, own[0]);
    for (int r = 0; r < claim_rounds; ++r) {
        LAUNCH(k_claim_owner, stride_grid(cap), a_list, counts, a_s, a_t, a_type, a_state, own[r & 1]);
        LAUNCH(k_claim_decide, stride_grid(cap), a_list, counts, a_s, a_t, a_type, a_state, own[r & 1],
               r + 1 < claim_rounds ? own[(r + 1) & 1] : nullptr);
    }
    LAUNCH1(k_zero_claim_counts, counts);
    LAUNCH(k_claim_apply, stride_grid(cap), a_list, counts, a_s, a_t, a_type, a_state, transferring[cur]);
    LAUNCH1(k_claim_sequential, a_list, counts, a_s, a_t, a_type, a_state, transferring[cur], a_t0, tcap);
    // emit merges (transfer list) and splits (split list) in id order
    LAUNCH(k_claim_flags, stride_grid(cap), a_list, counts, a_type, a_state, a_owner, a_flag);
    scan_u32(a_owner, u2, counts + C_PROPOSALS, cap, false, nullptr, ws_mem, s);
    scan_u32(a_flag, u3, counts + C_PROPOSALS, cap, false, counts + C_TMP0, ws_mem, s);
    float merge_ttl = fmax_g(0.0001f, P.s.mMergeDuration);
    LAUNCH(k_claim_emit, stride_grid(cap), a_list, counts, a_s, a_t, a_owner, u2, a_flag, u3,
           t_source[tcur], t_target[tcur], t_time_left[tcur], merge_ttl, u4);
    LAUNCH1(k_claim_counts, counts);
    if (!P.s.mSplit) return;  // update_transfers.cpp:59
    LAUNCH(k_split_finish, stride_grid(cap), u4, counts, cap, tcap, arrays(cur), t_source[tcur],
           t_target[tcur], t_time_left[tcur], -P.split_duration);
    LAUNCH1(k_split_counts, counts, cap, tcap, particle_substeps);
    pairs_valid = false;
}

// ------------------------------------------------------------ state I/O
#define HIP_OK(x)                                  \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) return APBF_ERR_HIP; \
    } while (0)

apbf_status Scene::set_state(uint32_t n, const int32_t* h_pos, const float* h_vel, const float* h_im,
                             const float* h_r, const int32_t* h_backup, const uint32_t* h_tr,
                             const uint32_t* h_fidx, const float* h_trad, const float* h_kw,
                             const float* h_b, const uint32_t* h_bd, uint32_t nt, const uint32_t* h_ts,
                             const uint32_t* h_tt, const float* h_ttl, float last_delta_time) {
    if (n > cap || nt > tcap) return APBF_ERR_CAPACITY;
    if (n && (!h_pos || !h_vel || !h_im || !h_r || !h_backup || !h_tr || !h_trad || !h_kw || !h_b || !h_bd))
        return APBF_ERR_INVALID_ARGUMENT;
    if (nt && (!h_ts || !h_tt || !h_ttl)) return APBF_ERR_INVALID_ARGUMENT;
    // fluid id k lives in slot k: slot k takes hidden particle fidx[k]
    std::vector<uint32_t> order(n), inv(n, 0xFFFFFFFFu);
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t h = h_fidx ? h_fidx[k] : k;
        if (h >= n || inv[h] != 0xFFFFFFFFu) return APBF_ERR_INVALID_ARGUMENT;
        order[k] = h;
        inv[h] = k;
    }
    std::vector<int4> p4(n), b4(n);
    std::vector<float4> v4(n);
    std::vector<float> im(n), r(n);
    std::vector<uint32_t> tr(n);
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t h = order[k];
        p4[k] = make_int4(h_pos[4 * h], h_pos[4 * h + 1], h_pos[4 * h + 2], h_pos[4 * h + 3]);
        b4[k] = make_int4(h_backup[4 * h], h_backup[4 * h + 1], h_backup[4 * h + 2], h_backup[4 * h + 3]);
        v4[k] = make_float4(h_vel[4 * h], h_vel[4 * h + 1], h_vel[4 * h + 2], h_vel[4 * h + 3]);
        im[k] = h_im[h];
        r[k] = h_r[h];
        tr[k] = h_tr[h];
    }
    std::vector<uint32_t> ts(nt), tt(nt);
    for (uint32_t t = 0; t < nt; ++t) {
        if (h_ts[t] >= n || h_tt[t] >= n) return APBF_ERR_INVALID_ARGUMENT;
        ts[t] = inv[h_ts[t]];
        tt[t] = inv[h_tt[t]];
    }
    HIP_OK(hipStreamSynchronize(stream));
    HIP_OK(hipMemcpy(pos[cur], p4.data(), (size_t)n * 16, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(backup[cur], b4.data(), (size_t)n * 16, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(vel[cur], v4.data(), (size_t)n * 16, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(inv_mass[cur], im.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(radius[cur], r.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(transferring[cur], tr.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(target_radius[cur], h_trad, (size_t)n * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(kernel_width[cur], h_kw, (size_t)n * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(boundariness[cur], h_b, (size_t)n * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(boundary_distance[cur], h_bd, (size_t)n * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemset(lam_prev[cur], 0, (size_t)cap * 4));
    if (nt) {
        HIP_OK(hipMemcpy(t_source[tcur], ts.data(), (size_t)nt * 4, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(t_target[tcur], tt.data(), (size_t)nt * 4, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(t_time_left[tcur], h_ttl, (size_t)nt * 4, hipMemcpyHostToDevice));
    }
    uint32_t c[C_COUNT] = {0};
    c[C_N] = n;
    c[C_N_PLUS_1] = n + 1;
    c[C_N4] = 4 * n;
    c[C_TRANSFERS] = nt;
    HIP_OK(hipMemcpy(counts, c, sizeof(c), hipMemcpyHostToDevice));
    // no neighbour list yet: empty rows
    HIP_OK(hipMemset(row, 0, (size_t)(cap + 1) * 4));
    last_dt = last_delta_time;
    pairs_valid = false;
    HIP_OK(hipDeviceSynchronize());
    return APBF_OK;
}

apbf_status Scene::get_state(int32_t* h_pos, float* h_vel, float* h_im, float* h_r, int32_t* h_backup,
                             uint32_t* h_tr, float* h_trad, float* h_kw, float* h_b, uint32_t* h_bd,
                             uint32_t* h_ts, uint32_t* h_tt, float* h_ttl, uint32_t* h_pairs,
                             float* h_ldt) {
    uint32_t c[C_COUNT];
    HIP_OK(hipStreamSynchronize(stream));
    HIP_OK(hipMemcpy(c, counts, sizeof(c), hipMemcpyDeviceToHost));
    uint32_t n = c[C_N], nt = c[C_TRANSFERS];
    if (h_pos) HIP_OK(hipMemcpy(h_pos, pos[cur], (size_t)n * 16, hipMemcpyDeviceToHost));
    if (h_vel) HIP_OK(hipMemcpy(h_vel, vel[cur], (size_t)n * 16, hipMemcpyDeviceToHost));
    if (h_im) HIP_OK(hipMemcpy(h_im, inv_mass[cur], (size_t)n * 4, hipMemcpyDeviceToHost));
    if (h_r) HIP_OK(hipMemcpy(h_r, radius[cur], (size_t)n * 4, hipMemcpyDeviceToHost));
    if (h_backup) HIP_OK(hipMemcpy(h_backup, backup[cur], (size_t)n * 16, hipMemcpyDeviceToHost));
    if (h_tr) HIP_OK(hipMemcpy(h_tr, transferring[cur], (size_t)n * 4, hipMemcpyDeviceToHost));
    if (h_trad) HIP_OK(hipMemcpy(h_trad, target_radius[cur], (size_t)n * 4, hipMemcpyDeviceToHost));
    if (h_kw) HIP_OK(hipMemcpy(h_kw, kernel_width[cur], (size_t)n * 4, hipMemcpyDeviceToHost));
    if (h_b) HIP_OK(hipMemcpy(h_b, boundariness[cur], (size_t)n * 4, hipMemcpyDeviceToHost));
    if (h_bd) HIP_OK(hipMemcpy(h_bd, boundary_distance[cur], (size_t)n * 4, hipMemcpyDeviceToHost));
    if (nt) {
        if (h_ts) HIP_OK(hipMemcpy(h_ts, t_source[tcur], (size_t)nt * 4, hipMemcpyDeviceToHost));
        if (h_tt) HIP_OK(hipMemcpy(h_tt, t_target[tcur], (size_t)nt * 4, hipMemcpyDeviceToHost));
        if (h_ttl) HIP_OK(hipMemcpy(h_ttl, t_time_left[tcur], (size_t)nt * 4, hipMemcpyDeviceToHost));
    }
    if (h_pairs) {
        // rows cover the particles present when the list was built (later splits
        // append particles without neighbours, as the reference's pair list)
        uint32_t rows = std::min(c[C_ROWS], cap);
        std::vector<uint32_t> hr(rows + 1);
        HIP_OK(hipMemcpy(hr.data(), row, (size_t)(rows + 1) * 4, hipMemcpyDeviceToHost));
        uint32_t np = c[C_PAIRS];
        std::vector<uint32_t> hc(np);
        if (np) HIP_OK(hipMemcpy(hc.data(), col, (size_t)np * 4, hipMemcpyDeviceToHost));
        uint64_t w = 0;
        for (uint32_t i = 0; i < rows; ++i) {
            uint32_t b = std::min(hr[i], np), e = std::min(hr[i + 1], np);
            for (uint32_t k = b; k < e && w < np; ++k) {
                h_pairs[2 * w] = i;
                h_pairs[2 * w + 1] = hc[k];
                ++w;
            }
        }
    }
    if (h_ldt) *h_ldt = last_dt;
    return APBF_OK;
}

}  // namespace apbf
