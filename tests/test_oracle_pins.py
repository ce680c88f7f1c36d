"""Pins of the oracle's split/merge find, its kernel width from boundary
distance and the warm-start network's gradient, by what the mathematics fixes
(brute force over the pair list, per-particle closed forms, the maximality of
a greedy claim schedule, central differences), and proof that each pin bites:
every check runs again on an oracle built with one planted mistake
(ORACLE_MUTATION in oracle/apbf_oracle.cpp) and must fail there. CPU only.

Reference lines restated by the checks:
  find_split_and_merge_1.comp:17-35  bd_i = min_j(bd_old_j + uint(|x_j - x_i|)), minD_i = min_j uint(|x_j - x_i|)
  find_split_and_merge_2.comp:19-38  nearest_i = some j with uint(|x_j - x_i|) == minD_i
  find_split_and_merge_3.comp:67-121 target radius (both modes), boundary-distance decay, predicates, claims
  uint_to_float_with_indexed_lower_bound.comp:20-44 kernel width = max(step toward bd * f, 4 r)
"""
import numpy as np
import pytest

from apbf_amd import config as cfg
from oracle import mlp_ref
from oracle.oracle import OracleScene, build_variant

F = np.float32
RES = F(262144.0)
U32 = 0xFFFFFFFF


def _dist_fixed(pos, a, b):
    """uint(length(vec3(pos_b - pos_a))) in fixed-point units: int32 differences
    (wrapping), float32 squares summed x, y, then z, the float32 square root,
    truncated to uint."""
    d = (pos[b, :3].astype(np.int64) - pos[a, :3].astype(np.int64)).astype(np.int32).astype(F)
    sq = ((d[:, 0] * d[:, 0]) + (d[:, 1] * d[:, 1])) + d[:, 2] * d[:, 2]
    return np.floor(np.sqrt(sq.astype(F))).astype(np.uint64), np.sqrt(sq.astype(F))


# (boundary-distance mode, DIMENSIONS) of the split/merge cases
CASES = [(False, 3), (True, 3), (False, 2), (True, 2)]


def _case(bd_mode: bool, offset: float = 3.0, seed: int = 11, library=None, substeps: int = 0, dims: int = 3):
    """A small water drop whose drop is squeezed to 0.45 of its size (nearest
    neighbours closer than the radius: many merges competing for the same
    particles) and whose every third pool particle has radius 1.4 (splits
    wherever the target radius is below 1.4 / 1.247; in 2D 1.6, below
    1.6 / 1.400), run
    stage by stage up to update_transfers after `substeps` full substeps, in
    3D or in the planar build. Returns the parameters, the scene, the states
    before and after update_transfers and the intermediates of the find."""
    p, st = cfg.small_water_drop(seed=seed, dims=dims)
    n = len(st["radius"])
    p["cap_transfers"] = 8 * n  # no claim is dropped by a capacity clamp
    p["cap_particles"] = 8 * n
    s = p["settings"]
    s["mBaseKernelWidthOnBoundaryDistance"] = int(bd_mode)
    s["mTargetRadiusOffset"] = float(offset)
    if bd_mode:
        s["mTargetRadiusScaleFactor"] = 0.3
    st = dict(st)
    pos = st["position"][:, :3].astype(np.float64)
    fine = st["radius"] < 0.75
    c = pos[fine].mean(0)
    pos[fine] = c + (pos[fine] - c) * 0.45
    st["position"] = st["position"].copy()
    st["position"][:, :3] = np.trunc(pos).astype(np.int32)
    st["pos_backup"] = st["position"].copy()
    big = (~fine) & (np.arange(n) % 3 == 0)
    st["radius"] = st["radius"].copy()
    st["radius"][big] = 1.4 if dims == 3 else 1.6
    st["inverse_mass"] = (np.float32(1.0) / (np.float32(2.0) * st["radius"]) ** dims).astype(np.float32)
    o = OracleScene(p, library=library)
    o.set_state(st)
    dt = p["dt"]
    for _ in range(substeps):
        o.update(dt)
    o.stage("velocity", dt)
    o.stage("particle_transfer", dt)
    if bd_mode:
        o.stage("kernel_width_bd", dt)
    o.stage("neighborhood", dt)
    if not bd_mode:
        o.stage("spread", dt)
    for _ in range(p["solver_iterations"]):
        o.stage("box", dt)
        o.stage("incompressibility", dt)
    before = o.get_state()
    o.stage("update_transfers", dt)
    after = o.get_state()
    return p, o, before, after, o.fsm()


def pin_flood_fill(library=None):
    """find_split_and_merge_1: the flood-filled distance and the shortest
    neighbour distance equal a brute-force minimum over each particle's pairs;
    particles without pairs keep UINT_MAX."""
    for bd_mode, dims in CASES:
        p, o, before, after, (bd_fill, min_d, _) = _case(bd_mode, library=library, dims=dims)
        assert np.array_equal(before["fluid_index"], np.arange(len(before["fluid_index"])))
        pairs = before["pairs"].astype(np.int64)
        a, b = pairs[:, 0], pairs[:, 1]
        dist, _ = _dist_fixed(before["position"], a, b)
        nf = len(before["fluid_index"])
        want_bd = np.full(nf, U32, np.uint64)
        np.minimum.at(want_bd, a, (before["boundary_distance"][b].astype(np.uint64) + dist) & U32)
        want_min = np.full(nf, U32, np.uint64)
        np.minimum.at(want_min, a, dist)
        assert len(pairs) > 10 * nf
        assert np.array_equal(bd_fill.astype(np.uint64), want_bd)
        assert np.array_equal(min_d.astype(np.uint64), want_min)


def pin_nearest(library=None):
    """find_split_and_merge_2: the nearest neighbour of a particle with pairs is
    one of its neighbours at the shortest distance (ties: any of them; the
    oracle takes the last in list order); without pairs it is the particle."""
    for bd_mode, dims in CASES:
        p, o, before, after, (_, min_d, nearest) = _case(bd_mode, library=library, dims=dims)
        pairs = before["pairs"].astype(np.int64)
        a, b = pairs[:, 0], pairs[:, 1]
        dist, _ = _dist_fixed(before["position"], a, b)
        nf = len(before["fluid_index"])
        want_min = np.full(nf, U32, np.uint64)
        np.minimum.at(want_min, a, dist)
        tied = dist == want_min[a]
        ok = np.zeros(nf, bool)
        hit = tied & (b == nearest[a].astype(np.int64))
        ok[a[hit]] = True
        has = np.bincount(a, minlength=nf) > 0
        assert np.all(ok[has])
        assert np.array_equal(nearest[~has], np.nonzero(~has)[0].astype(np.uint32))
        # the documented tie rule: the last tied pair in row order
        last = np.full(nf, -1, np.int64)
        last[a[tied]] = b[tied]  # fancy assignment keeps the last write per index
        assert np.array_equal(nearest[has].astype(np.int64), last[has])


def _split_factor(dims):
    return F(F(2.0 ** float(F(1.0) / F(dims))) * F(0.99))  # pow(2.0, 1.0 / DIMENSIONS) * 0.99


def _target_radius(s, bdist, bd_mode):
    f, rmin, off = F(s["mTargetRadiusScaleFactor"]), F(s["mSmallestTargetRadius"]), F(s["mTargetRadiusOffset"])
    if bd_mode:
        return np.maximum((f / (F(4.0) + F(4.0) * f)).astype(F) * bdist, rmin).astype(F)
    return (rmin + np.maximum(F(0.0), ((bdist - off).astype(F) * f).astype(F))).astype(F)


def pin_find_closed_forms(library=None):
    """find_split_and_merge_3 per particle: the target radius in both modes
    (:69-74), the boundary-distance decay toward the radius for boundary
    particles (:79-81), the snapped boundariness, and the split and merge
    predicates (:90-91) of every new split and merge, except where the deciding
    quantity lies within 1e-6 (relative) of its threshold."""
    for bd_mode, dims in CASES:
        p, o, before, after, (bd_fill, _, nearest) = _case(bd_mode, library=library, dims=dims)
        s = p["settings"]
        nf = len(before["fluid_index"])
        r = before["radius"].astype(F)
        bdist = (bd_fill.astype(F) / RES).astype(F)
        tr = _target_radius(s, bdist, bd_mode)
        assert np.array_equal(after["target_radius"][:nf], tr)
        bnd = before["boundariness"] >= F(1.0)
        assert np.array_equal(after["boundariness"][:nf], bnd.astype(F))
        want_bd = np.where(bnd, np.floor(r * RES), np.floor(bdist * RES)).astype(np.uint64)
        assert np.array_equal(after["boundary_distance"][:nf].astype(np.uint64), want_bd)
        assert (~bnd).sum() > 0 and bnd.sum() > 0
        # predicates, float64 with a margin around each threshold
        split_q = tr.astype(np.float64) * float(_split_factor(dims)) - r
        idx = np.arange(nf)
        _, nn_len = _dist_fixed(before["position"], idx, nearest.astype(np.int64))
        nn_d = nn_len.astype(np.float64) / 262144.0
        rn = r[nearest].astype(np.float64)
        vol_q = r.astype(np.float64) ** dims + rn ** dims - tr.astype(np.float64) ** dims
        near = (np.abs(split_q) <= 1e-6 * r) | (np.abs(vol_q) <= 1e-6 * tr.astype(np.float64) ** dims) | \
               (np.abs(nn_d - r) <= 1e-6 * r)
        split = split_q <= 0
        merge = (nn_d < r) & (vol_q <= 0)
        nt0 = len(before["t_source"])
        src, tgt, ttl = after["t_source"][nt0:], after["t_target"][nt0:], after["t_time_left"][nt0:]
        m = ttl > 0
        splits = o.splits()
        assert m.sum() > 0 and len(splits) > 0
        for i in splits:  # a split particle satisfies the split predicate and not the merge predicate
            assert near[i] or (split[i] and not merge[i]), i
        for sidx, tidx in zip(src[m], tgt[m]):
            # a merge is decided by one of its two particles, whose nearest neighbour is the other
            deciders = [i for i in (sidx, tidx) if i < nf and nearest[i] == (tidx if i == sidx else sidx)]
            assert deciders, (sidx, tidx)
            assert any(near[i] or merge[i] for i in deciders), (sidx, tidx)
            for i in deciders:  # orientation: a larger (or equal, higher-slot) neighbour is the source
                j = int(nearest[i])
                larger = r[j] > r[i] or (r[j] == r[i] and j > i)
                if merge[i] and not near[i]:
                    assert (sidx, tidx) == ((j, i) if larger else (i, j)), (sidx, tidx, i)


def pin_claims_legal_and_maximal(library=None):
    """find_split_and_merge_3 claims (:94-121): no particle takes part in two
    transfers, exactly the claimed particles are flagged, new merges have a
    positive time left and splits a non-positive one, and the schedule is
    maximal: every particle whose predicate holds is claimed or was blocked by
    a particle already flagged (source or target of its transfer). A particle
    without pairs is its own nearest neighbour; its merge claims the same
    particle twice, which the second exchange refuses, so it is left out."""
    for bd_mode, dims in CASES:
        p, o, before, after, (bd_fill, _, nearest) = _case(bd_mode, library=library, dims=dims)
        s = p["settings"]
        nf = len(before["fluid_index"])
        src, tgt, ttl = after["t_source"], after["t_target"], after["t_time_left"]
        ids = np.concatenate([src, tgt])
        assert len(np.unique(ids)) == len(ids)
        flagged = np.nonzero(after["transferring"])[0]
        assert np.array_equal(np.sort(flagged), np.sort(ids))
        nt0 = len(before["t_source"])
        assert np.all(ttl[nt0:][np.isin(src[nt0:], o.splits(), invert=True)] > 0)
        assert np.all(ttl[nt0:][np.isin(src[nt0:], o.splits())] <= 0)
        r = before["radius"].astype(F)
        tr = after["target_radius"][:nf]
        idx = np.arange(nf)
        _, nn_len = _dist_fixed(before["position"], idx, nearest.astype(np.int64))
        nn_d = nn_len.astype(np.float64) / 262144.0
        vol_q = r.astype(np.float64) ** dims + r[nearest].astype(np.float64) ** dims - tr.astype(np.float64) ** dims
        split_q = tr.astype(np.float64) * float(_split_factor(dims)) - r
        margin = (np.abs(split_q) <= 1e-6 * r) | (np.abs(vol_q) <= 1e-6 * tr.astype(np.float64) ** dims)
        lonely = nearest == idx
        merge = (nn_d < r) & (vol_q <= 0) & ~margin & ~lonely
        split = (split_q <= 0) & ~((nn_d < r) & (vol_q <= 0)) & ~margin
        fl = after["transferring"][:nf] == 1
        assert np.all(fl[merge] | fl[nearest[merge]])
        assert np.all(fl[split])
        assert merge.sum() > 0 and split.sum() > 0
        # no capacity clamp in this case: every requested split was kept
        assert o.counts()[2] < p["cap_transfers"] and o.counts()[0] < p["cap_particles"]


def pin_kernel_width_from_bd(library=None):
    """uint_to_float_with_indexed_lower_bound.comp:20-44: the new width is the
    multiplicative step of 1% toward bd * (f / 2^18), snapped to the target when
    the step crosses it, and at least 4 r. Known answers by hand, then the
    closed form over random states."""
    p, st = cfg.pool_scene(mn=(-8.0, -4.0, -8.0), mx=(8.0, 4.0, 8.0))
    p["settings"]["mTargetRadiusScaleFactor"] = 0.25
    n = len(st["radius"])
    st = dict(st)
    # by hand: f = 0.25, bd = 4 * 2^18 -> target 1.0 exactly
    st["boundary_distance"] = np.full(n, 4 * 262144, np.uint32)
    st["radius"] = np.full(n, 0.2, np.float32)
    kw = np.full(n, 1.005, np.float32)
    kw[1] = 2.0   # 1.98 after one step down
    kw[2] = 0.5   # 0.505 after one step up (radius 0.1: lower bound 0.4)
    kw[3] = 0.995  # crosses 1.0 upward: snapped to 1.0
    st["kernel_width"] = kw
    st["radius"][2] = 0.1
    st["radius"][4] = 0.3  # lower bound 1.2 above the stepped width
    o = OracleScene(p, library=library)
    o.set_state(st)
    o.stage("kernel_width_bd", 0.0)
    got = o.get_state()["kernel_width"]
    assert got[0] == F(1.0) and got[3] == F(1.0)
    assert got[1] == F(F(2.0) * F(0.99))
    assert got[2] == F(F(0.5) * F(1.01)) and got[2] < F(1.0)
    assert got[4] == F(F(0.3) * F(4.0))
    # closed form over random states
    rng = np.random.default_rng(5)
    st["boundary_distance"] = rng.integers(0, 40 * 262144, n).astype(np.uint32)
    st["radius"] = rng.uniform(0.1, 1.5, n).astype(np.float32)
    st["kernel_width"] = rng.uniform(0.2, 12.0, n).astype(np.float32)
    o = OracleScene(p, library=library)
    o.set_state(st)
    o.stage("kernel_width_bd", 0.0)
    got = o.get_state()["kernel_width"]
    val = (st["boundary_distance"].astype(F) * F(F(0.25) / RES)).astype(F)
    old = st["kernel_width"]
    step = F(p["settings"]["mKernelWidthAdaptionSpeed"])
    moved = np.where(val > old, old * (F(1) + step), np.where(val < old, old * (F(1) - step), old)).astype(F)
    crossed = (old < val) != (moved < val)
    want = np.maximum(np.where(crossed, val, moved), (st["radius"] * F(4.0)).astype(F)).astype(F)
    assert np.array_equal(got, want)
    assert crossed.sum() > 0 and (~crossed).sum() > 0


def pin_mlp_gradient(mutation: int = 0):
    """The warm-start network's gradient (oracle/mlp_ref.py) with the BF16
    roundings switched off equals central differences of its float64 loss."""
    rng = np.random.default_rng(3)
    params = (rng.standard_normal(mlp_ref.OFF_B3 + 1) * 0.3).astype(np.float32)
    norm = np.concatenate([rng.standard_normal(16) * 0.1, rng.uniform(0.5, 2.0, 16), [1.7]]).astype(np.float32)
    x = rng.standard_normal((40, 16)).astype(np.float32)
    t = rng.standard_normal(40).astype(np.float32)
    old = mlp_ref._MUTATION
    mlp_ref._MUTATION = mutation
    try:
        g, _ = mlp_ref.gradient(params, norm, x, t, exact=True)
    finally:
        mlp_ref._MUTATION = old
    p64 = params.astype(np.float64)
    eps = 1e-6
    # every bias and output weight, and a sample of the matrix weights
    pick = np.concatenate([np.arange(mlp_ref.OFF_B1, mlp_ref.OFF_W2), np.arange(mlp_ref.OFF_B2, mlp_ref.OFF_B3 + 1),
                           rng.choice(mlp_ref.OFF_B1, 120, replace=False),
                           mlp_ref.OFF_W2 + rng.choice(4096, 200, replace=False)])
    for k in pick:
        a, b = p64.copy(), p64.copy()
        a[k] += eps
        b[k] -= eps
        fd = (mlp_ref.loss(a, norm, x, t) - mlp_ref.loss(b, norm, x, t)) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


# --------------------------------------------- solver loop: tolerance and warm start
def _solver_case(library=None, iterations=4):
    """A compressed dam break (lattice at 0.48 h: the density starts above rest)
    after one substep, ready for the next one's stages."""
    p, st = cfg.dam_break(n_side=9, spacing=0.48, iterations=iterations)
    o = OracleScene(p, library=library)
    o.set_state(st)
    o.update(p["dt"])
    return p, o.get_state()


def _stages_until_solver(o, dt):
    o.stage("velocity", dt)
    o.stage("neighborhood", dt)


def pin_tolerance_loop(library=None):
    """Tolerance mode against the same loop driven from outside through the
    oracle's stages and its density-error pass: before each iteration the box
    collision runs, then the mean error of the current positions decides -
    at or below the tolerance nothing more moves; else one incompressibility
    pass. Also: a tolerance below every error gives the fixed count."""
    p, st = _solver_case(library)
    dt = p["dt"]
    ref = OracleScene(p, library=library)
    ref.set_state(st)
    _stages_until_solver(ref, dt)
    errs = []
    for _ in range(8):
        ref.stage("box", dt)
        errs.append(ref.density_error()[0])
        ref.stage("incompressibility", dt)
    tol = 0.5 * (errs[2] + errs[3])  # three iterations are above it, the fourth starts below
    assert errs[2] > tol >= errs[3]
    manual = OracleScene(p, library=library)
    manual.set_state(st)
    _stages_until_solver(manual, dt)
    applied = 0
    for k in range(8):
        manual.stage("box", dt)
        if manual.density_error()[0] <= tol:
            break
        manual.stage("incompressibility", dt)
        applied += 1
    auto = OracleScene(p, library=library)
    auto.set_state(st)
    auto.set_solver_tolerance(tol, 8)
    auto.update(dt)
    a, m = auto.get_state(), manual.get_state()
    assert np.array_equal(a["position"], m["position"])
    assert auto.solver_stats()[0] == applied == 3
    # a tolerance no error reaches: the fixed count, bit for bit
    fixed = OracleScene(dict(p, solver_iterations=5), library=library)
    fixed.set_state(st)
    fixed.update(dt)
    tiny = OracleScene(p, library=library)
    tiny.set_state(st)
    tiny.set_solver_tolerance(1e-30, 5)
    tiny.update(dt)
    assert np.array_equal(fixed.get_state()["position"], tiny.get_state()["position"])
    assert np.array_equal(fixed.get_state()["boundariness"].view(np.uint32),
                          tiny.get_state()["boundariness"].view(np.uint32))


def pin_warm_start(library=None):
    """The warm-start shift pass and its judgement, against the stage sequence:
    zero predicted lambdas move nothing, so the substep equals box, box, then
    the plain iterations with a box before each (or, when the second box
    collision alone raised the error, one iteration fewer); lambdas that raise
    the compression are undone, and the substep equals box, box, then one
    iteration fewer."""
    p, st = _solver_case(library)
    dt = p["dt"]
    iters = p["solver_iterations"]

    def staged(n_solves):
        o = OracleScene(p, library=library)
        o.set_state(st)
        _stages_until_solver(o, dt)
        o.stage("box", dt)
        for k in range(n_solves):
            o.stage("box", dt)
            o.stage("incompressibility", dt)
        return o.get_state()["position"]

    n = len(st["radius"])
    # zero lambdas: the judgement compares the error after the second box
    # collision with the error after the first one (the warm pass moves nothing)
    probe = OracleScene(p, library=library)
    probe.set_state(st)
    _stages_until_solver(probe, dt)
    probe.stage("box", dt)
    e0 = probe.density_error()[0]
    probe.stage("box", dt)
    undone = probe.density_error()[0] > e0
    zero = OracleScene(p, library=library)
    zero.set_state(st)
    zero.set_warm_lambdas(np.zeros(n, np.float32))
    zero.update(dt)
    want = staged(iters - 1) if undone else staged(iters)
    assert np.array_equal(zero.get_state()["position"], want)
    assert zero.solver_stats() == ((iters - 1, 1, 1) if undone else (iters, 1, 0))
    # a few particles with a far too large lambda push their neighbours into
    # the others: the compression after the warm shift exceeds the one before
    rng = np.random.default_rng(0)
    bad_lam = np.where(rng.random(n) < 0.02, -20.0, 0.0).astype(np.float32)
    bad = OracleScene(p, library=library)
    bad.set_state(st)
    bad.set_warm_lambdas(bad_lam)
    bad.update(dt)
    assert bad.solver_stats() == (iters - 1, 1, 1)
    assert np.array_equal(bad.get_state()["position"], staged(iters - 1))


def pin_planar_transfer(library=None):
    """particle_transfer.comp:37-84 in the planar build (DIMENSIONS 2): the
    volume a transfer moves is r^2, so the total of r^2 and the total mass
    are conserved up to float rounding (r^3 is not)."""
    p, st = cfg.small_water_drop(seed=9, active=True, dims=2)
    o = OracleScene(p, library=library)
    o.set_state(st)
    for _ in range(6):
        o.update(p["dt"])
        if o.counts()[2] > 0:
            break
    before = o.get_state()
    assert len(before["t_source"]) > 0
    o.stage("particle_transfer", p["dt"])
    after = o.get_state()
    m0 = (1.0 / before["inverse_mass"].astype(np.float64)).sum()
    m1 = (1.0 / after["inverse_mass"].astype(np.float64)).sum()
    a0 = (before["radius"].astype(np.float64) ** 2).sum()
    a1 = (after["radius"].astype(np.float64) ** 2).sum()
    assert abs(m1 - m0) <= 1e-5 * m0
    assert abs(a1 - a0) <= 1e-5 * a0
    assert not np.allclose(before["radius"][: len(after["radius"])], after["radius"])


def pin_planar_rest_density(library=None):
    """The planar rest density (incompressibility_2.comp:81, pow(2r, D) m^-1
    with D = 2) and the 2D Gaussian normalisation (kernels.glsl:84-89,104):
    a single layer of particles on a square lattice at spacing 2r with mass
    (2r)^2 is a Riemann sum of the unit-area kernel, so every particle's
    density is within 3 % of the rest density and one solver iteration moves
    no particle by more than 0.05 r."""
    p, st = cfg.pool_scene(dims=2, mn=(-20.0, -10.0, -20.0), mx=(20.0, 10.0, -10.0))
    o = OracleScene(p, library=library)
    o.set_state(st)
    o.stage("neighborhood", p["dt"])
    mean_err, max_err = o.density_error()
    assert mean_err < 0.03 and max_err < 0.03, (mean_err, max_err)
    before = o.get_state()
    o.stage("incompressibility", p["dt"])
    after = o.get_state()
    d = np.abs(after["position"][:, :3].astype(np.int64) - before["position"][:, :3].astype(np.int64)) / 262144.0
    assert d.max() <= 0.05 * float(st["radius"][0]), d.max()
    assert np.array_equal(after["position"][:, 2], before["position"][:, 2])  # the layer stays planar


PINS = {
    "planar_transfer": pin_planar_transfer,
    "planar_rest_density": pin_planar_rest_density,
    "tolerance_loop": pin_tolerance_loop,
    "warm_start": pin_warm_start,
    "flood_fill": pin_flood_fill,
    "nearest": pin_nearest,
    "find_closed_forms": pin_find_closed_forms,
    "claims": pin_claims_legal_and_maximal,
    "kernel_width_from_bd": pin_kernel_width_from_bd,
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_oracle_pin(name):
    PINS[name]()


def test_mlp_gradient_matches_central_differences():
    pin_mlp_gradient()


# mutation -> the pin that must fail on it (oracle/apbf_oracle.cpp, MUT(k))
MUTATIONS = {
    1: "flood_fill",          # the particle's own old distance instead of the neighbour's
    2: "flood_fill",          # shortest distance taken over the fill candidates
    3: "nearest",             # tie test against the neighbour's shortest distance
    4: "find_closed_forms",   # target radius offset form without its clamp at zero
    5: "find_closed_forms",   # boundary-distance decay with the weights swapped
    6: "find_closed_forms",   # merge predicate without the distance condition
    7: "kernel_width_from_bd",  # no snap where the step crosses the target
    8: "kernel_width_from_bd",  # lower bound r instead of 4 r
    9: "claims",              # a merge does not check that its target is free
    10: "tolerance_loop",     # the iteration that reaches the tolerance still shifts
    11: "warm_start",         # the warm shift is undone when it helped, kept when it hurt
    12: "warm_start",         # no box collision between the warm shift and the first iteration
    13: "find_closed_forms",  # the 3D split factor 2^(1/3) in the planar build
    14: "planar_transfer",    # 3D volumes r^3 moved by planar transfers
    15: "planar_rest_density",  # the 3D rest density (2r)^3 m^-1 in the planar solver
}


@pytest.mark.parametrize("mutation", sorted(MUTATIONS))
def test_pin_fails_on_planted_mistake(mutation, tmp_path):
    lib = build_variant(mutation, str(tmp_path))
    with pytest.raises(AssertionError):
        PINS[MUTATIONS[mutation]](library=lib)


def test_mlp_gradient_pin_fails_on_planted_mistake():
    for mutation in (1, 2):  # 1: ReLU mask of layer 1 dropped; 2: W2 used untransposed
        with pytest.raises(AssertionError):
            pin_mlp_gradient(mutation)
