"""Scene parameters and initial particle states for the benchmark workloads.

The settings dict carries the fields of the reference's ``apbf_settings``
uniform block (shaders/cpu_gpu_shared_config.h:74-94) with the defaults of
source/settings.cpp:5-31 for a 3D build. Scene dicts add the host-side knobs of
source/settings.cpp (basicPbf, solverIterations, splitDuration) and the scene
wiring of source/pool.cpp / source/waterdrop.cpp (walls, gravity, hash domain,
capacities).

Initial states are plain numpy dicts (see ``STATE_KEYS``): positions are int32
fixed point with 2^18 units per world unit (cpu_gpu_shared_config.h:7), h = 1
world unit.
"""
from __future__ import annotations

import math

import numpy as np

POS_RESOLUTION = 262144.0
KERNEL_SCALE = 4.0

REFERENCE_SETTINGS = dict(
    mHeightKernelId=1, mGradientKernelId=1, mMerge=1, mSplit=1,
    mBaseKernelWidthOnTargetRadius=1, mBaseKernelWidthOnBoundaryDistance=1,
    mUpdateTargetRadius=1, mUpdateBoundariness=1, mNeighborListSorted=0,
    mBoundarinessCalculationMethod=2, mBoundarinessAdaptionSpeed=0.5,
    mKernelWidthAdaptionSpeed=0.01, mBoundarinessSelfGradLengthFactor=2.0,
    mBoundarinessUnderpressureFactor=4.0, mMergeDuration=2.0,
    mSmallestTargetRadius=1.0, mTargetRadiusOffset=10.0, mTargetRadiusScaleFactor=0.3,
)

STATE_KEYS = ("position", "velocity", "inverse_mass", "radius", "pos_backup", "transferring",
              "fluid_index", "target_radius", "kernel_width", "boundariness",
              "boundary_distance", "t_source", "t_target", "t_time_left")


def walls(box_min, box_max, radius, ceiling=True, thickness=4.0, dims: int = 3):
    """Wall boxes around [box_min, box_max] (pool.cpp:30-40, waterdrop.cpp:28-38);
    in 2D without the two z walls (pool.cpp:37-40)."""
    mn = np.asarray(box_min, np.float32)
    mx = np.asarray(box_max, np.float32)
    t = np.float32(thickness)
    wmin = mn - t
    wmax = mx + t + np.array([0, radius * 2, 0], np.float32)
    b = [
        (wmin, np.array([wmin[0] + t, wmax[1], wmax[2]], np.float32)),
        (wmin, np.array([wmax[0], wmin[1] + t, wmax[2]], np.float32)),
        (np.array([wmax[0] - t, wmin[1], wmin[2]], np.float32), wmax),
    ]
    if ceiling:
        b.append((np.array([wmin[0], wmax[1] - t, wmin[2]], np.float32), wmax))
    if dims > 2:
        b.append((wmin, np.array([wmax[0], wmax[1], wmin[2] + t], np.float32)))
        b.append((np.array([wmin[0], wmin[1], wmax[2] - t], np.float32), wmax))
    return [(tuple(float(x) for x in lo), tuple(float(x) for x in hi)) for lo, hi in b]


def base_params(**kw) -> dict:
    p = dict(
        settings=dict(REFERENCE_SETTINGS),
        basic_pbf=0, solver_iterations=3, split_duration=0.0,
        gravity=(0.0, -9.81, 0.0),
        hash_min=(0.0, 0.0, 0.0), hash_max=(1.0, 1.0, 1.0), hash_res_log2=4,
        boxes=[], cap_particles=0, cap_transfers=10000, cap_pairs=0,
        dt=1.0 / 120.0,
    )
    p.update(kw)
    return p


def _fixed(xyz: np.ndarray) -> np.ndarray:
    """World float -> int32 fixed point (truncation, as ivec3(pos * 2^18))."""
    return np.trunc(xyz.astype(np.float64) * POS_RESOLUTION).astype(np.int32)


def make_state(pos_world: np.ndarray, radius: np.ndarray, bd: np.ndarray | None = None,
               boundariness: np.ndarray | None = None, target_radius: np.ndarray | None = None,
               velocity: np.ndarray | None = None, dims: int = 3) -> dict:
    """Inverse mass 1 / (2r)^D (initialize.cpp:25,60 at inverse density 1)."""
    n = len(pos_world)
    radius = np.asarray(radius, np.float32)
    pos = np.zeros((n, 4), np.int32)
    pos[:, :3] = _fixed(pos_world)
    vel = np.zeros((n, 4), np.float32)
    if velocity is not None:
        vel[:, :3] = velocity
    inv_mass = (np.float32(1.0) / (np.float32(2.0) * radius) ** dims).astype(np.float32)
    st = dict(
        position=pos, velocity=vel, inverse_mass=inv_mass, radius=radius.copy(),
        pos_backup=pos.copy(), transferring=np.zeros(n, np.uint32),
        fluid_index=np.arange(n, dtype=np.uint32),
        target_radius=(radius if target_radius is None else target_radius).astype(np.float32).copy(),
        kernel_width=(radius * np.float32(KERNEL_SCALE)).astype(np.float32),
        boundariness=(np.ones(n, np.float32) if boundariness is None else boundariness.astype(np.float32)),
        boundary_distance=(np.trunc(radius.astype(np.float64) * POS_RESOLUTION).astype(np.uint32)
                           if bd is None else np.trunc(bd * POS_RESOLUTION).astype(np.uint32)),
        t_source=np.zeros(0, np.uint32), t_target=np.zeros(0, np.uint32),
        t_time_left=np.zeros(0, np.float32),
        last_delta_time=1.0,
    )
    return st


def lattice(lo, counts, spacing, jitter, rng) -> np.ndarray:
    """Particle centres lo + spacing/2 + spacing*(i,j,k) (initialize_box.comp:33-38),
    each coordinate jittered by U(-jitter, jitter)."""
    counts = [int(c) for c in counts]
    g = np.stack(np.meshgrid(*[np.arange(c) for c in counts], indexing="ij"), -1).reshape(-1, 3)
    p = np.asarray(lo, np.float64) + spacing * 0.5 + spacing * g
    if jitter > 0:
        p = p + rng.uniform(-jitter, jitter, p.shape)
    return p


# ------------------------------------------------------------------ workloads
def dam_break(n_side: int = 32, spacing: float = 0.5, seed: int = 0, box_scale=(4, 2, 2),
              iterations: int = 3, res: int | None = None):
    """Plain PBF dam break: an n_side^3 lattice at `spacing` (jitter +-0.05h) in
    the low corner of a box whose interior is box_scale times the lattice extent.
    basicPbf, Gauss kernel, r = spacing/2, kernel width 4r = 2*spacing."""
    rng = np.random.default_rng(seed)
    r = spacing / 2.0
    extent = n_side * spacing
    box_max = (extent * box_scale[0], extent * box_scale[1], extent * box_scale[2])
    pos = lattice((0.0, 0.0, 0.0), (n_side, n_side, n_side), spacing, 0.05, rng)
    n = len(pos)
    st = make_state(pos, np.full(n, r, np.float32))
    if res is None:
        # largest z-order resolution whose cells are at least one kernel width wide on every axis
        res = int(math.floor(math.log2(min(box_max) / (4 * r))))
    s = dict(REFERENCE_SETTINGS)
    s["mSmallestTargetRadius"] = r
    p = base_params(
        settings=s, basic_pbf=1, solver_iterations=iterations,
        hash_min=(0.0, 0.0, 0.0), hash_max=box_max, hash_res_log2=res,
        boxes=walls((0, 0, 0), box_max, r, ceiling=True),
        cap_particles=n, cap_transfers=16, cap_pairs=n * 64,
        workload=f"dam_break_{n}",
    )
    return p, st


def _sphere_points(center, radius_world, spacing, jitter, rng, dims: int = 3):
    """Lattice points inside a ball (in 2D: a disc in the x-y plane)."""
    m = int(math.ceil(radius_world / spacing)) + 1
    g = np.arange(-m, m + 1) * spacing
    X, Y, Z = np.meshgrid(g, g, g if dims == 3 else np.zeros(1), indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], -1)
    pts = pts[np.linalg.norm(pts, axis=1) <= radius_world - spacing * 0.5]
    if jitter > 0:
        pts = pts + rng.uniform(-jitter, jitter, pts.shape)
        if dims == 2:
            pts[:, 2] = 0.0
    return pts + np.asarray(center, np.float64)


def water_drop(pool_counts=(150, 40, 150), drop_particles: int = 100_000, r_coarse: float = 1.0,
               r_fine: float = 0.5, seed: int = 0, iterations: int = 4, res: int = 6, cap_factor: float = 1.5,
               pairs_per_particle: int = 800, target_scale: float | None = None, merge_duration: float = 2.0,
               dims: int = 3):
    """Adaptive water drop: a sphere of `drop_particles` fine particles (radius
    r_fine = 0.5h) released at t = 0 above a pool of coarse particles (radius
    r_coarse = 1h, a pool_counts lattice at spacing 2h: 150 x 40 x 150 = 900 000
    particles in a 300 x 80 x 300 h box). Split, merge and kernel-width spreading
    on (baseKernelWidthOnBoundaryDistance off, neighbour range scale 1.5),
    4 solver iterations. smallestTargetRadius = r_coarse (the reference scenes use
    their base radius, pool.cpp:52 / waterdrop.cpp:54); target radius
    r_coarse + s * boundary distance with targetRadiusOffset 0 and s chosen so it
    reaches 2 r_coarse at the pool's centre depth, so radii stay in [0.5h, 2h].
    Initial boundary distances are the geometric distances to the pool faces /
    the drop surface, boundariness 1 within one particle diameter of them.
    The pair capacity, 800 per particle of the particle capacity (slots of
    about 1 200 entries per particle in the neighbour walk), leaves almost no
    row longer than its slot, so the exact refill has nearly nothing to do;
    the list never reaches the capacity, so the results do not depend on it
    (400: 7.98 against 7.86 ms per substep, same states).

    dims=2 builds the planar version of the reference's default build
    (DIMENSIONS 2): one layer of pool particles at z = r_coarse (pool_counts[2]
    is ignored), a disc-shaped drop in the same plane, no z walls, inverse
    masses 1 / (2r)^2."""
    rng = np.random.default_rng(seed)
    sc = 2 * r_coarse
    if dims == 2:
        pool_counts = (pool_counts[0], pool_counts[1], 1)
    pool_max = tuple(c * sc for c in pool_counts)
    pool = lattice((0, 0, 0), pool_counts, sc, 0.05, rng)
    sf = 2 * r_fine
    if dims == 2:
        pool[:, 2] = r_coarse
        drop_r = math.sqrt(drop_particles * sf ** 2 / math.pi) + sf * 0.5
        center = (pool_max[0] * 0.5, pool_max[1] + 8.0 + drop_r, r_coarse)
    else:
        drop_r = (drop_particles * sf ** 3 * 3.0 / (4.0 * math.pi)) ** (1.0 / 3.0) + sf * 0.5
        center = (pool_max[0] * 0.5, pool_max[1] + 8.0 + drop_r, pool_max[2] * 0.5)
    drop = _sphere_points(center, drop_r, sf, 0.05, rng, dims)
    pos = np.concatenate([pool, drop], 0)
    radius = np.concatenate([np.full(len(pool), r_coarse), np.full(len(drop), r_fine)]).astype(np.float32)
    faces = [pool[:, 0], pool_max[0] - pool[:, 0], pool[:, 1], pool_max[1] - pool[:, 1]]
    if dims == 3:
        faces += [pool[:, 2], pool_max[2] - pool[:, 2]]
    d_pool = np.minimum.reduce(faces)
    d_drop = drop_r - np.linalg.norm(drop - np.asarray(center), axis=1)
    bd = np.maximum(np.concatenate([d_pool, d_drop]), radius)
    if target_scale is None:
        target_scale = r_coarse / (0.5 * min(pool_max[:dims]))
    s = dict(REFERENCE_SETTINGS)
    s.update(mBaseKernelWidthOnBoundaryDistance=0, mSmallestTargetRadius=r_coarse,
             mTargetRadiusOffset=0.0, mTargetRadiusScaleFactor=target_scale, mMergeDuration=merge_duration)
    tr = (r_coarse + bd * target_scale).astype(np.float32)
    bness = (bd < 2.0 * radius).astype(np.float32)
    st = make_state(pos, radius, bd=bd, boundariness=bness, target_radius=tr, dims=dims)
    n = len(pos)
    top = center[1] + drop_r + 4
    cap = int(n * cap_factor)
    p = base_params(
        settings=s, basic_pbf=0, solver_iterations=iterations,
        hash_min=(-4.0, -4.0, -4.0), hash_max=(pool_max[0] + 4.0, top, pool_max[2] + 4.0), hash_res_log2=res,
        boxes=walls((0, 0, 0), pool_max, r_coarse, ceiling=False, dims=dims),
        cap_particles=cap, cap_transfers=max(1000, n // 4), cap_pairs=cap * pairs_per_particle,
        workload=f"water_drop{'_2d' if dims == 2 else ''}_{n}", dimensions=dims,
    )
    return p, st


def _subset(st: dict, keep: np.ndarray) -> dict:
    """The particles `keep` of a state without transfers, fluid ids renumbered."""
    out = {k: (st[k][:0] if k.startswith("t_") else st[k][keep]) for k in STATE_KEYS if k != "fluid_index"}
    out["fluid_index"] = np.arange(int(np.count_nonzero(keep)), dtype=np.uint32)
    out["last_delta_time"] = st["last_delta_time"]
    return out


def _join(a: dict, b: dict) -> dict:
    """Two states as one: b's particles after a's, b's transfer endpoints moved with them."""
    na = len(a["radius"])
    out = {}
    for k in STATE_KEYS:
        if k == "fluid_index":
            out[k] = np.arange(na + len(b["radius"]), dtype=np.uint32)
        elif k in ("t_source", "t_target"):
            out[k] = np.concatenate([a[k], b[k] + np.uint32(na)]).astype(np.uint32)
        else:
            out[k] = np.concatenate([a[k], b[k]])
    out["last_delta_time"] = a["last_delta_time"]
    return out


def settled_water_drop(pool_counts=(180, 50, 180), drop_particles: int = 100_000, settle: int = 300,
                       drop_settle: int = 100, gap: float = 3.0, merge_duration: float = 0.05, seed: int = 0,
                       iterations: int = 4, res: int = 6, pairs_per_particle: int = 800):
    """The adaptive water drop from rest (runs the library on the GPU).

    The coarse pool of :func:`water_drop` first runs `settle` substeps on its
    own with split and merge on: it settles under gravity and coarsens
    towards its adaptive equilibrium (merged particles of up to 2h in the
    deep interior, the surface at the base radius), 1.62 M lattice particles
    becoming about 0.9 M. The fine drop settles separately for `drop_settle`
    substeps without gravity (the reference holds its drop in place until it
    is released, waterdrop.cpp:129-131). The drop is then put `gap` units
    above the pool's highest particle and released. Both parts start this
    workload at their rest density, so its density error is the solver's
    steady error, and particles split and merge in every substep. Merges take
    `merge_duration` seconds (6 substeps at dt = 1/120; the reference's
    default is 2 s). The pair capacity of 800 per particle makes the
    neighbour walk's fixed-stride slots (capacity / current particles, about
    1 200 entries) longer than the longest rows (820 entries), so no row
    needs the exact refill (7.14 against 8.17 ms per substep at 300, whose
    slots of 450 entries leave some rows to the refill every substep).
    Returns (params, state, report)."""
    from .scene import Scene
    p, st = water_drop(pool_counts=pool_counts, drop_particles=drop_particles, seed=seed, iterations=iterations,
                       res=res, pairs_per_particle=pairs_per_particle, merge_duration=merge_duration,
                       cap_factor=1.0)
    n0 = len(st["radius"])
    fine = st["radius"] < np.float32(0.75)
    p["cap_transfers"] = max(1000, n0 // 4)
    dt = p["dt"]
    report = {"lattice_pool": int(np.count_nonzero(~fine)), "drop": int(np.count_nonzero(fine)),
              "settle_substeps": settle, "drop_settle_substeps": drop_settle}
    parts = []
    for keep, steps, gravity in ((~fine, settle, p["gravity"]), (fine, drop_settle, (0.0, 0.0, 0.0))):
        q = dict(p, gravity=gravity)
        sc = Scene(q)
        sc.set_state(_subset(st, keep))
        for _ in range(steps):
            sc.update(dt)
        parts.append(sc.get_state(pairs=False))
        sc.close()
    pool, drop = parts
    n_pool = len(pool["radius"])
    pool["fluid_index"] = pool["fluid_index"][:n_pool]
    drop["fluid_index"] = drop["fluid_index"][: len(drop["radius"])]
    top = int(pool["position"][:, 1].max())
    shift = np.int32(top + int(gap * POS_RESOLUTION) - int(drop["position"][:, 1].min()))
    for k in ("position", "pos_backup"):
        drop[k] = drop[k].copy()
        drop[k][:, 1] += shift
    both = _join(pool, drop)
    both["last_delta_time"] = pool["last_delta_time"]
    n = len(both["radius"])
    p["cap_particles"] = int(n * 1.5)
    p["cap_pairs"] = p["cap_particles"] * pairs_per_particle
    p["cap_transfers"] = max(1000, n // 4)
    p["hash_max"] = (p["hash_max"][0], max(p["hash_max"][1], top / POS_RESOLUTION + 40.0), p["hash_max"][2])
    p["workload"] = f"water_drop_settled_{n}"
    report.update(settled_pool=n_pool, settled_drop=len(drop["radius"]), particles=n)
    return p, both, report


def small_water_drop(seed: int = 0, pool_counts=None, drop_particles: int | None = None, res: int | None = None,
                     active: bool = False, dims: int = 3):
    """A ~8k-particle version of :func:`water_drop` for parity tests (2D: a
    80 x 30 pool layer and a 600-particle disc, 3 000 particles). With
    active=True the smallest target radius is the fine radius, so most pool
    particles split (a stress case for split, transfer and claim handling)."""
    if pool_counts is None:
        pool_counts = (24, 10, 24) if dims == 3 else (80, 30, 1)
    if drop_particles is None:
        drop_particles = 2000 if dims == 3 else 600
    if res is None:
        res = 4 if dims == 3 else 6
    p, st = water_drop(pool_counts=pool_counts, drop_particles=drop_particles, seed=seed, res=res,
                       pairs_per_particle=400 if active else 200, cap_factor=3.0 if active else 2.0, dims=dims)
    if active:
        p["settings"]["mSmallestTargetRadius"] = 0.5
        p["settings"]["mTargetRadiusScaleFactor"] = 0.05
    return p, st


def box_shape(mn, mx, radius: float, dims: int = 3):
    """initialize::add_box_shape (initialize.cpp:5-33, initialize_box.comp):
    floor((max - min) / 2r) particles per axis, centred at min + r + 2r * index.
    In 2D the z range is first narrowed to [mid - r, mid + 1.5 r] around the
    box's middle, one layer (initialize.cpp:16-19)."""
    mn = np.asarray(mn, np.float64).copy()
    mx = np.asarray(mx, np.float64).copy()
    if dims < 3:
        mid = (mn[2] + mx[2]) / 2.0
        mn[2], mx[2] = mid - radius, mid + radius * 1.5
    counts = np.floor((mx - mn) / (2.0 * radius)).astype(int)
    g = np.stack(np.meshgrid(*[np.arange(c) for c in counts], indexing="ij"), -1).reshape(-1, 3)
    return mn + radius + 2.0 * radius * g


def _sphere_shells(center, radius: float, particle_radius: float, dims: int = 3):
    """initialize::append_sphere (initialize.cpp:71-100): the points of a sphere
    surface, recursively as rings of circles of the next lower dimension."""
    c = np.asarray(center, np.float64)
    if dims <= 1:
        if radius < particle_radius:
            return [c.copy()]
        off = np.array([radius, 0.0, 0.0])
        return [c - off, c + off]
    d = dims - 1
    pts = _sphere_shells(c, radius, particle_radius, d)
    count = math.floor(radius * math.pi / (2.0 * particle_radius))
    if count <= 0:
        return pts
    da = math.pi / count
    angle = 0.0
    for _ in range(int(count) // 2):
        angle += da
        layer_off = math.sin(angle) * radius
        layer_r = math.cos(angle) * radius
        lc = c.copy()
        lc[d] += layer_off
        pts += _sphere_shells(lc, layer_r, particle_radius, d)
        lc[d] -= 2 * layer_off
        pts += _sphere_shells(lc, layer_r, particle_radius, d)
    return pts


def sphere_shape(center, outer_radius: float, particle_radius: float, dims: int = 3):
    """initialize::add_sphere_shape(center, outer, 0, true, r) (initialize.cpp:35-69):
    concentric shells from the outside in, each shell's radius chosen so the
    particles placed so far fill the volume between it and the outer surface
    (in 2D: rings in the x-y plane filling the disc's area)."""
    a = outer_radius ** dims
    b = -(particle_radius ** dims) * 2 * dims / math.pi
    shell = outer_radius - particle_radius
    out = []
    while 0.0 <= shell <= outer_radius - particle_radius:
        out += _sphere_shells(center, shell, particle_radius, dims)
        if shell < particle_radius:
            break  # the centre particle is placed
        v = a + b * len(out)
        shell = (v ** (1.0 / dims) if v >= 0 else math.nan) - particle_radius  # pow of a negative: NaN, the loop ends
    return np.array(out, np.float64)


def _reference_scene(pos, r, res, hash_min, hash_max, boxes, name, cap_pairs_per_particle=64, dims=3):
    n = len(pos)
    st = make_state(pos, np.full(n, r, np.float32), dims=dims)
    s = dict(REFERENCE_SETTINGS)
    s["mSmallestTargetRadius"] = r
    if dims == 2:
        s["mTargetRadiusScaleFactor"] = 0.4  # settings.cpp:26
    p = base_params(
        settings=s, basic_pbf=0, solver_iterations=3, gravity=(0.0, -10.0, 0.0),
        hash_min=tuple(float(x) for x in hash_min), hash_max=tuple(float(x) for x in hash_max), hash_res_log2=res,
        boxes=boxes, cap_particles=max(2 * n, 16), cap_transfers=max(n, 16),
        cap_pairs=max(2 * n, 16) * cap_pairs_per_particle, workload=f"{name}{'_2d' if dims == 2 else ''}_{n}",
        dimensions=dims,
    )
    return p, st


def pool_scene(mn=(-40.0, -10.0, -80.0), mx=(40.0, 30.0, -40.0), radius: float = 1.0, res: int = 4,
               dims: int = 3):
    """The reference's pool (pool.cpp:6-56 as main.cpp:62 builds it): a box of
    particles filling [mn, mx], six walls 4 units thick around it, gravity
    (0, -10, 0), the hash domain [mn, mx] at 2^4 cells per axis, the reference
    settings (adaptive, boundary-distance kernel widths). dims=2 is the
    reference's default build (DIMENSIONS 2, cpu_gpu_shared_config.h:2): one
    particle layer, four walls, targetRadiusScaleFactor 0.4."""
    pos = box_shape(mn, mx, radius, dims)
    return _reference_scene(pos, radius, res, mn, mx, walls(mn, mx, radius, ceiling=True, dims=dims), "pool",
                            dims=dims)


def waterfall_scene(mn=(-40.0, -10.0, -80.0), mx=(40.0, 30.0, -40.0), radius: float = 1.0, res: int = 4,
                    dims: int = 3):
    """The reference's waterfall (waterfall.cpp:6-67): the pool's particles and
    an upper basin whose x-max side is open, over a lower basin shifted by
    (min - max) in x and y; eleven wall boxes in 3D, seven in 2D (no z walls,
    waterfall.cpp:41-47)."""
    pos = box_shape(mn, mx, radius, dims)
    t = 4.0
    w0 = np.asarray(mn, np.float64) - t
    w1 = np.asarray(mx, np.float64) + t + np.array([0.0, 2 * radius, 0.0])
    shift = w0 - w1
    shift[2] = 0.0
    b = [(w0, np.array([w0[0] + t, w1[1], w1[2]])),
         (w0, np.array([w1[0], w0[1] + t, w1[2]])),
         (np.array([w1[0] - t, w0[1], w0[2]]), w1),
         (np.array([w0[0], w1[1] - t, w0[2]]), w1),
         (shift + w0, np.array([w0[0] + t, w1[1], w1[2]]) + shift),
         (shift + w0, np.array([w1[0], w0[1] + t, w1[2]]) + shift),
         (shift + np.array([w1[0] - t, w0[1], w0[2]]), w1 + shift)]
    if dims > 2:
        b += [(w0, np.array([w1[0], w1[1], w0[2] + t])),
              (np.array([w0[0], w0[1], w1[2] - t]), w1),
              (shift + w0, np.array([w1[0], w1[1], w0[2] + t]) + shift),
              (shift + np.array([w0[0], w0[1], w1[2] - t]), w1 + shift)]
    boxes = [(tuple(float(x) for x in lo), tuple(float(x) for x in hi)) for lo, hi in b]
    return _reference_scene(pos, radius, res, mn, mx, boxes, "waterfall", dims=dims)


def spherical_pool_scene(center=(0.0, 10.0, -60.0), pool_radius: float = 100.0, particle_radius: float = 1.0,
                         res: int = 4, max_particles: int | None = None, dims: int = 3):
    """The reference's spherical pool (spherical_pool.cpp:6-45): a ball of
    particles on concentric shells inside a hollow sphere collider of the pool
    radius, gravity (0, -10, 0), the hash domain center +- pool radius at 2^4
    cells per axis. max_particles keeps the first particles (the reference's
    100 000-particle capacity clamps a 3D ball of radius 100)."""
    pos = sphere_shape(center, pool_radius, particle_radius, dims)
    if max_particles is not None:
        pos = pos[:max_particles]
    c = np.asarray(center, np.float64)
    p, st = _reference_scene(pos, particle_radius, res, c - pool_radius, c + pool_radius, [], "spherical_pool",
                             dims=dims)
    p["sphere"] = (tuple(float(x) for x in center), float(pool_radius), True)
    return p, st


def large_pool(box=(200.0, 100.0, 100.0), n_target: int = 8_000_000, spacing: float | None = None,
               kernel_width: float = 1.0, impulse_width: float = 20.0, impulse_speed: float = 4.0,
               seed: int = 0, iterations: int = 3, res: int = 8, pairs_per_particle: int = 128,
               jitter: float = 0.01):
    """Large uniform pool, plain PBF, at rest density: particles of radius
    r = kernel_width/4 and inverse mass 1/(2r)^3 (initialize.cpp:25) on a
    lattice at spacing 2.068 r (jitter +-0.01h), filling the box floor to the
    depth that holds ~n_target particles (386 x 107 x 193 = 7 971 286 at the
    defaults, depth 55h). The reference's own box lattice at spacing 2r starts
    5.4% compressed under the Gaussian kernel (the oracle's density of a
    perfect lattice); at 2.068 r an interior particle's density is the rest
    density, and the mean compression error of the jittered lattice is 0.07%.
    A wave impulse: the slab x < impulse_width moves at impulse_speed along
    +x. The velocity is inferred from the displacement over the last substep
    (velocity_handling.cpp:26), so it enters as pos_backup = pos - v dt with
    last_delta_time = dt. The hash domain is the cube [0, max(box)]^3 at
    2^res cells per axis (cells 0.78h at the defaults, so a particle's search
    box spans 3-4 cells per axis instead of 3 cells of 1.56h at 2^7).
    128 pairs per particle of capacity keep every row inside its slot of the
    neighbour walk."""
    rng = np.random.default_rng(seed)
    r = kernel_width / KERNEL_SCALE
    if spacing is None:
        spacing = 2.068 * r
    nx = int(box[0] / spacing)
    nz = int(box[2] / spacing)
    ny = max(1, int(round(n_target / (nx * nz))))
    pos = lattice((0.0, 0.0, 0.0), (nx, ny, nz), spacing, jitter * kernel_width, rng)
    n = len(pos)
    st = make_state(pos, np.full(n, r, np.float32))
    dt = 1.0 / 120.0
    moving = pos[:, 0] < impulse_width
    shift = np.zeros((n, 4), np.int32)
    shift[moving, 0] = np.int32(math.trunc(impulse_speed * dt * POS_RESOLUTION))
    st["pos_backup"] = st["position"] - shift
    st["last_delta_time"] = dt
    side = float(max(box))
    s = dict(REFERENCE_SETTINGS)
    s["mSmallestTargetRadius"] = r
    p = base_params(
        settings=s, basic_pbf=1, solver_iterations=iterations, dt=dt,
        hash_min=(0.0, 0.0, 0.0), hash_max=(side, side, side), hash_res_log2=res,
        boxes=walls((0, 0, 0), box, r, ceiling=True),
        cap_particles=n, cap_transfers=16, cap_pairs=n * pairs_per_particle,
        workload=f"large_pool_{n}",
    )
    return p, st


def batch_scenes(count: int = 64, n_min: int = 10_000, n_max: int = 500_000, seed: int = 0):
    """Independent scenes for batched stepping: particle counts drawn
    log-uniformly from [n_min, n_max] (seeded), alternating a plain dam break
    (basic PBF lattice) and an adaptive water drop whose pool and drop are
    scaled to the drawn count (one target-radius scale for all drops, so the
    scenes of each kind share their settings and can be packed into one
    list). Returns a list of (params, state)."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(n_min), math.log(n_max), count)).round().astype(int)
    out = []
    for k, n in enumerate(counts):
        if k % 2 == 0:
            side = max(4, int(round(n ** (1.0 / 3.0))))
            p, st = dam_break(n_side=side, seed=seed + k)
        else:
            # pool of a x b x a coarse particles with b = 0.27 a, plus a drop of n / 10
            a = max(6, int(round((0.9 * n / 0.27) ** (1.0 / 3.0))))
            b = max(3, int(round(0.27 * a)))
            p, st = water_drop(pool_counts=(a, b, a), drop_particles=max(100, int(n) // 10), seed=seed + k,
                               res=5 if n < 100_000 else 6, pairs_per_particle=300, cap_factor=1.5,
                               target_scale=0.05)
        p["workload"] = f"batch_{k}_{p.get('workload', '')}"
        out.append((p, st))
    return out
