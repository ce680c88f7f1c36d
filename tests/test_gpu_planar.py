"""The planar build (DIMENSIONS 2, the reference's default,
cpu_gpu_shared_config.h:2) on the GPU against the oracle, bit for bit: the
2D z-curve hash and hash-grid walk operators, the reference's default
configuration (pool scene, binary-search neighbourhood, FIXED_TIME_STEP 1/60),
the other reference scenes, the adaptive water drop with splits and merges
(2D volumes, split factor and rest density), the cone and spike kernels'
2D constants, and packed batches of planar scenes."""
import numpy as np
import pytest

from apbf_amd import config as cfg
from oracle import oracle as O
from oracle.oracle import OracleScene
from statecmp import compare

pytestmark = pytest.mark.gpu


def _pair(p, st):
    from apbf_amd.scene import Scene
    g = Scene(p)
    g.set_state(st)
    r = OracleScene(p)
    r.set_state(st)
    return g, r


def _check(g, r, where, z0=None):
    gs, rs = g.get_state(), r.get_state()
    bad = compare(gs, rs, check_pairs=True)
    assert not bad, f"{where}: " + "; ".join(bad)
    if z0 is not None:  # one layer: nothing moves a particle off its plane
        assert np.all(gs["position"][:, 2] == z0), where
    return gs


def _stages(p):
    s = p["settings"]
    out = ["velocity", "particle_transfer"]
    out.append("kernel_width_bd" if s["mBaseKernelWidthOnBoundaryDistance"] else None)
    out.append("neighborhood")
    if not s["mBaseKernelWidthOnBoundaryDistance"]:
        out.append("spread")
    out += ["box", "incompressibility"] * p["solver_iterations"] + ["update_transfers"]
    return [x for x in out if x]


def test_planar_hash_and_walk_operators(device):
    import torch
    from apbf_amd import ops
    rng = np.random.default_rng(4)
    n = 6000
    pw = np.zeros((n, 3))
    pw[:, :2] = rng.uniform(0, 40, (n, 2))
    pw[:, 2] = rng.choice([5.0, 5.5, 7.0], n)
    p = np.zeros((n, 4), np.int32)
    p[:, :3] = np.trunc(pw * 262144).astype(np.int32)
    mn, mx = (0.0, 0.0, 0.0), (40.0, 40.0, 40.0)
    for res in (3, 6, 9, 13):
        got = ops.position_hash(torch.from_numpy(p).cuda(), mn, mx, res, dims=2).cpu().numpy().view(np.uint32)
        assert np.array_equal(got, O.position_hash(p, mn, mx, res, dims=2)), res
    res = 5
    h = O.position_hash(p, mn, mx, res, dims=2)
    order = np.argsort(h, kind="stable")
    ps, hs = p[order], h[order]
    rad = rng.uniform(0.8, 2.0, n).astype(np.float32)
    s, e = O.find_value_ranges(hs, 1 << (2 * res))
    want = O.neighborhood_green(ps, rad, s, e, 1.5, mn, mx, res, 4_000_000, dims=2)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()
    row, col, npairs = ops.neighborhood_green(t(ps), torch.from_numpy(rad).cuda(), t(s), t(e), 1.5, mn, mx, res,
                                              4_000_000, dims=2)
    k = int(npairs.item())
    row = row.cpu().numpy().view(np.uint32)
    col = col.cpu().numpy().view(np.uint32)[:k]
    first = np.repeat(np.arange(n, dtype=np.uint32), np.diff(row.astype(np.int64)))
    assert k == len(want) > 0
    assert np.array_equal(first, want[:, 0]) and np.array_equal(col, want[:, 1])


def test_planar_reference_default_build(device):
    """The reference as configured by default: DIMENSIONS 2, the pool scene
    (SCENE 0), the binary-search neighbourhood (NEIGHBORHOOD_TYPE 3), a fixed
    time step of 1/60 s (cpu_gpu_shared_config.h:2-12), 20 substeps."""
    p, st = cfg.pool_scene(dims=2)
    p["dt"] = 1.0 / 60.0
    g, r = _pair(p, st)
    g.set_neighborhood_method("binary_search")
    r.set_neighborhood_method("binary_search")
    z0 = st["position"][0, 2]
    for step in range(20):
        g.update(p["dt"])
        r.update(p["dt"])
        _check(g, r, f"substep {step}", z0)


@pytest.mark.parametrize("kind", ["pool", "waterfall", "spherical_pool"])
def test_planar_reference_scenes(device, kind):
    """The reference scenes in the planar build with the hash-grid walk:
    every stage of a substep, then six substeps."""
    if kind == "spherical_pool":
        p, st = cfg.spherical_pool_scene(pool_radius=30.0, dims=2)
    elif kind == "waterfall":
        p, st = cfg.waterfall_scene(dims=2)
    else:
        p, st = cfg.pool_scene(dims=2, res=6)
    g, r = _pair(p, st)
    z0 = st["position"][0, 2]
    for name in _stages(p) + (["sphere"] if kind == "spherical_pool" else ["box"]):
        g.stage(name, p["dt"])
        r.stage(name, p["dt"])
        _check(g, r, f"after {name}", z0)
    for step in range(6):
        g.update(p["dt"])
        r.update(p["dt"])
        _check(g, r, f"substep {step}", z0)


@pytest.mark.parametrize("active,steps", [(False, 30), (True, 12)])
def test_planar_water_drop(device, active, steps):
    """The adaptive water drop in the plane (spreading kernel widths; with
    active=True most pool particles split, so transfers move 2D volumes
    every substep)."""
    p, st = cfg.small_water_drop(seed=5, active=active, dims=2)
    g, r = _pair(p, st)
    z0 = st["position"][0, 2]
    for name in _stages(p):
        g.stage(name, p["dt"])
        r.stage(name, p["dt"])
        _check(g, r, f"after {name}", z0)
    transfers = 0
    for step in range(steps):
        g.update(p["dt"])
        r.update(p["dt"])
        _check(g, r, f"substep {step}", z0)
        transfers += g.counts()[1]
    assert transfers > 0


@pytest.mark.parametrize("kernels", [(3, 3), (1, 4), (4, 1)])
def test_planar_kernel_constants(device, kernels):
    """The cone and quadratic-spike kernels carry D-dependent constants
    (kernels.glsl:48-82): solver stages and substeps of the planar pool with
    them as height / gradient kernels."""
    p, st = cfg.pool_scene(dims=2, res=5)
    p["settings"]["mHeightKernelId"], p["settings"]["mGradientKernelId"] = kernels
    g, r = _pair(p, st)
    for step in range(4):
        g.update(p["dt"])
        r.update(p["dt"])
        _check(g, r, f"substep {step}")


def test_planar_packed_batch(device):
    """Four planar drops packed into one list, against the oracle per scene."""
    from apbf_amd.scene import PackedBatch
    cases = [cfg.small_water_drop(seed=s, dims=2, active=(s % 2 == 1)) for s in range(4)]
    for p, _ in cases:  # one packed list: the same settings for all
        p["settings"] = dict(cases[0][0]["settings"])
    dt = cases[0][0]["dt"]
    batch = PackedBatch(cases)
    assert len(batch.groups) == 1
    for _ in range(3):
        batch.update(dt)
    got = batch.scene_states()
    batch.close()
    for k, (p, st) in enumerate(cases):
        r = OracleScene(p)
        r.set_state(st)
        for _ in range(3):
            r.update(dt)
        bad = compare(got[k], r.get_state(), check_pairs=True)
        assert not bad, f"scene {k}: " + "; ".join(bad)
