"""GPU parity of whole substeps and of every stage of a substep against the
oracle, from identical input states. The float state is expected to match bit
for bit (shared exp/pow algorithms, no contraction, IEEE division and sqrt on
both sides); the north-star bound 1e-5 h on positions is checked as well.
Transfers are compared as sets (their list order carries no meaning)."""
import numpy as np
import pytest

from apbf_amd import config as cfg
from oracle.oracle import OracleScene
from statecmp import compare

pytestmark = pytest.mark.gpu

PLAIN_STAGES = ["velocity", "neighborhood", "box", "incompressibility", "box", "incompressibility",
                "box", "incompressibility"]


def adaptive_stages(p):
    s = p["settings"]
    out = ["velocity", "particle_transfer"]
    if s["mBaseKernelWidthOnBoundaryDistance"]:
        out.append("kernel_width_bd")
    out.append("neighborhood")
    if not s["mBaseKernelWidthOnBoundaryDistance"]:
        out.append("spread")
    out += ["box", "incompressibility"] * p["solver_iterations"]
    out.append("update_transfers")
    return out


def _pair(p, st):
    from apbf_amd.scene import Scene
    g = Scene(p)
    g.set_state(st)
    r = OracleScene(p)
    r.set_state(st)
    return g, r


def _check(g, r, where, pairs=True):
    gs, rs = g.get_state(), r.get_state()
    bad = compare(gs, rs, check_pairs=pairs)
    assert not bad, f"{where}: " + "; ".join(bad)
    # north-star bound on positions (1e-5 h, h = 1 world unit = 2^18 units)
    d = np.abs(gs["position"][:, :3].astype(np.int64) - rs["position"][:, :3].astype(np.int64))
    assert d.max(initial=0) <= 2, where
    return gs, rs


def test_dam_break_stages(device):
    p, st = cfg.dam_break(n_side=16)
    g, r = _pair(p, st)
    dt = p["dt"]
    for step in range(3):
        for name in PLAIN_STAGES:
            g.stage(name, dt)
            r.stage(name, dt)
            _check(g, r, f"substep {step} after {name}")


def test_dam_break_substeps(device):
    p, st = cfg.dam_break(n_side=16)
    g, r = _pair(p, st)
    for step in range(12):
        g.update(p["dt"])
        r.update(p["dt"])
    gs, rs = _check(g, r, "after 12 substeps")
    assert len(gs["pairs"]) > 16 ** 3 * 10
    m_g, x_g = g.density_error()
    m_r, x_r = r.density_error()
    assert x_g == x_r
    assert abs(m_g - m_r) <= 1e-12 + 1e-9 * abs(m_r)


@pytest.mark.parametrize("active", [False, True])
def test_water_drop_stages(device, active):
    p, st = cfg.small_water_drop(active=active)
    g, r = _pair(p, st)
    dt = p["dt"]
    for step in range(4):
        for name in adaptive_stages(p):
            g.stage(name, dt)
            r.stage(name, dt)
            _check(g, r, f"substep {step} after {name}")


@pytest.mark.parametrize("active,steps", [(False, 30), (True, 10)])
def test_water_drop_substeps(device, active, steps):
    p, st = cfg.small_water_drop(seed=3, active=active)
    g, r = _pair(p, st)
    transfers = 0
    for step in range(steps):
        g.update(p["dt"])
        r.update(p["dt"])
        gs, _ = _check(g, r, f"substep {step}")
        transfers = max(transfers, len(gs["t_source"]))
    assert transfers > 0  # merges or splits were in flight


def test_water_drop_boundary_distance_mode(device):
    """Reference default mode: kernel width from boundary distance, range scale 1."""
    p, st = cfg.small_water_drop(seed=5, active=True)
    p["settings"]["mBaseKernelWidthOnBoundaryDistance"] = 1
    p["settings"]["mTargetRadiusScaleFactor"] = 0.3
    g, r = _pair(p, st)
    for step in range(6):
        g.update(p["dt"])
        r.update(p["dt"])
        _check(g, r, f"substep {step}")


def test_transfer_capacity_clamp(device):
    """A tiny transfer capacity forces the clamps of find_split_and_merge_3 and
    remove_impossible_splits (and the sequential claim replay)."""
    p, st = cfg.small_water_drop(seed=2, active=True)
    p["cap_transfers"] = 40
    g, r = _pair(p, st)
    for step in range(5):
        g.update(p["dt"])
        r.update(p["dt"])
        _check(g, r, f"substep {step}")


@pytest.mark.parametrize("bd_mode", [False, True])
def test_pair_capacity_clamp(device, bd_mode):
    """A pair capacity below the walk's pair count cuts the neighbour list in
    the middle of a row: the spreading walk (bd_mode False) and the plain walk
    (bd_mode True) must drop the same entries as the sequential schedule."""
    p, st = cfg.small_water_drop(seed=4, active=True)
    if bd_mode:
        p["settings"]["mBaseKernelWidthOnBoundaryDistance"] = 1
    g0 = OracleScene(p)
    g0.set_state(st)
    g0.stage("velocity", p["dt"])
    if bd_mode:
        g0.stage("kernel_width_bd", p["dt"])
    g0.stage("neighborhood", p["dt"])
    full = len(g0.get_state()["pairs"])
    p["cap_pairs"] = int(full * 0.6) + 7  # not a row boundary in general
    g, r = _pair(p, st)
    for step in range(3):
        g.update(p["dt"])
        r.update(p["dt"])
        _check(g, r, f"substep {step}")
    assert g.overflow_flags() & 1


def test_spreading_boundary_widths(device):
    """Kernel widths equal to a neighbour's rounded distance (and one float step
    either side) put pairs exactly on the cutoff of the spreading filter, and
    an exact lattice puts pairs exactly at the original kernel width, where
    the spread influence changes formula. Gravity and velocities are zero so
    the positions reach the neighbour search unchanged."""
    p, st = cfg.small_water_drop(seed=6)
    p["gravity"] = (0.0, 0.0, 0.0)
    n = len(st["radius"])
    rng = np.random.default_rng(12)
    st["velocity"][:] = 0.0
    # an exact lattice block at spacing 2 (= 2 r for the coarse particles)
    k = min(n, 512)
    ijk = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3)[:k]
    base = st["position"][:, :3].min(0)
    st["position"][:k, :3] = base + (ijk * 2 * 262144).astype(np.int32)
    st["position"][:k, 3] = 0
    st["radius"][:k] = 1.0
    st["inverse_mass"][:k] = 0.125
    st["target_radius"][:k] = 1.0
    st["pos_backup"] = st["position"].copy()
    # widths at exact rounded distances to a random other particle
    pos = st["position"]
    partner = np.clip(np.arange(n) + rng.integers(-30, 30, n), 0, n - 1)
    partner[partner == np.arange(n)] = (partner[partner == np.arange(n)] + 1) % n
    diff = (pos[partner, :3].astype(np.int64) - pos[:, :3].astype(np.int64)).astype(np.int32)
    dw = diff.astype(np.float32) / np.float32(262144.0)
    d2 = (dw[:, 0] * dw[:, 0] + dw[:, 1] * dw[:, 1]) + dw[:, 2] * dw[:, 2]
    dist = np.sqrt(d2.astype(np.float32))
    step = rng.integers(-1, 2, n)
    kw = np.where(step < 0, np.nextafter(dist, np.float32(0)),
                  np.where(step > 0, np.nextafter(dist, np.float32(np.inf)), dist)).astype(np.float32)
    use = dist > st["kernel_width"]  # only where the width, not the radius, sets the cutoff
    st["kernel_width"] = np.where(use, kw, st["kernel_width"]).astype(np.float32)
    g, r = _pair(p, st)
    for step_i in range(2):
        g.update(p["dt"])
        r.update(p["dt"])
        _check(g, r, f"substep {step_i}")


@pytest.mark.parametrize("kind", ["dam_break", "water_drop_active"])
def test_binary_search_neighborhood_matches_oracle(device, kind):
    """neighborhood_binary_search: the z-order sort, the 27-cell lower-bound
    walk and the substeps built on it, bit for bit."""
    if kind == "dam_break":
        p, st = cfg.dam_break(n_side=14)
        rng = np.random.default_rng(3)
        st["kernel_width"] = (st["kernel_width"] * rng.uniform(0.5, 1.7, len(st["radius"]))).astype(np.float32)
    else:
        p, st = cfg.small_water_drop(seed=8, active=True)
    g, r = _pair(p, st)
    g.set_neighborhood_method("binary_search")
    r.set_neighborhood_method("binary_search")
    g.stage("velocity", p["dt"])
    r.stage("velocity", p["dt"])
    g.stage("neighborhood", p["dt"])
    r.stage("neighborhood", p["dt"])
    _check(g, r, "after the neighbourhood stage")
    for step in range(4):
        g.update(p["dt"])
        r.update(p["dt"])
        _check(g, r, f"substep {step}")


@pytest.mark.parametrize("kind", ["dam_break", "water_drop_active"])
def test_brute_force_neighborhood_matches_oracle(device, kind):
    """neighborhood_brute_force as the scene's search: stages and substeps bit
    for bit (no sort, rows in index order)."""
    if kind == "dam_break":
        p, st = cfg.dam_break(n_side=12)
    else:
        p, st = cfg.small_water_drop(seed=9, active=True)
    g, r = _pair(p, st)
    g.set_neighborhood_method("brute_force")
    r.set_neighborhood_method("brute_force")
    g.stage("velocity", p["dt"])
    r.stage("velocity", p["dt"])
    g.stage("neighborhood", p["dt"])
    r.stage("neighborhood", p["dt"])
    _check(g, r, "after the neighbourhood stage")
    for step in range(3):
        g.update(p["dt"])
        r.update(p["dt"])
        _check(g, r, f"substep {step}")


@pytest.mark.parametrize("kind", ["spherical_pool", "waterfall", "pool"])
def test_reference_scenes_match_oracle(device, kind):
    """The reference's other scenes at reduced size: the spherical pool's
    sphere collider (sphere_collision.comp:42-60, no boxes), the waterfall's
    eleven boxes and the pool, stages and substeps bit for bit."""
    if kind == "spherical_pool":
        p, st = cfg.spherical_pool_scene(pool_radius=12.0)
    elif kind == "waterfall":
        p, st = cfg.waterfall_scene(mn=(-10.0, -4.0, -20.0), mx=(10.0, 6.0, -10.0))
    else:
        p, st = cfg.pool_scene(mn=(-10.0, -4.0, -20.0), mx=(10.0, 6.0, -10.0))
    g, r = _pair(p, st)
    for name in adaptive_stages(p) + (["sphere"] if kind == "spherical_pool" else ["box"]):
        g.stage(name, p["dt"])
        r.stage(name, p["dt"])
        _check(g, r, f"after {name}")
    for step in range(4):
        g.update(p["dt"])
        r.update(p["dt"])
        _check(g, r, f"substep {step}")


@pytest.mark.parametrize("seed,res", [(3, 4), (11, 5)])
def test_fused_walk_raw_count_matches_oracle(device, seed, res):
    """The single-pass walk with spreading counts the full (range 1.5 kw) list
    exactly, also where it only counts the entries of far cells: its raw total
    equals the length of the oracle's unfiltered neighbour list."""
    p, st = cfg.small_water_drop(seed=seed, res=res)
    g, r = _pair(p, st)
    for step in range(3):
        for name in ("velocity", "particle_transfer", "neighborhood"):
            r.stage(name, p["dt"])
        raw = len(r.get_state()["pairs"])
        g.update(p["dt"])
        assert g.diagnostics()["pairs_before_clamp"] == raw, f"substep {step}"
        for name in ["spread"] + ["box", "incompressibility"] * p["solver_iterations"] + ["update_transfers"]:
            r.stage(name, p["dt"])
        _check(g, r, f"substep {step}")


def test_adaptive_density_error_matches_oracle(device):
    """The density error measured inside the adaptive substep (after the solver
    loop, in the split/merge find's pass over the rows) equals the oracle's
    density error of the same state."""
    p, st = cfg.small_water_drop(seed=4, active=True)
    g, r = _pair(p, st)
    g.set_error_measurement(True)
    run_max = 0.0
    for step in range(4):
        g.update(p["dt"])
        for name in adaptive_stages(p)[:-1]:  # up to the end of the solver loop
            r.stage(name, p["dt"])
        m_r, x_r = r.density_error()
        r.stage("update_transfers", p["dt"])
        _check(g, r, f"substep {step}")
        _, x_g, last_g, _, n_g = g.error_stats()
        run_max = max(run_max, x_r)
        assert n_g == step + 1
        assert abs(last_g - m_r) <= 1e-12 + 1e-9 * abs(m_r), (step, last_g, m_r)
        assert x_g == run_max


def test_solver_long_rows_match_oracle(device):
    """Rows of thousands of entries: every particle of a 16^3 lattice reaches
    most of the others, so the neighbours of a block's rows outnumber the
    solver's table of shifts in LDS and most shifts take the direct atomics.
    The first substep still matches the oracle bit for bit; it blows the
    overlapping lattice apart, and the second one (nearly no pairs left, far
    out-of-box positions) matches as well."""
    p, st = cfg.dam_break(n_side=16)
    n = len(st["radius"])
    rad = np.full(n, 1.5, np.float32)
    st.update(radius=rad, target_radius=rad.copy(),
              inverse_mass=(np.float32(1.0) / (np.float32(2.0) * rad) ** 3).astype(np.float32),
              kernel_width=(rad * np.float32(cfg.KERNEL_SCALE)).astype(np.float32),
              boundary_distance=np.trunc(rad.astype(np.float64) * cfg.POS_RESOLUTION).astype(np.uint32))
    p["cap_pairs"] = n * n
    p["hash_res_log2"] = 1  # cells of 16 x 8 x 8 world units, wider than the 6-unit kernel
    g, r = _pair(p, st)
    for step in range(2):
        g.update(p["dt"])
        r.update(p["dt"])
        gs, _ = _check(g, r, f"substep {step}")
        if step == 0:
            assert len(gs["pairs"]) > 1500 * n, len(gs["pairs"]) / n


def test_solver_mass_quotients_match_oracle(device):
    """Inverse masses that are not powers of two (as after merges of unequal
    masses) in every neighbour row: most within a factor of two of the
    lattice's, some 2^-44 and 2^44 times it (outside the range of the
    solver's reciprocal quotient, so their rows' slots take the full IEEE
    division). The solver's quotients by them give the oracle's substeps bit
    for bit."""
    p, st = cfg.small_water_drop(seed=5, active=False)
    n = len(st["radius"])
    rng = np.random.default_rng(11)
    f = rng.uniform(0.5, 2.0, n).astype(np.float32)
    far = rng.random(n) < 0.01
    f[far] *= np.where(rng.random(int(far.sum())) < 0.5, np.float32(2.0 ** -44), np.float32(2.0 ** 44))
    im = (st["inverse_mass"] * f).astype(np.float32)
    assert ((im.view(np.uint32) & 0x7FFFFF) != 0).mean() > 0.99
    st["inverse_mass"] = im
    g, r = _pair(p, st)
    for step in range(2):
        g.update(p["dt"])
        r.update(p["dt"])
        _check(g, r, f"substep {step}")


@pytest.mark.parametrize("mode", [0, 1])
def test_solver_table_variants_match_oracle(device, mode):
    """The solver's two shift-table variants (flush every 64 or every 32 rows),
    pinned instead of chosen by row length, give the oracle's substeps."""
    p, st = cfg.small_water_drop(seed=6, active=True)
    g, r = _pair(p, st)
    g.set_long_rows(mode)
    for step in range(6):
        g.update(p["dt"])
        r.update(p["dt"])
    _check(g, r, f"variant {mode}")
