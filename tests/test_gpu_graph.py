"""HIP-graph substeps, batches of scenes and launch shapes: every path must
give the state of plain eager updates bit for bit."""
import numpy as np
import pytest

from apbf_amd import config as cfg
from statecmp import compare

pytestmark = pytest.mark.gpu


def _scene(p, st):
    from apbf_amd.scene import Scene
    s = Scene(p)
    s.set_state(st)
    return s


def _same(a, b, where):
    bad = compare(a.get_state(), b.get_state(), check_pairs=True)
    assert not bad, f"{where}: " + "; ".join(bad)


@pytest.mark.parametrize("kind", ["dam_break", "water_drop_active"])
def test_scene_graph_replay_matches_eager(device, kind):
    if kind == "dam_break":
        p, st = cfg.dam_break(n_side=12)
    else:
        p, st = cfg.small_water_drop(seed=2, active=True)
    dt = p["dt"]
    eager, graphed = _scene(p, st), _scene(p, st)
    eager.update(dt)  # the graph is captured with last_dt == dt
    graphed.update(dt)
    graphed.capture(dt)
    for r in range(4):  # eager work and synchronisation between the replays
        eager.update(dt)
        eager.update(dt)
        graphed.replay(1)
        _same(eager, graphed, f"replay {r}")


def test_batch_graph_matches_individual_scenes(device):
    from apbf_amd.scene import Batch
    cases = [cfg.dam_break(n_side=10), cfg.small_water_drop(seed=1), cfg.small_water_drop(seed=3, active=True),
             cfg.dam_break(n_side=7, seed=2)]
    singles = [_scene(p, st) for p, st in cases]
    batched = [_scene(p, st) for p, st in cases]
    dt = cases[0][0]["dt"]
    for s in singles + batched:  # queued on the null stream, not waited for:
        for _ in range(3):       # the batch's streams must start after this work
            s.update(dt)
    batch = Batch(batched)
    batch.update(dt)  # eager batch substep
    for s in singles:
        s.update(dt)
    batch.synchronize()
    for k, (a, b) in enumerate(zip(singles, batched)):
        _same(a, b, f"eager batch, scene {k}")
    batch.capture(dt)
    for r in range(3):
        batch.replay(1)
        batch.synchronize()
        for s in singles:
            s.update(dt)
            s.update(dt)
        for k, (a, b) in enumerate(zip(singles, batched)):
            _same(a, b, f"batch replay {r}, scene {k}")
    batch.close()


@pytest.mark.parametrize("walk,solve", [(4, 8), (4, 16), (8, 8), (8, 32), (16, 8), (16, 16), (64, 16), (64, 32)])
def test_launch_shapes_match_the_oracle(device, walk, solve):
    from oracle.oracle import OracleScene
    p, st = cfg.small_water_drop(seed=4, active=True)
    g = _scene(p, st)
    g.set_launch_shape(walk, solve)
    r = OracleScene(p)
    r.set_state(st)
    for k in range(4):
        g.update(p["dt"])
        r.update(p["dt"])
    bad = compare(g.get_state(), r.get_state(), check_pairs=True)
    assert not bad, "; ".join(bad)


@pytest.mark.parametrize("kind", ["small_water_drop", "large_keys"])
def test_sort_shapes_match_the_oracle(device, kind):
    """Every digit width (4, 8, 11 bits) and tile (4, 8, 16 values per
    thread) of the hash sort gives the oracle's substeps bit for bit: once on
    the small water drop (hash keys of 12 bits at 2^4 cells per axis: 3, 2
    and 2 passes) and once with 2^7 cells per axis (21-bit keys: 6, 3 and 2
    passes, the last 11-bit pass partial)."""
    from oracle.oracle import OracleScene
    p, st = cfg.small_water_drop(seed=6, active=True)
    if kind == "large_keys":
        p = dict(p, hash_res_log2=7)
    r = OracleScene(p)
    r.set_state(st)
    for _ in range(3):
        r.update(p["dt"])
    want = r.get_state()
    for bits in (4, 8, 11):
        for items in (4, 8, 16):
            g = _scene(p, st)
            g.set_sort_shape(bits, items)
            assert g.sort_shape() == (bits, items)
            for _ in range(3):
                g.update(p["dt"])
            bad = compare(g.get_state(), want, check_pairs=True)
            g.close()
            assert not bad, f"sort shape {bits} bits x {items}: " + "; ".join(bad)


def test_autotune_keeps_results(device):
    """The autotuner times every shape from the same saved state and puts the
    scene and its statistics back: afterwards the scene equals one that was
    never tuned (the neighbour list aside, which the next substep rebuilds)."""
    p, st = cfg.small_water_drop(seed=5)
    dt = p["dt"]
    tuned, plain = _scene(p, st), _scene(p, st)
    for s in (tuned, plain):
        s.set_error_measurement(True)
        for _ in range(3):
            s.update(dt)
    w, s, ms = tuned.autotune(dt, substeps=2)
    assert w in (4, 8, 64) and s in (8, 16, 32) and ms > 0
    bits, items = tuned.sort_shape()
    assert bits in (4, 8, 11) and items in (4, 8, 16)
    bad = compare(tuned.get_state(), plain.get_state(), check_pairs=False)
    assert not bad, "after autotune: " + "; ".join(bad)
    assert tuned.counts()[2] == 0
    assert tuned.error_stats() == plain.error_stats()
    assert tuned.particle_substeps() == plain.particle_substeps()
    # per-iteration errors are double sums in atomic order: equal up to rounding
    np.testing.assert_allclose(tuned.iteration_errors(), plain.iteration_errors(), rtol=1e-12, atol=0)
    for k in range(2):
        tuned.update(dt)
        plain.update(dt)
        _same(tuned, plain, f"tuned substep {k}")


def test_walk_candidate_counter(device):
    """The candidate counter (a diagnostic counting walk) changes no result,
    counts at least every in-range pair of the walk, and never more than the
    particles of the whole scene per particle."""
    p, st = cfg.small_water_drop(seed=2, active=True)
    dt = p["dt"]
    a, b = _scene(p, st), _scene(p, st)
    a.count_walk_candidates(True)
    a.update(dt)
    b.update(dt)
    d = a.diagnostics()
    n = d["rows"]
    assert d["pairs_before_clamp"] <= d["walk_candidates"] <= n * n
    a.update(dt)
    b.update(dt)
    _same(a, b, "with and without candidate counting")
    a.count_walk_candidates(False)
    assert a.diagnostics()["walk_candidates"] == 0
    a.close()
    b.close()


@pytest.mark.parametrize("walk", [4, 8, 64])
def test_walk_shapes_on_wide_and_stray_particles(device, walk):
    """Every walk shape against the oracle where the search box leaves the
    grid: particles with kernel widths far beyond the domain (the walk is cut
    to 4 * 2^res cells per axis, and cells past the grid end alias onto its
    cells, so a neighbour can be listed twice), particles outside the hash
    domain (clamped into the border cells) and a zero width; plain PBF with
    the box colliders off, so only the neighbour search differs between
    shapes."""
    from oracle.oracle import OracleScene
    p, st = cfg.dam_break(n_side=10, res=3)
    p["boxes"] = []
    st = dict(st)
    rng = np.random.default_rng(7)
    n = len(st["radius"])
    p["cap_pairs"] = 400 * n  # no clamp: every duplicate listed
    kw = st["kernel_width"].copy()
    kw[rng.choice(n, 6, replace=False)] = 12.0   # boxes wider than the grid
    kw[rng.choice(n, 5, replace=False)] = 0.0
    st["kernel_width"] = kw
    pos = st["position"].copy()
    stray = rng.choice(n, 20, replace=False)
    pos[stray, :3] += (rng.uniform(-30, 30, (20, 3)) * 262144).astype(np.int32)
    st["position"] = pos
    st["pos_backup"] = pos.copy()
    g = _scene(p, st)
    g.set_launch_shape(walk, 16)
    r = OracleScene(p)
    r.set_state(st)
    for name in ("velocity", "neighborhood"):
        g.stage(name, p["dt"])
        r.stage(name, p["dt"])
    gs, rs = g.get_state(), r.get_state()
    bad = compare(gs, rs, check_pairs=True)
    assert not bad, "; ".join(bad)
    # an aliased cell was visited twice: some row lists a neighbour twice
    pairs = gs["pairs"]
    assert len(np.unique(pairs, axis=0)) < len(pairs)
    for step in range(2):
        g.update(p["dt"])
        r.update(p["dt"])
        bad = compare(g.get_state(), r.get_state(), check_pairs=True)
        assert not bad, f"substep {step}: " + "; ".join(bad)
