"""The solver loop's tolerance mode and the judged warm start against the
oracle, bit for bit: every substep's state and the applied-iteration and
fallback counts. The warm start's predicted lambdas come from the device
network (apbf_scene_get_warm_lambdas) and are handed to the oracle, whose
warm-shift pass, judgement and plain iterations must then give the same
state."""
import numpy as np
import pytest

from apbf_amd import config as cfg
from statecmp import compare

pytestmark = pytest.mark.gpu


def _pair(p, st):
    from apbf_amd.scene import Scene
    from oracle.oracle import OracleScene
    g = Scene(p)
    g.set_state(st)
    r = OracleScene(p)
    r.set_state(st)
    return g, r


@pytest.mark.parametrize("kind,tol", [("dam_break", 2e-3), ("water_drop_active", 3e-4), ("water_drop_active", 1e-9)])
def test_tolerance_mode_matches_the_oracle(device, kind, tol):
    p, st = cfg.dam_break(n_side=10, spacing=0.48) if kind == "dam_break" else cfg.small_water_drop(seed=8, active=True)
    g, r = _pair(p, st)
    g.set_solver_tolerance(tol, 8)
    r.set_solver_tolerance(tol, 8)
    for k in range(4):
        g.update(p["dt"])
        r.update(p["dt"])
        bad = compare(g.get_state(), r.get_state(), check_pairs=True)
        assert not bad, f"substep {k}: " + "; ".join(bad)
    gi, gs, gf = g.solver_stats()
    assert (gi, gs, gf) == r.solver_stats()
    assert gs == 4 and 0 < gi <= 32
    if tol == 1e-9:
        assert gi == 32  # never reached: every substep runs the limit


def _warm_model(seed, bias, spread):
    from apbf_amd import learned as Lm
    rng = np.random.default_rng(seed)
    prm = Lm.init_params(seed) * np.float32(spread)
    prm[Lm.OFF_B3] = bias
    norm = np.concatenate([np.zeros(16), np.full(16, 0.1), [1.0]]).astype(np.float32)
    return Lm.WarmStartModel(prm, norm)


@pytest.mark.parametrize("seed,bias,spread", [(1, -0.02, 0.05), (2, -0.5, 1.0), (3, 0.0, 4.0)])
def test_warm_start_and_its_judgement_match_the_oracle(device, seed, bias, spread):
    from apbf_amd import learned as Lm
    p, st = cfg.small_water_drop(seed=9, active=True)
    g, r = _pair(p, st)
    model = _warm_model(seed, bias, spread)
    Lm.set_mode(g, model, "warm")
    for k in range(4):
        g.update(p["dt"])
        lam = g.warm_lambdas()
        assert len(lam) and np.all(np.isfinite(lam))
        r.set_warm_lambdas(lam)
        r.update(p["dt"])
        bad = compare(g.get_state(), r.get_state(), check_pairs=True)
        assert not bad, f"substep {k}: " + "; ".join(bad)
    assert g.solver_stats() == r.solver_stats()
    Lm.set_mode(g, None, "off")
    model.close()


def test_warm_start_fallback_is_taken_and_undone(device):
    """A wild network (large random weights) pushes some particles far into
    their neighbours and raises the compression: the plain solver's first
    iteration judges the warm shift worse than no shift, the positions from
    before it come back, and every substep equals the oracle's."""
    from apbf_amd import learned as Lm
    p, st = cfg.dam_break(n_side=10, spacing=0.48)
    g, r = _pair(p, st)
    g.update(p["dt"])
    r.update(p["dt"])
    model = _warm_model(7, 0.0, 8.0)
    Lm.set_mode(g, model, "warm")
    for k in range(3):
        g.update(p["dt"])
        r.set_warm_lambdas(g.warm_lambdas())
        r.update(p["dt"])
        bad = compare(g.get_state(), r.get_state(), check_pairs=True)
        assert not bad, f"substep {k}: " + "; ".join(bad)
    gi, gs, gf = g.solver_stats()
    print(f"warm-start fallbacks {gf} of {gs} substeps, {gi} iterations applied", flush=True)
    assert (gi, gs, gf) == r.solver_stats()
    Lm.set_mode(g, None, "off")
    model.close()


@pytest.mark.parametrize("tol", [0.5, 3e-2, 1e-3, 1e-9])
def test_warm_start_with_tolerance_matches_the_oracle(device, tol):
    """Warm start and tolerance together: iteration 0 judges the warm shift,
    and when the shift helped and the error after it is within the tolerance
    no plain iteration applies anything (0.5: every substep); with 1e-9 the
    tolerance is never reached and every substep runs the limit."""
    from apbf_amd import learned as Lm
    p, st = cfg.small_water_drop(seed=9, active=True)
    g, r = _pair(p, st)
    model = _warm_model(1, -0.02, 0.05)
    Lm.set_mode(g, model, "warm")
    g.set_solver_tolerance(tol, 8)
    r.set_solver_tolerance(tol, 8)
    for k in range(4):
        g.update(p["dt"])
        r.set_warm_lambdas(g.warm_lambdas())
        r.update(p["dt"])
        bad = compare(g.get_state(), r.get_state(), check_pairs=True)
        assert not bad, f"substep {k}: " + "; ".join(bad)
    stats = g.solver_stats()
    print(f"tolerance {tol}: (iterations, substeps, warm starts undone) = {stats}", flush=True)
    assert stats == r.solver_stats()
    if tol == 0.5 and stats[2] == 0:  # no shift undone: the warm shift alone met the tolerance
        assert stats[0] == 0
    Lm.set_mode(g, None, "off")
    model.close()
