"""The learned lambda warm start on the benchmark's 1,000,137-particle water
drop at full size (its first substep from the lattice state):

* the warm-shift pass with the network's lambdas handed to the oracle, the
  plain solver's judgement of it and the following iterations give the
  oracle's state bit for bit (every list, the ordered neighbour pairs
  included) and the same iteration and fallback counts;
* the 16 features of every particle equal their float64 NumPy restatement
  (relative 1e-3 of each feature's range: the device truncates float32
  terms of the integer sums), and the MFMA network's predictions over all
  1 M feature rows equal the NumPy network of oracle/mlp_ref.py within the
  bound of its BF16 operands (2e-3 of the largest prediction)."""
import numpy as np
import pytest

from apbf_amd import config as cfg
from statecmp import compare
from warm_features import device_features, restate_features

pytestmark = pytest.mark.gpu


def _model(seed, bias, spread, norm=None):
    from apbf_amd import learned as Lm
    prm = Lm.init_params(seed) * np.float32(spread)
    prm[Lm.OFF_B3] = bias
    if norm is None:
        norm = np.concatenate([np.zeros(16), np.full(16, 0.1), [1.0]]).astype(np.float32)
    return Lm.WarmStartModel(prm, norm), prm, norm


@pytest.mark.timeout(900)
def test_warm_start_full_size_substep_matches_oracle(device):
    from apbf_amd import learned as Lm
    from apbf_amd.scene import Scene
    from oracle.oracle import OracleScene
    p, st = cfg.water_drop()
    assert len(st["radius"]) == 1_000_137
    g = Scene(p)
    g.set_state(st)
    model, _, _ = _model(1, -0.02, 0.05)
    Lm.set_mode(g, model, "warm")
    g.update(p["dt"])
    lam = g.warm_lambdas()
    assert len(lam) == 1_000_137 and np.all(np.isfinite(lam))
    assert np.count_nonzero(lam) > 900_000
    gs = g.get_state()
    stats = g.solver_stats()
    Lm.set_mode(g, None, "off")
    model.close()
    g.close()
    r = OracleScene(p)
    r.set_state(st)
    r.set_warm_lambdas(lam)
    r.update(p["dt"])
    rs = r.get_state()
    print(f"solver stats (iterations, substeps, warm starts undone): {stats}", flush=True)
    assert stats == r.solver_stats()
    assert len(gs["pairs"]) > 50_000_000
    bad = compare(gs, rs, check_pairs=True)
    assert not bad, "; ".join(bad)


@pytest.mark.timeout(900)
def test_warm_features_and_network_full_size(device):
    import torch
    from apbf_amd import learned as Lm
    from oracle import mlp_ref
    p, st = cfg.water_drop()
    F, sb = device_features(p, st, substeps_before=0)
    assert len(F) == 1_000_137
    R = restate_features(sb, p["settings"], p["dt"], F[:, 8])
    scale = np.maximum(np.abs(R).max(axis=0), 1e-12)
    err = (np.abs(F - R) / scale).max(axis=0)
    assert (err < 1e-3).all(), {k: float(e) for k, e in enumerate(err) if e >= 1e-3}
    x = F.astype(np.float32)
    rng = np.random.default_rng(4)
    norm = Lm.fit_norm(x, rng.standard_normal(len(x)).astype(np.float32) * 0.3)
    model, prm, norm = _model(4, 0.1, 1.0, norm)
    got = model.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    model.close()
    want, _ = mlp_ref.forward(prm, norm, x)
    big = np.abs(want).max() + 1e-6
    worst = np.max(np.abs(got - want)) / big
    print(f"network over {len(x)} rows: worst error {worst:.2e} of the largest prediction", flush=True)
    assert worst <= 2e-3
