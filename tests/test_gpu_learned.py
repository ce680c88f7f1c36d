"""Learned lambda warm start: the MFMA network against its NumPy restatement
(forward and gradients), training, and the solver gain on a scene."""
import numpy as np
import pytest

from oracle import mlp_ref
from warm_features import device_features, restate_features

pytestmark = pytest.mark.gpu


def _model_and_data(n=1000, seed=0):
    import torch
    from apbf_amd import learned as Lm
    rng = np.random.default_rng(seed)
    p = Lm.init_params(seed)
    p[Lm.OFF_B1:Lm.OFF_W2] = rng.standard_normal(64) * 0.1
    p[Lm.OFF_B2:Lm.OFF_W3] = rng.standard_normal(64) * 0.1
    p[Lm.OFF_B3] = 0.3
    x = (rng.standard_normal((n, 16)) * rng.uniform(0.1, 10, 16) + rng.uniform(-5, 5, 16)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.1 * x[:, 3] - 0.02 * x[:, 7] ** 2).astype(np.float32)
    norm = Lm.fit_norm(x, t)
    m = Lm.WarmStartModel(p, norm)
    return m, p, norm, x, t, torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()


@pytest.mark.parametrize("n", [1, 16, 1000, 4099])
def test_mlp_forward_matches_reference(device, n):
    m, p, norm, x, t, xd, td = _model_and_data(n=n, seed=n)
    got = m.forward(xd).cpu().numpy()
    want, _ = mlp_ref.forward(p, norm, x)
    scale = np.abs(want).max() + 1e-6
    assert np.max(np.abs(got - want)) <= 2e-3 * scale, np.max(np.abs(got - want)) / scale


def test_mlp_gradient_matches_reference(device):
    m, p, norm, x, t, xd, td = _model_and_data(n=2000, seed=3)
    g, loss = m.gradient(xd, td)
    gr, lr = mlp_ref.gradient(p, norm, x, t)
    assert abs(loss - lr) <= 1e-3 * abs(lr)
    cos = float(g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr)))
    assert cos > 0.9999, cos
    # every block of the parameter vector, relative to its own magnitude
    from apbf_amd import learned as Lm
    for a, b in [(Lm.OFF_W1, Lm.OFF_B1), (Lm.OFF_B1, Lm.OFF_W2), (Lm.OFF_W2, Lm.OFF_B2), (Lm.OFF_B2, Lm.OFF_W3),
                 (Lm.OFF_W3, Lm.OFF_B3), (Lm.OFF_B3, Lm.PARAMS)]:
        err = np.linalg.norm(g[a:b] - gr[a:b]) / (np.linalg.norm(gr[a:b]) + 1e-12)
        assert err < 2e-2, (a, b, err)


def test_mlp_training_reduces_loss(device):
    m, p, norm, x, t, xd, td = _model_and_data(n=20000, seed=5)
    _, l0 = m.gradient(xd, td)
    losses = m.train(xd, td, epochs=6, batch=1024, lr=3e-3)
    _, l1 = m.gradient(xd, td)
    assert np.all(np.isfinite(losses))
    assert l1 < 0.2 * l0, (l0, l1)


def _train_call(m, xd, td, perm, batch, steps, lr):
    import ctypes as C
    from apbf_amd import learned as Lm
    from apbf_amd._lib import check
    out = np.zeros(max(steps, 1), np.float32)
    check(m._L.apbf_mlp_train(m.handle, Lm._dp(xd), Lm._dp(td), Lm._dp(perm), xd.shape[0], batch, steps,
                              C.c_float(lr), Lm._fp(out), Lm._stream()), "apbf_mlp_train")
    return out


def test_captured_training_equals_step_by_step(device):
    """A call of 24 steps runs as one captured graph (replayed on the second
    epoch); the same steps as calls of 8 run launch by launch. The sample
    order of each short call is the long call's rotated by its first step,
    and the device step count carries the Adam corrections across calls, so
    parameters and losses must agree bit for bit."""
    import torch
    n, batch, steps, chunk = 24 * 512 + 100, 512, 24, 8
    ma, p, norm, x, t, xd, td = _model_and_data(n=n, seed=11)
    mb = type(ma)(p, norm)
    g = torch.Generator(device="cpu").manual_seed(4)
    la, lb = [], []
    for epoch in range(2):
        perm = torch.randperm(n, generator=g).to(torch.int32)
        la.append(_train_call(ma, xd, td, perm.cuda(), batch, steps, 3e-3))
        for c in range(steps // chunk):
            rot = torch.roll(perm, -c * chunk * batch).cuda()
            lb.append(_train_call(mb, xd, td, rot, batch, chunk, 3e-3))
    pa, _ = ma.get_params()
    pb, _ = mb.get_params()
    assert np.array_equal(np.concatenate(la), np.concatenate(lb))
    assert np.array_equal(pa, pb)
    assert not np.array_equal(pa, p)


def test_warm_start_gain_on_a_scene(device):
    """Trained on pairs from plain substeps of a small water drop, the warm
    start followed by 2 Jacobi iterations ends each substep with a lower mean
    density error than the plain solve with 4 iterations from the same state:
    at least 2x fewer plain iterations for the same error."""
    import torch
    from apbf_amd import config as cfg
    from apbf_amd import learned as Lm
    from apbf_amd.scene import Scene
    p, st = cfg.small_water_drop(seed=1)
    dt = p["dt"]
    sc = Scene(p)
    sc.set_state(st)
    for _ in range(10):
        sc.update(dt)
    model, rep = Lm.fit_warm_start(sc, dt, harvest_substeps=12, epochs=8, damping=1.0)
    assert rep["relative_rms_error"] < 0.5
    s0 = sc.get_state(pairs=False)
    s0["fluid_index"] = s0["fluid_index"][: len(s0["radius"])]

    def run(mode, iters):
        q = dict(p)
        q["solver_iterations"] = iters
        s = Scene(q)
        s.set_state(s0)
        s.set_error_measurement(True)
        if mode == "warm":
            Lm.set_mode(s, model, "warm")
        for _ in range(2):
            s.update(dt)
        s.error_stats(reset=True)
        for _ in range(16):
            s.update(dt)
        torch.cuda.synchronize()
        return s.error_stats()[0]

    plain4 = run("off", 4)
    plain2 = run("off", 2)
    warm2 = run("warm", 2)
    assert warm2 < plain4 < plain2, (warm2, plain4, plain2)


def test_features_match_numpy_restatement(device):
    """The warm-start features of a substep (k_features_rows) against a NumPy
    restatement in float64, from the state the solver starts from: the same
    scene stepped through the stage API up to the first collision pass. The
    previous-lambda feature, internal to the scene, is left out. The integer
    sums truncate every pair's term, so float64 and the device's float32 may
    differ by a unit in the last place of a term: the bound is relative 1e-3."""
    from apbf_amd import config as cfg
    p, st = cfg.small_water_drop(seed=2, active=True)
    F, sb = device_features(p, st, substeps_before=1)
    R = restate_features(sb, p["settings"], p["dt"], F[:, 8])
    scale = np.maximum(np.abs(R).max(axis=0), 1e-12)
    err = (np.abs(F - R) / scale).max(axis=0)
    assert (err < 1e-3).all(), {k: float(e) for k, e in enumerate(err) if e >= 1e-3}
