"""Learned lambda warm start: the MFMA network against its NumPy restatement
(forward and gradients), training, and the solver gain on a scene."""
import numpy as np
import pytest

from oracle import mlp_ref

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
    import ctypes as C
    import torch
    from apbf_amd import _lib
    from apbf_amd import config as cfg
    from apbf_amd import learned as Lm
    from apbf_amd.scene import Scene
    p, st = cfg.small_water_drop(seed=2, active=True)
    dt, s = p["dt"], p["settings"]
    a, b = Scene(p), Scene(p)
    for sc in (a, b):
        sc.set_state(st)
        sc.update(dt)
    L = _lib.lib()
    _lib.check(L.apbf_scene_set_warm_start(a._h, None, Lm.MODES["harvest"]), "set_warm_start")
    a.update(dt)
    stages = ["velocity", "particle_transfer"] + (["kernel_width_bd"] if s["mBaseKernelWidthOnBoundaryDistance"]
                                                    else []) + ["neighborhood"]
    stages += ([] if s["mBaseKernelWidthOnBoundaryDistance"] else ["spread"]) + ["box"]
    for name in stages:
        b.stage(name, dt)
    # the substep's splits append particles after the solver: the features
    # cover the particles the solver started with, the first rows of the list
    na = a.counts()[0]
    f = torch.empty((na, 16), dtype=torch.float32, device="cuda")
    t = torch.empty(na, dtype=torch.float32, device="cuda")
    cnt = C.c_uint32()
    _lib.check(L.apbf_scene_harvest(a._h, Lm._dp(f), Lm._dp(t), na, C.byref(cnt)), "apbf_scene_harvest")
    sb = b.get_state()
    n = len(sb["radius"])
    assert 0 < n <= cnt.value
    F = f[:n].cpu().numpy().astype(np.float64)
    res = 2.0 ** 18
    pos = sb["position"][:, :3].astype(np.int64)
    im = sb["inverse_mass"].astype(np.float64)
    r = sb["radius"].astype(np.float64)
    kw = sb["kernel_width"].astype(np.float64)
    ivdwp = (0.6 / (kw / 2.0) ** 3) ** (2.0 / 3.0)
    ivd, norm = ivdwp * np.pi, ivdwp ** 1.5
    ird = (2.0 * r) ** 3 * im
    i, j = sb["pairs"][:, 0].astype(np.int64), sb["pairs"][:, 1].astype(np.int64)
    d = (pos[j] - pos[i]) / res
    dd = np.einsum("ij,ij->i", d, d)
    w = np.exp(-dd * ivd[i]) * norm[i]
    grad = np.where((np.sqrt(dd) >= 1e-4)[:, None], -2.0 * (w * ivd[i])[:, None] * d, 0.0)
    q = np.trunc
    dens = np.bincount(i, q(w / im[j] * res), n) + q(norm / im * res)
    sv = np.stack([np.bincount(i, q(grad[:, k] / im[j] * res), n) for k in range(3)], 1) / res
    sq = np.bincount(i, q(np.einsum("ij,ij->i", grad, grad) / im[j] * res), n) / res
    c = np.bincount(i, None, n).astype(np.float64)
    densf = dens / res
    wgss = np.einsum("ij,ij->i", sv, sv)
    sgs = sq + wgss * im
    under = 1.0 - densf * ird
    lam = under / (ird * ird * (sgs + 0.01)) / ((2.0 * s["mSmallestTargetRadius"]) ** 3 / ird * im)
    rsum = np.bincount(i, r[j], n)
    dsum = np.bincount(i, np.sqrt(dd), n)
    v = sb["velocity"][:, :3].astype(np.float64)
    R = np.zeros((n, 16))
    R[:, 0] = c / 64.0
    R[:, 1] = densf * ird
    R[:, 2] = under
    R[:, 3] = np.sqrt(wgss) * ird
    R[:, 4] = sq * ird * ird
    R[:, 5] = ird * ird * (sgs + 0.01)
    R[:, 6] = lam
    R[:, 7] = np.minimum(lam, 0.0)
    R[:, 8] = F[:, 8]
    R[:, 9] = np.where(c > 0, r * c / np.where(rsum > 0, rsum, 1.0), 1.0)
    R[:, 10] = kw / r
    R[:, 11] = sb["boundariness"]
    R[:, 12] = ird
    R[:, 13] = np.log2(r)
    R[:, 14] = np.where(c > 0, dsum / np.where(c > 0, c * kw, 1.0), 0.0)
    R[:, 15] = np.sqrt(np.einsum("ij,ij->i", v, v)) * dt / r
    scale = np.maximum(np.abs(R).max(axis=0), 1e-12)
    err = (np.abs(F - R) / scale).max(axis=0)
    assert (err < 1e-3).all(), {k: float(e) for k, e in enumerate(err) if e >= 1e-3}
