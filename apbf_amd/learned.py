"""Learned lambda warm start: the 16-64-64-1 network of libapbf_hip
(include/apbf.h, apbf_mlp_*), trained on the GPU by the library's own
forward/backward kernels on (feature, total-lambda) pairs that the plain
solver produces (apbf_scene_set_warm_start mode 1), then used to warm-start
the Jacobi solve (mode 2).

The reference has no counterpart: it recomputes lambda from zero in every
iteration and discards it (incompressibility.cpp:22-25)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

FEATURES = 16
HIDDEN = 64
PARAMS = 5313
NORM = 33
OFF_W1, OFF_B1 = 0, 1024
OFF_W2, OFF_B2 = 1088, 5184
OFF_W3, OFF_B3 = 5248, 5312

MODES = {"off": 0, "harvest": 1, "warm": 2}


def _fp(a):
    return a.ctypes.data_as(C.c_void_p)


def _dp(t):
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised weights, zero biases, as the flat parameter vector."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, np.float32)
    p[OFF_W1:OFF_B1] = rng.standard_normal(HIDDEN * FEATURES) * np.sqrt(2.0 / FEATURES)
    p[OFF_W2:OFF_B2] = rng.standard_normal(HIDDEN * HIDDEN) * np.sqrt(2.0 / HIDDEN)
    p[OFF_W3:OFF_B3] = rng.standard_normal(HIDDEN) * np.sqrt(1.0 / HIDDEN)
    return p


def fit_norm(features, targets) -> np.ndarray:
    """Feature means and inverse standard deviations, and the target scale
    (NumPy arrays, or torch tensors reduced on their own device in FP64)."""
    if type(features).__module__.startswith("torch"):
        import torch
        f = features.reshape(-1, FEATURES).to(torch.float64)
        mean = f.mean(0)
        std = f.std(0, unbiased=False)
        scale = torch.where(std > 1e-12, 1.0 / std.clamp_min(1e-12), torch.ones_like(std))
        t = targets.to(torch.float64)
        out = float(torch.sqrt(torch.mean(t * t))) or 1.0
        return np.concatenate([mean.cpu().numpy(), scale.cpu().numpy(), [out]]).astype(np.float32)
    f = np.asarray(features, np.float64).reshape(-1, FEATURES)
    mean = f.mean(0)
    std = f.std(0)
    scale = np.where(std > 1e-12, 1.0 / np.maximum(std, 1e-12), 1.0)
    t = np.asarray(targets, np.float64)
    out = float(np.sqrt(np.mean(t * t))) or 1.0
    return np.concatenate([mean, scale, [out]]).astype(np.float32)


class WarmStartModel:
    """One network on the device (apbf_mlp)."""

    def __init__(self, params: np.ndarray | None = None, norm: np.ndarray | None = None, seed: int = 0):
        self._L = _lib.lib()
        h = C.c_void_p()
        check(self._L.apbf_mlp_create(C.byref(h)), "apbf_mlp_create")
        self._h = h
        self.set_params(init_params(seed) if params is None else params,
                        np.concatenate([np.zeros(FEATURES), np.ones(FEATURES), [1.0]]).astype(np.float32)
                        if norm is None else norm)

    def close(self):
        if getattr(self, "_h", None):
            self._L.apbf_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_params(self, params: np.ndarray, norm: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        n = np.ascontiguousarray(norm, np.float32)
        assert p.size == PARAMS and n.size == NORM
        check(self._L.apbf_mlp_set_params(self._h, _fp(p), _fp(n)), "apbf_mlp_set_params")

    def get_params(self):
        p = np.zeros(PARAMS, np.float32)
        n = np.zeros(NORM, np.float32)
        check(self._L.apbf_mlp_get_params(self._h, _fp(p), _fp(n)), "apbf_mlp_get_params")
        return p, n

    def set_output_damping(self, gamma: float) -> None:
        """Scale the predicted lambdas by gamma. The one-shot warm shift applies
        the total lambda of a whole solve with the gradients of its first
        iteration; gamma = 1 applies the prediction as it is, gamma < 1 damps
        it and gamma > 1 pushes the particles further apart (less compression
        error, slightly more expansion)."""
        p, n = self.get_params()
        if not hasattr(self, "_out_scale"):
            self._out_scale = float(n[32])
        n[32] = np.float32(self._out_scale * gamma)
        self.set_params(p, n)

    def forward(self, features):
        """Predictions for a (n, 16) float32 CUDA tensor."""
        import torch
        x = features.contiguous()
        n = x.shape[0]
        out = torch.empty(n, dtype=torch.float32, device=x.device)
        d_len = torch.tensor([n], dtype=torch.int32, device=x.device)
        check(self._L.apbf_mlp_forward(self._h, _dp(x), _dp(d_len), n, _dp(out), _stream()), "apbf_mlp_forward")
        torch.cuda.synchronize()
        return out

    def gradient(self, features, targets):
        """(gradient vector, loss) over all samples, without an update."""
        import torch
        n = features.shape[0]
        perm = torch.arange(n, dtype=torch.int32, device=features.device)
        g = np.zeros(PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        check(self._L.apbf_mlp_gradient(self._h, _dp(features.contiguous()), _dp(targets.contiguous()), _dp(perm), n,
                                        _fp(g), _fp(loss), _stream()), "apbf_mlp_gradient")
        return g, float(loss[0])

    def train(self, features, targets, epochs: int = 4, batch: int = 4096, lr: float = 2e-3, seed: int = 0):
        """Adam over shuffled batches; returns the per-step losses."""
        import torch
        n = features.shape[0]
        batch = min(batch, n)
        steps = max(1, n // batch)
        g = torch.Generator(device=features.device).manual_seed(seed)
        losses = []
        x = features.contiguous()
        t = targets.contiguous()
        for e in range(epochs):
            perm = torch.randperm(n, generator=g, device=features.device).to(torch.int32)
            out = np.zeros(steps, np.float32)
            check(self._L.apbf_mlp_train(self._h, _dp(x), _dp(t), _dp(perm), n, batch, steps, C.c_float(lr),
                                         _fp(out), _stream()), "apbf_mlp_train")
            losses.append(out)
        return np.concatenate(losses)


def fit_warm_start(scene, dt: float, harvest_substeps: int = 6, epochs: int = 12, batch: int = 4096,
                   lr: float = 2e-3, damping: float = 1.0, seed: int = 0):
    """Harvest (feature, total-lambda) pairs from plain substeps of `scene`,
    train a network on them on the GPU and return (model, report)."""
    import time
    import torch
    t0 = time.perf_counter()
    f, t = harvest(scene, harvest_substeps, dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    norm = fit_norm(f, t)
    model = WarmStartModel(None, norm, seed=seed)
    losses = model.train(f, t, epochs=epochs, batch=batch, lr=lr, seed=seed)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    pred = model.forward(f)
    rel = float(torch.sqrt(torch.mean((pred - t) ** 2)) / torch.sqrt(torch.mean(t * t)))
    model.set_output_damping(damping)
    set_mode(scene, None, "off")
    return model, {"samples": int(f.shape[0]), "harvest_s": t1 - t0, "train_s": t2 - t1,
                   "epochs": epochs, "final_loss": float(losses[-1]), "relative_rms_error": rel,
                   "damping": damping}


def harvest(scene, substeps: int, dt: float):
    """Runs `substeps` plain substeps with harvesting on and returns the
    concatenated (features, targets) CUDA tensors of all but the first
    (whose previous-lambda feature is still empty)."""
    import torch
    L = _lib.lib()
    check(L.apbf_scene_set_warm_start(scene._h, None, MODES["harvest"]), "set_warm_start")
    feats, targs = [], []
    for k in range(substeps):
        scene.update(dt)
        if k == 0:
            continue
        n = scene.counts()[0]
        f = torch.empty((n, FEATURES), dtype=torch.float32, device="cuda")
        t = torch.empty(n, dtype=torch.float32, device="cuda")
        cnt = C.c_uint32()
        check(L.apbf_scene_harvest(scene._h, _dp(f), _dp(t), n, C.byref(cnt)), "apbf_scene_harvest")
        feats.append(f[: cnt.value])
        targs.append(t[: cnt.value])
    return torch.cat(feats), torch.cat(targs)


def set_mode(scene, model: WarmStartModel | None, mode: str) -> None:
    L = _lib.lib()
    check(L.apbf_scene_set_warm_start(scene._h, model.handle if model is not None else None, MODES[mode]),
          "apbf_scene_set_warm_start")
