"""Comparison helpers for scene states (GPU vs oracle)."""
from __future__ import annotations

import numpy as np

PARTICLE_KEYS = ("position", "velocity", "inverse_mass", "radius", "pos_backup", "transferring",
                 "target_radius", "kernel_width", "boundariness", "boundary_distance")


def bits(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a)
    if a.dtype == np.float32:
        return a.view(np.uint32)
    return a


def first_diff(a, b):
    a, b = bits(a), bits(b)
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    if a.size == 0:
        return None
    d = np.nonzero((a != b).reshape(len(a), -1).any(axis=1))[0] if a.ndim > 1 else np.nonzero(a != b)[0]
    if len(d) == 0:
        return None
    i = int(d[0])
    return f"{len(d)} rows differ, first at {i}: {a[i]!r} vs {b[i]!r}"


def transfers_set(st):
    return sorted(zip(st["t_source"].tolist(), st["t_target"].tolist(),
                      bits(st["t_time_left"]).tolist()))


def compare(g: dict, r: dict, check_pairs: bool = True) -> list[str]:
    """Bit-exact comparison; returns a list of mismatch descriptions."""
    out = []
    assert np.array_equal(r["fluid_index"], np.arange(len(r["fluid_index"]))), "oracle fluid index not identity"
    for k in PARTICLE_KEYS:
        d = first_diff(g[k], r[k])
        if d:
            out.append(f"{k}: {d}")
    if transfers_set(g) != transfers_set(r):
        out.append(f"transfers differ ({len(g['t_source'])} vs {len(r['t_source'])})")
    if check_pairs:
        d = first_diff(g["pairs"], r["pairs"])
        if d:
            out.append(f"pairs: {d}")
    return out
