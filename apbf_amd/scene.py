"""Python front end of a GPU scene (apbf_scene_* in include/apbf.h).

A :class:`Scene` holds one simulation in HBM; ``update(dt)`` runs one substep
(the reference's ``pool::update`` / ``waterdrop::update``) entirely on the
device. State upload/download is for setup, tests and inspection only.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

STAGES = {
    "velocity": 0, "particle_transfer": 1, "kernel_width_bd": 2, "neighborhood": 3,
    "spread": 4, "box": 5, "incompressibility": 6, "update_transfers": 7, "sphere": 8,
}


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else None


class Scene:
    def __init__(self, params: dict, stream: int | None = None):
        self._L = _lib.lib()
        self.params = params
        self._p = _lib.make_params(params)
        h = C.c_void_p()
        check(self._L.apbf_scene_create(C.byref(self._p), C.byref(h)), "apbf_scene_create")
        self._h = h
        if stream is not None:
            check(self._L.apbf_scene_set_stream(self._h, C.c_void_p(stream)), "apbf_scene_set_stream")
        if params.get("sphere"):  # (center, radius, hollow) of the scene's sphere collider
            c, r, hollow = params["sphere"]
            self.set_sphere(c, r, hollow)

    def close(self):
        if getattr(self, "_h", None):
            self._L.apbf_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream: int) -> None:
        check(self._L.apbf_scene_set_stream(self._h, C.c_void_p(stream)), "apbf_scene_set_stream")

    def set_state(self, st: dict) -> None:
        n = len(st["radius"])
        nt = len(st["t_source"])
        a = dict(
            position=np.ascontiguousarray(st["position"], np.int32).reshape(n, 4),
            velocity=np.ascontiguousarray(st["velocity"], np.float32).reshape(n, 4),
            inverse_mass=np.ascontiguousarray(st["inverse_mass"], np.float32),
            radius=np.ascontiguousarray(st["radius"], np.float32),
            pos_backup=np.ascontiguousarray(st["pos_backup"], np.int32).reshape(n, 4),
            transferring=np.ascontiguousarray(st["transferring"], np.uint32),
            fluid_index=np.ascontiguousarray(st["fluid_index"], np.uint32),
            target_radius=np.ascontiguousarray(st["target_radius"], np.float32),
            kernel_width=np.ascontiguousarray(st["kernel_width"], np.float32),
            boundariness=np.ascontiguousarray(st["boundariness"], np.float32),
            boundary_distance=np.ascontiguousarray(st["boundary_distance"], np.uint32),
            t_source=np.ascontiguousarray(st["t_source"], np.uint32),
            t_target=np.ascontiguousarray(st["t_target"], np.uint32),
            t_time_left=np.ascontiguousarray(st["t_time_left"], np.float32))
        if len(a["fluid_index"]) != n:
            raise ValueError("every particle must be a fluid particle exactly once")
        check(self._L.apbf_scene_set_state(
            self._h, n, _ptr(a["position"]), _ptr(a["velocity"]), _ptr(a["inverse_mass"]),
            _ptr(a["radius"]), _ptr(a["pos_backup"]), _ptr(a["transferring"]), _ptr(a["fluid_index"]),
            _ptr(a["target_radius"]), _ptr(a["kernel_width"]), _ptr(a["boundariness"]),
            _ptr(a["boundary_distance"]), nt, _ptr(a["t_source"]), _ptr(a["t_target"]),
            _ptr(a["t_time_left"]), float(st.get("last_delta_time", 1.0))), "apbf_scene_set_state")

    def counts(self):
        n, nt, npairs = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(self._L.apbf_scene_get_counts(self._h, C.byref(n), C.byref(nt), C.byref(npairs)),
              "apbf_scene_get_counts")
        return n.value, nt.value, npairs.value

    def get_state(self, pairs: bool = True) -> dict:
        n, nt, npairs = self.counts()
        st = dict(
            position=np.zeros((n, 4), np.int32), velocity=np.zeros((n, 4), np.float32),
            inverse_mass=np.zeros(n, np.float32), radius=np.zeros(n, np.float32),
            pos_backup=np.zeros((n, 4), np.int32), transferring=np.zeros(n, np.uint32),
            fluid_index=np.arange(n, dtype=np.uint32), target_radius=np.zeros(n, np.float32),
            kernel_width=np.zeros(n, np.float32), boundariness=np.zeros(n, np.float32),
            boundary_distance=np.zeros(n, np.uint32), t_source=np.zeros(nt, np.uint32),
            t_target=np.zeros(nt, np.uint32), t_time_left=np.zeros(nt, np.float32),
            pairs=np.zeros((npairs if pairs else 0, 2), np.uint32))
        ldt = C.c_float()
        check(self._L.apbf_scene_get_state(
            self._h, _ptr(st["position"]), _ptr(st["velocity"]), _ptr(st["inverse_mass"]),
            _ptr(st["radius"]), _ptr(st["pos_backup"]), _ptr(st["transferring"]),
            _ptr(st["target_radius"]), _ptr(st["kernel_width"]), _ptr(st["boundariness"]),
            _ptr(st["boundary_distance"]), _ptr(st["t_source"]), _ptr(st["t_target"]),
            _ptr(st["t_time_left"]), _ptr(st["pairs"]) if pairs else None, C.byref(ldt)),
            "apbf_scene_get_state")
        st["last_delta_time"] = ldt.value
        return st

    def lambdas(self) -> np.ndarray:
        n = self.counts()[0]
        out = np.zeros(max(n, 1), np.float32)
        check(self._L.apbf_scene_get_lambda(self._h, _ptr(out)), "apbf_scene_get_lambda")
        return out[:n]

    def stage(self, name: str, dt: float = 0.0) -> None:
        check(self._L.apbf_scene_stage(self._h, STAGES[name], float(dt)), f"stage {name}")

    def update(self, dt: float) -> None:
        check(self._L.apbf_scene_update(self._h, float(dt)), "apbf_scene_update")

    def density_error(self):
        m, x = C.c_double(), C.c_double()
        check(self._L.apbf_scene_density_error(self._h, C.byref(m), C.byref(x)), "density_error")
        return m.value, x.value

    def set_error_measurement(self, enable: bool) -> None:
        check(self._L.apbf_scene_set_error_measurement(self._h, int(enable)), "set_error_measurement")

    def error_stats(self, reset: bool = False):
        """(mean over substeps of the mean density error, max error, last mean,
        mean over substeps of the mean absolute deviation, substeps)."""
        m, x, l, a = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        n = C.c_uint32()
        check(self._L.apbf_scene_error_stats(self._h, C.byref(m), C.byref(x), C.byref(l), C.byref(a), C.byref(n),
                                             int(reset)), "error_stats")
        return m.value, x.value, l.value, a.value, n.value

    def save_particle_info(self, folder: str = "particle_data") -> None:
        """save_particle_info::apply: the reference's text dumps of the fluid."""
        check(self._L.apbf_scene_save_particle_info(self._h, folder.encode()), "save_particle_info")

    def save_svg(self, folder: str, svg_id: int, view_min, view_max, render_scale: float = 0.7,
                 max_expected_boundary_distance: float = 100.0, color: int = 0, render_boxes: bool = True) -> None:
        """save_particle_info::save_as_svg's drawing: particles_<svg_id>.svg of
        a planar scene (nothing in 3D). color: settings::color (0 boundariness,
        1 boundary-distance gradient, 2 transfers, 3 with kernel circles)."""
        vmin = (C.c_float * 2)(*view_min)
        vmax = (C.c_float * 2)(*view_max)
        check(self._L.apbf_scene_save_svg(self._h, folder.encode(), int(svg_id), vmin, vmax, float(render_scale),
                                          float(max_expected_boundary_distance), int(color), int(bool(render_boxes))),
              "save_svg")

    def save_measurements(self, folder: str, svg_id: int, step_ms: float = 0.0, neighborhood_ms: float = 0.0,
                          solver_ms: float = 0.0) -> None:
        """save_particle_info::save_additional_info: measurements_<svg_id>.txt."""
        check(self._L.apbf_scene_save_measurements(self._h, folder.encode(), int(svg_id), float(step_ms),
                                                   float(neighborhood_ms), float(solver_ms)), "save_measurements")

    def set_sphere(self, center, radius: float, hollow: bool = True, enable: bool = True) -> None:
        """The scene's sphere collider (sphere_collision::set_sphere)."""
        c = np.asarray(center, np.float32)
        check(self._L.apbf_scene_set_sphere(self._h, _ptr(c), float(radius), int(hollow), int(enable)),
              "set_sphere")

    def set_neighborhood_method(self, method: str) -> None:
        """'green' (hash grid), 'binary_search' (z-order codes) or 'brute_force'."""
        code = {"green": 0, "binary_search": 1, "brute_force": 2}[method]
        check(self._L.apbf_scene_set_neighborhood_method(self._h, code), "set_neighborhood_method")

    def capture(self, dt: float) -> None:
        """Capture two substeps into a HIP graph (replay advances by two)."""
        check(self._L.apbf_scene_capture(self._h, float(dt)), "capture")

    def replay(self, times: int = 1) -> None:
        check(self._L.apbf_scene_replay(self._h, int(times)), "replay")

    def set_launch_shape(self, walk_lanes: int, solve_lanes: int) -> None:
        check(self._L.apbf_scene_set_launch_shape(self._h, int(walk_lanes), int(solve_lanes)), "set_launch_shape")

    def set_sort_shape(self, digit_bits: int, items_per_thread: int) -> None:
        check(self._L.apbf_scene_set_sort_shape(self._h, int(digit_bits), int(items_per_thread)), "set_sort_shape")

    def sort_shape(self):
        b, i = C.c_int32(), C.c_int32()
        check(self._L.apbf_scene_sort_shape(self._h, C.byref(b), C.byref(i)), "sort_shape")
        return b.value, i.value

    def set_long_rows(self, mode: int) -> None:
        """The solver's shift-table variant: -1 by row length (default), 0 short rows, 1 long rows."""
        check(self._L.apbf_scene_set_long_rows(self._h, int(mode)), "set_long_rows")

    def autotune(self, dt: float, substeps: int = 3):
        """Fastest (walk lanes, solver lanes, ms per substep), every shape timed
        from the same saved state; the scene is put back afterwards (its
        neighbour list stays empty until the next substep)."""
        w, s, ms = C.c_int32(), C.c_int32(), C.c_float()
        check(self._L.apbf_scene_autotune(self._h, float(dt), int(substeps), C.byref(w), C.byref(s), C.byref(ms)),
              "autotune")
        return w.value, s.value, ms.value

    def iteration_errors(self, max_iters: int = 16, reset: bool = False) -> np.ndarray:
        """Mean density error before each solver iteration, averaged over the
        measured substeps."""
        out = np.zeros(max_iters, np.float64)
        check(self._L.apbf_scene_iteration_errors(self._h, _ptr(out), max_iters, int(reset)), "iteration_errors")
        return out

    def set_solver_tolerance(self, tolerance: float, max_iterations: int = 16) -> None:
        """Iterate until the mean density error is <= tolerance (at most
        max_iterations per substep); tolerance <= 0: the fixed count."""
        check(self._L.apbf_scene_set_solver_tolerance(self._h, float(tolerance), int(max_iterations)),
              "apbf_scene_set_solver_tolerance")

    def warm_lambdas(self) -> np.ndarray:
        """The network's lambdas of the last warm-started substep (per particle
        of that substep's solver)."""
        rows = self.diagnostics()["rows"]
        out = np.zeros(max(rows, 1), np.float32)
        check(self._L.apbf_scene_get_warm_lambdas(self._h, _ptr(out)), "apbf_scene_get_warm_lambdas")
        return out[:rows]

    def solver_stats(self, reset: bool = False):
        """(iterations applied, substeps, warm starts undone) summed over substeps."""
        v = [C.c_uint64() for _ in range(3)]
        check(self._L.apbf_scene_solver_stats(self._h, *[C.byref(x) for x in v], int(reset)), "apbf_scene_solver_stats")
        return tuple(x.value for x in v)

    def substep_totals(self, reset: bool = False) -> dict:
        """Sums over substeps since the last reset: particles the solver ran on,
        its pairs, the walk's pairs before the spreading filter, particles
        created by splits and deleted by finished merges."""
        out = np.zeros(5, np.uint64)
        check(self._L.apbf_scene_substep_totals(self._h, _ptr(out), 5, int(reset)), "apbf_scene_substep_totals")
        return dict(zip(("particles", "pairs", "raw_pairs", "created", "deleted"), (int(x) for x in out)))

    def time_kernels(self, names, max_launches: int = 4096) -> None:
        mask = 0
        for n in names:
            mask |= 1 << self.KERNELS[n]
        check(self._L.apbf_scene_time_kernels(self._h, mask, max_launches), "time_kernels")

    def kernel_time_of(self, name: str):
        """(total milliseconds, launches) of one timed kernel, without restarting."""
        t = C.c_double()
        n = C.c_uint32()
        check(self._L.apbf_scene_kernel_time_of(self._h, self.KERNELS[name], C.byref(t), C.byref(n)),
              "kernel_time_of")
        return t.value, n.value

    def particle_substeps(self, reset: bool = False):
        """(sum of particle counts, sum of pair counts) over substeps since the last reset."""
        v, p = C.c_uint64(), C.c_uint64()
        check(self._L.apbf_scene_particle_substeps(self._h, C.byref(v), C.byref(p), int(reset)),
              "particle_substeps")
        return v.value, p.value

    KERNELS = {"solve_rows": 0, "solve_apply": 1, "walk_count": 2, "walk_fill": 3, "spread": 4, "fsm12": 5,
               "permute": 6}

    def time_kernel(self, name: str | None, max_launches: int = 4096) -> None:
        kid = -1 if name is None else self.KERNELS[name]
        check(self._L.apbf_scene_time_kernel(self._h, kid, max_launches), "time_kernel")

    def kernel_time(self):
        """(total milliseconds, launches) of the timed kernel since the last call."""
        t = C.c_double()
        n = C.c_uint32()
        check(self._L.apbf_scene_kernel_time(self._h, C.byref(t), C.byref(n)), "kernel_time")
        return t.value, n.value

    def diagnostics(self) -> dict:
        c = np.zeros(32, np.uint32)
        check(self._L.apbf_scene_diagnostics(self._h, _ptr(c), 32), "diagnostics")
        return {"particles": int(c[0]), "pairs": int(c[2]), "transfers": int(c[3]), "splits": int(c[4]),
                "overflow_flags": int(c[6]), "pairs_before_clamp": int(c[7]), "rows": int(c[14]),
                "peak_pairs_before_clamp": int(c[17]), "longest_outgrown_row": int(c[27]),
                "rows_outgrew_slots": int(c[22]), "walk_candidates": int(c[28]) | (int(c[29]) << 32)}

    def count_walk_candidates(self, enable: bool) -> None:
        check(self._L.apbf_scene_count_walk_candidates(self._h, int(bool(enable))), "count_walk_candidates")

    def overflow_flags(self) -> int:
        f = C.c_uint32()
        check(self._L.apbf_scene_overflow_flags(self._h, C.byref(f)), "overflow_flags")
        return f.value


def _group_key(p: dict):
    """Scenes that can share a packed list: the same settings, flags, solver,
    gravity, time step and dimension count (grids and boxes may differ)."""
    return (tuple(sorted(p["settings"].items())), int(p["basic_pbf"]), int(p["solver_iterations"]),
            float(p["split_duration"]), tuple(float(g) for g in p["gravity"]), float(p.get("dt", 0.0)),
            int(p.get("dimensions", 3)))


def _split_states(st: dict, ids: np.ndarray, n_scenes: int) -> list:
    """The states of the scenes of a packed state: each scene's particles in
    list order, its transfers and neighbour pairs with endpoints renumbered
    to the scene's own list."""
    n = len(ids)
    order = np.argsort(ids, kind="stable")
    counts = np.bincount(ids, minlength=n_scenes)
    start = np.concatenate([[0], np.cumsum(counts)])
    local = np.empty(n, np.int64)
    local[order] = np.arange(n) - start[ids[order]]
    keys = ("position", "velocity", "inverse_mass", "radius", "pos_backup", "transferring", "target_radius",
            "kernel_width", "boundariness", "boundary_distance")
    sorted_fields = {k: st[k][order] for k in keys}
    ts = np.asarray(st["t_source"], np.int64)
    t_sc = ids[ts] if len(ts) else np.zeros(0, np.int64)
    t_order = np.argsort(t_sc, kind="stable")
    t_start = np.concatenate([[0], np.cumsum(np.bincount(t_sc, minlength=n_scenes))])
    pr = st.get("pairs")
    if pr is not None and len(pr):
        p_sc = ids[pr[:, 0]]
        p_order = np.argsort(p_sc, kind="stable")
        p_start = np.concatenate([[0], np.cumsum(np.bincount(p_sc, minlength=n_scenes))])
    out = []
    for k in range(n_scenes):
        a, b = start[k], start[k + 1]
        one = {key: sorted_fields[key][a:b] for key in keys}
        one["fluid_index"] = np.arange(b - a, dtype=np.uint32)
        sel = t_order[t_start[k]:t_start[k + 1]]
        one["t_source"] = local[ts[sel]].astype(np.uint32)
        one["t_target"] = local[np.asarray(st["t_target"], np.int64)[sel]].astype(np.uint32)
        one["t_time_left"] = st["t_time_left"][sel]
        if pr is not None and len(pr):
            ps = pr[p_order[p_start[k]:p_start[k + 1]]]
            one["pairs"] = np.stack([local[ps[:, 0]], local[ps[:, 1]]], 1).astype(np.uint32)
        else:
            one["pairs"] = np.zeros((0, 2), np.uint32)
        one["last_delta_time"] = st["last_delta_time"]
        out.append(one)
    return out


class PackedBatch:
    """Independent scenes stepped as packed lists (apbf_scene_create_packed):
    scenes with the same settings share one particle list, so every stage of
    a substep is one launch over all of them. ``cases`` is a list of
    (params, state); the groups' capacities are the sums of their scenes'.
    Each list is stepped on a stream of its own (default priority), so the
    lists overlap on the device."""

    def __init__(self, cases, stream: int | None = None):
        import torch
        self._L = _lib.lib()
        self._streams = []
        groups: dict = {}
        for k, (p, _) in enumerate(cases):
            groups.setdefault(_group_key(p), []).append(k)
        self.n_scenes = len(cases)
        self.groups = []
        for members in groups.values():
            ps = [cases[k][0] for k in members]
            gp = dict(ps[0])
            for cap in ("cap_particles", "cap_transfers", "cap_pairs"):
                gp[cap] = int(sum(int(p[cap]) for p in ps))
            gp["boxes"] = []
            gp["hash_res_log2"] = max(int(p["hash_res_log2"]) for p in ps)
            segs = (_lib.SegmentDesc * len(ps))(*[_lib.make_segment(p) for p in ps])
            cp = _lib.make_params(gp)
            h = C.c_void_p()
            check(self._L.apbf_scene_create_packed(C.byref(cp), len(ps), segs, C.byref(h)), "apbf_scene_create_packed")
            sc = Scene.__new__(Scene)
            sc._L, sc.params, sc._p, sc._h = self._L, gp, cp, h
            if stream is not None:
                sc.set_stream(stream)
            else:
                st_obj = torch.cuda.Stream()
                self._streams.append(st_obj)
                sc.set_stream(st_obj.cuda_stream)
            states = [cases[k][1] for k in members]
            joined = _join_states(states)
            sc.set_state(joined)
            ids = np.concatenate([np.full(len(st["radius"]), j, np.uint32) for j, st in enumerate(states)])
            check(self._L.apbf_scene_set_segment_ids(sc._h, _ptr(ids)), "apbf_scene_set_segment_ids")
            self.groups.append({"scene": sc, "members": members})

    def update(self, dt: float) -> None:
        for g in self.groups:
            g["scene"].update(dt)

    def synchronize(self) -> None:
        import torch
        torch.cuda.synchronize()

    def scene_states(self, pairs: bool = True) -> list:
        """Per-scene states in the order of ``cases``."""
        out = [None] * self.n_scenes
        for g in self.groups:
            sc = g["scene"]
            st = sc.get_state(pairs=pairs)
            n = len(st["radius"])
            ids = np.zeros(max(n, 1), np.uint32)
            check(self._L.apbf_scene_get_segment_ids(sc._h, _ptr(ids)), "apbf_scene_get_segment_ids")
            for k, one in zip(g["members"], _split_states(st, ids[:n].astype(np.int64), len(g["members"]))):
                out[k] = one
        return out

    def particle_substeps(self, reset: bool = False) -> int:
        return sum(g["scene"].particle_substeps(reset)[0] for g in self.groups)

    def close(self):
        for g in self.groups:
            g["scene"].close()
        self.groups = []


def _join_states(states: list) -> dict:
    """Scenes' states one after the other (transfer endpoints moved along)."""
    keys = ("position", "velocity", "inverse_mass", "radius", "pos_backup", "transferring", "target_radius",
            "kernel_width", "boundariness", "boundary_distance")
    out = {k: np.concatenate([np.asarray(st[k]) for st in states]) for k in keys}
    off = np.cumsum([0] + [len(st["radius"]) for st in states])[:-1]
    out["t_source"] = np.concatenate([np.asarray(st["t_source"], np.uint32) + np.uint32(o)
                                      for st, o in zip(states, off)]).astype(np.uint32)
    out["t_target"] = np.concatenate([np.asarray(st["t_target"], np.uint32) + np.uint32(o)
                                      for st, o in zip(states, off)]).astype(np.uint32)
    out["t_time_left"] = np.concatenate([np.asarray(st["t_time_left"], np.float32) for st in states])
    out["fluid_index"] = np.arange(len(out["radius"]), dtype=np.uint32)
    out["last_delta_time"] = states[0].get("last_delta_time", 1.0)
    return out


class Batch:
    """Independent scenes stepped together (apbf_batch_*)."""

    def __init__(self, scenes):
        self._L = _lib.lib()
        self.scenes = list(scenes)
        arr = (C.c_void_p * len(self.scenes))(*[s._h for s in self.scenes])
        h = C.c_void_p()
        check(self._L.apbf_batch_create(arr, len(self.scenes), C.byref(h)), "apbf_batch_create")
        self._h = h

    def update(self, dt: float) -> None:
        check(self._L.apbf_batch_update(self._h, float(dt)), "apbf_batch_update")

    def capture(self, dt: float) -> None:
        check(self._L.apbf_batch_capture(self._h, float(dt)), "apbf_batch_capture")

    def replay(self, times: int = 1) -> None:
        check(self._L.apbf_batch_replay(self._h, int(times)), "apbf_batch_replay")

    def synchronize(self) -> None:
        check(self._L.apbf_batch_synchronize(self._h), "apbf_batch_synchronize")

    def close(self):
        if getattr(self, "_h", None):
            self._L.apbf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
