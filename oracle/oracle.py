"""Test infrastructure only: ctypes front end of the scalar CPU oracle.

Loads ``oracle/liboracle.so`` (built by ``oracle/Makefile``) and exposes the
oracle scene (a sequential restatement of APBF's per-substep pipeline) and its
unit functions on numpy arrays. Only the tests, the smoke check and the
benchmark's CPU-baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def build_variant(mutation: int, out_dir: str):
    """Builds the oracle with the deliberately wrong variant `mutation` of one
    expression (ORACLE_MUTATION in apbf_oracle.cpp) into `out_dir` and returns
    the loaded library. Used only to show that a pin fails on a planted mistake."""
    out = os.path.join(out_dir, f"liboracle_m{mutation}.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-mfma",
                    f"-DORACLE_MUTATION={int(mutation)}", "-o", out, os.path.join(_HERE, "apbf_oracle.cpp")],
                   check=True)
    L = C.CDLL(out)
    _declare(L)
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class OrcSettings(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "mHeightKernelId", "mGradientKernelId", "mMerge", "mSplit",
        "mBaseKernelWidthOnTargetRadius", "mBaseKernelWidthOnBoundaryDistance",
        "mUpdateTargetRadius", "mUpdateBoundariness", "mNeighborListSorted",
        "mBoundarinessCalculationMethod")] + [(n, C.c_float) for n in (
        "mBoundarinessAdaptionSpeed", "mKernelWidthAdaptionSpeed",
        "mBoundarinessSelfGradLengthFactor", "mBoundarinessUnderpressureFactor",
        "mMergeDuration", "mSmallestTargetRadius", "mTargetRadiusOffset",
        "mTargetRadiusScaleFactor")]


MAX_BOXES = 16


class OrcSceneParams(C.Structure):
    _fields_ = [
        ("s", OrcSettings),
        ("basic_pbf", C.c_int32),
        ("solver_iterations", C.c_int32),
        ("split_duration", C.c_float),
        ("gravity", C.c_float * 3),
        ("hash_min", C.c_float * 3),
        ("hash_max", C.c_float * 3),
        ("hash_res_log2", C.c_uint32),
        ("n_boxes", C.c_uint32),
        ("box_min", (C.c_float * 4) * MAX_BOXES),
        ("box_max", (C.c_float * 4) * MAX_BOXES),
        ("cap_particles", C.c_uint32),
        ("cap_transfers", C.c_uint32),
        ("cap_pairs", C.c_uint32),
        ("dimensions", C.c_uint32),
    ]


def _p(t):
    return C.POINTER(t)


def _declare(L):
    u32, i32, f32, f64, vp = C.c_uint32, C.c_int32, C.c_float, C.c_double, C.c_void_p
    pu, pi, pf = _p(u32), _p(i32), _p(f32)
    L.orc_create.restype = vp
    L.orc_create.argtypes = [_p(OrcSceneParams)]
    L.orc_destroy.argtypes = [vp]
    L.orc_set_state.restype = C.c_int
    L.orc_set_state.argtypes = [vp, u32, pi, pf, pf, pf, pi, pu, u32, pu, pf, pf, pf, pu, u32, pu, pu, pf, f32]
    L.orc_get_counts.argtypes = [vp, pu, pu, pu, pu]
    L.orc_get_state.argtypes = [vp, pi, pf, pf, pf, pi, pu, pu, pf, pf, pf, pu, pu, pu, pf, pu, pf]
    L.orc_get_lambda.argtypes = [vp, pf]
    L.orc_get_splits.restype = u32
    L.orc_get_splits.argtypes = [vp, pu]
    L.orc_get_fsm.restype = u32
    L.orc_get_fsm.argtypes = [vp, pu, pu, pu]
    L.orc_stage.argtypes = [vp, C.c_int, f32]
    L.orc_set_neighborhood_method.argtypes = [vp, C.c_int]
    L.orc_set_sphere.argtypes = [vp, pf, f32, C.c_int, C.c_int]
    L.orc_update.argtypes = [vp, f32]
    L.orc_set_solver_tolerance.argtypes = [vp, f32, C.c_int]
    L.orc_set_warm_lambdas.argtypes = [vp, pf, u32]
    L.orc_solver_stats.argtypes = [vp, _p(C.c_uint64), _p(C.c_uint64), _p(C.c_uint64)]
    L.orc_density_error.argtypes = [vp, _p(f64), _p(f64)]
    L.orc_stable_radix_sort.argtypes = [pu, pu, u32, u32, pu, pu]
    L.orc_prefix_sum.argtypes = [pu, u32]
    L.orc_position_hash.argtypes = [pi, u32, pf, pf, u32, u32, pu]
    L.orc_position_code.argtypes = [pi, u32, u32, pu]
    L.orc_find_value_ranges.argtypes = [pu, u32, u32, pu, pu]
    L.orc_neighborhood_green.restype = u32
    L.orc_neighborhood_green.argtypes = [pi, pf, u32, pu, pu, f32, pf, pf, u32, u32, pu, u32]
    L.orc_neighborhood_brute_force.restype = u32
    L.orc_neighborhood_brute_force.argtypes = [pi, pf, u32, f32, pu, u32]
    L.orc_apply_hidden_edit.restype = u32
    L.orc_apply_hidden_edit.argtypes = [pu, u32, u32, pu, u32, u32, pu, pu]
    L.orc_delete_edit_list.restype = u32
    L.orc_delete_edit_list.argtypes = [pu, u32, u32, pu]
    L.orc_exp.restype = f32
    L.orc_exp.argtypes = [f32]
    L.orc_pow.restype = f32
    L.orc_pow.argtypes = [f32, f32]
    L.orc_kernel.argtypes = [C.c_int, pf, f32, u32, pf, pf]


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(_p(t))


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ------------------------------------------------------------------ params
def make_params(p: dict) -> OrcSceneParams:
    """Build the C parameter block from a plain dict (keys as in the struct;
    `settings` is a dict of the apbf_settings fields, `boxes` a list of
    (min3, max3) pairs)."""
    out = OrcSceneParams()
    for k, v in p["settings"].items():
        setattr(out.s, k, v)
    out.basic_pbf = int(p["basic_pbf"])
    out.solver_iterations = int(p["solver_iterations"])
    out.split_duration = float(p["split_duration"])
    for i in range(3):
        out.gravity[i] = p["gravity"][i]
        out.hash_min[i] = p["hash_min"][i]
        out.hash_max[i] = p["hash_max"][i]
    out.hash_res_log2 = int(p["hash_res_log2"])
    boxes = p["boxes"]
    if len(boxes) > MAX_BOXES:
        raise ValueError("too many boxes")
    out.n_boxes = len(boxes)
    for i, (bmin, bmax) in enumerate(boxes):
        for d in range(3):
            out.box_min[i][d] = bmin[d]
            out.box_max[i][d] = bmax[d]
    out.cap_particles = int(p["cap_particles"])
    out.cap_transfers = int(p["cap_transfers"])
    out.cap_pairs = int(p["cap_pairs"])
    out.dimensions = int(p.get("dimensions", 3))
    return out


STAGE = {
    "velocity": 0, "particle_transfer": 1, "kernel_width_bd": 2, "neighborhood": 3,
    "spread": 4, "box": 5, "incompressibility": 6, "update_transfers": 7, "sphere": 8,
}


class OracleScene:
    """One scene held by the oracle. State dicts use the keys of
    :data:`STATE_KEYS`."""

    def __init__(self, params: dict, library=None):
        self._L = library if library is not None else lib()
        self._params = make_params(params)
        self._h = self._L.orc_create(C.byref(self._params))
        if params.get("sphere"):  # (center, radius, hollow) of the scene's sphere collider
            c, r, hollow = params["sphere"]
            self.set_sphere(c, r, hollow)

    def close(self):
        if self._h:
            self._L.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, st: dict) -> None:
        n = len(st["radius"])
        nf = len(st["fluid_index"])
        nt = len(st["t_source"])
        arrs = dict(
            position=_i32(st["position"]).reshape(n, 4), velocity=_f32(st["velocity"]).reshape(n, 4),
            inverse_mass=_f32(st["inverse_mass"]), radius=_f32(st["radius"]),
            pos_backup=_i32(st["pos_backup"]).reshape(n, 4), transferring=_u32(st["transferring"]),
            fluid_index=_u32(st["fluid_index"]), target_radius=_f32(st["target_radius"]),
            kernel_width=_f32(st["kernel_width"]), boundariness=_f32(st["boundariness"]),
            boundary_distance=_u32(st["boundary_distance"]), t_source=_u32(st["t_source"]),
            t_target=_u32(st["t_target"]), t_time_left=_f32(st["t_time_left"]))
        self._keep = arrs
        a = arrs
        rc = self._L.orc_set_state(
            self._h, n, _ptr(a["position"], C.c_int32), _ptr(a["velocity"], C.c_float),
            _ptr(a["inverse_mass"], C.c_float), _ptr(a["radius"], C.c_float),
            _ptr(a["pos_backup"], C.c_int32), _ptr(a["transferring"], C.c_uint32), nf,
            _ptr(a["fluid_index"], C.c_uint32), _ptr(a["target_radius"], C.c_float),
            _ptr(a["kernel_width"], C.c_float), _ptr(a["boundariness"], C.c_float),
            _ptr(a["boundary_distance"], C.c_uint32), nt, _ptr(a["t_source"], C.c_uint32),
            _ptr(a["t_target"], C.c_uint32), _ptr(a["t_time_left"], C.c_float),
            float(st.get("last_delta_time", 1.0)))
        if rc != 0:
            raise ValueError("invalid oracle state")

    def counts(self):
        v = [C.c_uint32() for _ in range(4)]
        self._L.orc_get_counts(self._h, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def get_state(self) -> dict:
        nh, nf, nt, npairs = self.counts()
        st = dict(
            position=np.zeros((nh, 4), np.int32), velocity=np.zeros((nh, 4), np.float32),
            inverse_mass=np.zeros(nh, np.float32), radius=np.zeros(nh, np.float32),
            pos_backup=np.zeros((nh, 4), np.int32), transferring=np.zeros(nh, np.uint32),
            fluid_index=np.zeros(nf, np.uint32), target_radius=np.zeros(nf, np.float32),
            kernel_width=np.zeros(nf, np.float32), boundariness=np.zeros(nf, np.float32),
            boundary_distance=np.zeros(nf, np.uint32), t_source=np.zeros(nt, np.uint32),
            t_target=np.zeros(nt, np.uint32), t_time_left=np.zeros(nt, np.float32),
            pairs=np.zeros((npairs, 2), np.uint32))
        ldt = C.c_float()
        s = st
        self._L.orc_get_state(
            self._h, _ptr(s["position"], C.c_int32), _ptr(s["velocity"], C.c_float),
            _ptr(s["inverse_mass"], C.c_float), _ptr(s["radius"], C.c_float),
            _ptr(s["pos_backup"], C.c_int32), _ptr(s["transferring"], C.c_uint32),
            _ptr(s["fluid_index"], C.c_uint32), _ptr(s["target_radius"], C.c_float),
            _ptr(s["kernel_width"], C.c_float), _ptr(s["boundariness"], C.c_float),
            _ptr(s["boundary_distance"], C.c_uint32), _ptr(s["t_source"], C.c_uint32),
            _ptr(s["t_target"], C.c_uint32), _ptr(s["t_time_left"], C.c_float),
            _ptr(s["pairs"], C.c_uint32), C.byref(ldt))
        st["last_delta_time"] = ldt.value
        return st

    def lambdas(self) -> np.ndarray:
        nf = self.counts()[1]
        out = np.zeros(nf, np.float32)
        self._L.orc_get_lambda(self._h, _ptr(out, C.c_float))
        return out

    def fsm(self):
        """Intermediates of the last update_transfers stage per fluid id: the
        boundary distance after the flood fill, the shortest neighbour distance
        and the nearest neighbour (all u32)."""
        n = self._L.orc_get_fsm(self._h, None, None, None)
        bd, md, nn = (np.zeros(max(n, 1), np.uint32) for _ in range(3))
        self._L.orc_get_fsm(self._h, _ptr(bd, C.c_uint32), _ptr(md, C.c_uint32), _ptr(nn, C.c_uint32))
        return bd[:n], md[:n], nn[:n]

    def splits(self) -> np.ndarray:
        n = self._L.orc_get_splits(self._h, None)
        out = np.zeros(max(n, 1), np.uint32)
        self._L.orc_get_splits(self._h, _ptr(out, C.c_uint32))
        return out[:n]

    def set_neighborhood_method(self, method: str) -> None:
        self._L.orc_set_neighborhood_method(self._h, {"green": 0, "binary_search": 1, "brute_force": 2}[method])

    def set_sphere(self, center, radius: float, hollow: bool = True, enable: bool = True) -> None:
        c = np.asarray(center, np.float32)
        self._L.orc_set_sphere(self._h, _ptr(c, C.c_float), float(radius), int(hollow), int(enable))

    def stage(self, name: str, dt: float = 0.0) -> None:
        self._L.orc_stage(self._h, STAGE[name], float(dt))

    def update(self, dt: float) -> None:
        self._L.orc_update(self._h, float(dt))

    def set_solver_tolerance(self, tol: float, max_iterations: int = 16) -> None:
        self._L.orc_set_solver_tolerance(self._h, float(tol), int(max_iterations))

    def set_warm_lambdas(self, lam) -> None:
        a = _f32(lam)
        self._warm = a
        self._L.orc_set_warm_lambdas(self._h, _ptr(a, C.c_float), len(a))

    def solver_stats(self):
        v = [C.c_uint64() for _ in range(3)]
        self._L.orc_solver_stats(self._h, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def density_error(self):
        m, x = C.c_double(), C.c_double()
        self._L.orc_density_error(self._h, C.byref(m), C.byref(x))
        return m.value, x.value


# -------------------------------------------------------------- unit functions
def stable_radix_sort(keys, values, upper_bound=0xFFFFFFFF):
    k, v = _u32(keys), _u32(values)
    ok, ov = np.zeros_like(k), np.zeros_like(v)
    lib().orc_stable_radix_sort(_ptr(k, C.c_uint32), _ptr(v, C.c_uint32), len(k), upper_bound,
                                _ptr(ok, C.c_uint32), _ptr(ov, C.c_uint32))
    return ok, ov


def prefix_sum(values):
    v = _u32(values).copy()
    lib().orc_prefix_sum(_ptr(v, C.c_uint32), len(v))
    return v


def position_hash(position, mn, mx, res, dims=3):
    p = _i32(position).reshape(-1, 4)
    out = np.zeros(len(p), np.uint32)
    lib().orc_position_hash(_ptr(p, C.c_int32), len(p), _ptr(_f32(mn), C.c_float),
                            _ptr(_f32(mx), C.c_float), res, dims, _ptr(out, C.c_uint32))
    return out


def position_code(position, section):
    p = _i32(position).reshape(-1, 4)
    out = np.zeros(len(p), np.uint32)
    lib().orc_position_code(_ptr(p, C.c_int32), len(p), section, _ptr(out, C.c_uint32))
    return out


def find_value_ranges(sorted_keys, n_cells):
    k = _u32(sorted_keys)
    s = np.zeros(n_cells, np.uint32)
    e = np.zeros(n_cells, np.uint32)
    lib().orc_find_value_ranges(_ptr(k, C.c_uint32), len(k), n_cells, _ptr(s, C.c_uint32),
                                _ptr(e, C.c_uint32))
    return s, e


def neighborhood_green(position, rng, cell_start, cell_end, range_scale, mn, mx, res, cap, dims=3):
    p = _i32(position).reshape(-1, 4)
    r = _f32(rng)
    cs, ce = _u32(cell_start), _u32(cell_end)
    out = np.zeros((cap, 2), np.uint32)
    n = lib().orc_neighborhood_green(_ptr(p, C.c_int32), _ptr(r, C.c_float), len(p),
                                     _ptr(cs, C.c_uint32), _ptr(ce, C.c_uint32), range_scale,
                                     _ptr(_f32(mn), C.c_float), _ptr(_f32(mx), C.c_float), res, dims,
                                     _ptr(out, C.c_uint32), cap)
    return out[:n]


def neighborhood_brute_force(position, rng, range_scale, cap):
    p = _i32(position).reshape(-1, 4)
    r = _f32(rng)
    out = np.zeros((cap, 2), np.uint32)
    n = lib().orc_neighborhood_brute_force(_ptr(p, C.c_int32), _ptr(r, C.c_float), len(p),
                                           range_scale, _ptr(out, C.c_uint32), cap)
    return out[:n]


def apply_hidden_edit(index_list, n_hidden, edit, max_new):
    il, e = _u32(index_list), _u32(edit)
    ni = np.zeros(max(max_new, 1), np.uint32)
    ne = np.zeros(max(max_new, 1), np.uint32)
    n = lib().orc_apply_hidden_edit(_ptr(il, C.c_uint32), len(il), n_hidden, _ptr(e, C.c_uint32),
                                    len(e), max_new, _ptr(ni, C.c_uint32), _ptr(ne, C.c_uint32))
    return ni[:n], ne[:n]


def delete_edit_list(deleted, n_hidden):
    d = _u32(deleted)
    out = np.zeros(max(n_hidden, 1), np.uint32)
    n = lib().orc_delete_edit_list(_ptr(d, C.c_uint32), len(d), n_hidden, _ptr(out, C.c_uint32))
    return out[:n]


def exp(x: float) -> float:
    return lib().orc_exp(x)


def pow(x: float, y: float) -> float:  # noqa: A001 - mirrors the GLSL name
    return lib().orc_pow(x, y)


def kernel(kernel_id: int, r3, kernel_width: float, dims: int = 3):
    r = _f32(r3)
    h = C.c_float()
    g = np.zeros(3, np.float32)
    lib().orc_kernel(kernel_id, _ptr(r, C.c_float), kernel_width, dims, C.byref(h), _ptr(g, C.c_float))
    return h.value, g
