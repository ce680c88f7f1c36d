"""ctypes binding of libapbf_hip.so (include/apbf.h).

The library is the only compute path: importing a function that needs it
raises if the library is missing, and nothing falls back to a CPU
implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libapbf_hip.so")
_lib = None

STATUS = {0: "ok", -1: "invalid argument", -2: "HIP error", -3: "capacity exceeded",
          -4: "no HIP device", -5: "workspace too small"}


class ApbfError(RuntimeError):
    pass


def check(status: int, what: str = "") -> None:
    if status != 0:
        raise ApbfError(f"{what}: {STATUS.get(status, status)}")


class Settings(C.Structure):
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


class SceneParams(C.Structure):
    _fields_ = [
        ("s", Settings),
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


class SegmentDesc(C.Structure):
    _fields_ = [
        ("hash_min", C.c_float * 3),
        ("hash_max", C.c_float * 3),
        ("hash_res_log2", C.c_uint32),
        ("n_boxes", C.c_uint32),
        ("box_min", (C.c_float * 4) * MAX_BOXES),
        ("box_max", (C.c_float * 4) * MAX_BOXES),
    ]


def make_segment(p: dict) -> SegmentDesc:
    """The hash grid and boxes of one scene of a packed batch."""
    out = SegmentDesc()
    for i in range(3):
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
    return out


def make_params(p: dict) -> SceneParams:
    out = SceneParams()
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


# name -> (restype, argtypes)
_u32, _i32, _f32, _f64, _vp, _sz = C.c_uint32, C.c_int32, C.c_float, C.c_double, C.c_void_p, C.c_size_t
_pu, _pi, _pf, _pd = C.POINTER(_u32), C.POINTER(_i32), C.POINTER(_f32), C.POINTER(_f64)
PROTOTYPES = {
    "apbf_scene_create": (_i32, [C.POINTER(SceneParams), C.POINTER(_vp)]),
    "apbf_scene_destroy": (_i32, [_vp]),
    "apbf_scene_create_packed": (_i32, [C.POINTER(SceneParams), _u32, C.POINTER(SegmentDesc), C.POINTER(_vp)]),
    "apbf_scene_set_segment_ids": (_i32, [_vp, _vp]),
    "apbf_scene_get_segment_ids": (_i32, [_vp, _vp]),
    "apbf_scene_set_stream": (_i32, [_vp, _vp]),
    "apbf_scene_set_state": (_i32, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                    _u32, _vp, _vp, _vp, _f32]),
    "apbf_scene_get_counts": (_i32, [_vp, _pu, _pu, _pu]),
    "apbf_scene_get_state": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                    _vp, _vp, _pf]),
    "apbf_scene_get_lambda": (_i32, [_vp, _vp]),
    "apbf_scene_stage": (_i32, [_vp, _i32, _f32]),
    "apbf_scene_update": (_i32, [_vp, _f32]),
    "apbf_scene_density_error": (_i32, [_vp, _pd, _pd]),
    "apbf_scene_overflow_flags": (_i32, [_vp, _pu]),
    "apbf_scene_set_error_measurement": (_i32, [_vp, _i32]),
    "apbf_scene_error_stats": (_i32, [_vp, _pd, _pd, _pd, _pd, _pu, _i32]),
    "apbf_scene_iteration_errors": (_i32, [_vp, _vp, _u32, _i32]),
    "apbf_scene_particle_substeps": (_i32, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), _i32]),
    "apbf_scene_set_solver_tolerance": (_i32, [_vp, _f32, _i32]),
    "apbf_scene_get_warm_lambdas": (_i32, [_vp, _vp]),
    "apbf_scene_solver_stats": (_i32, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                       _i32]),
    "apbf_scene_time_kernel": (_i32, [_vp, _i32, _u32]),
    "apbf_scene_time_kernels": (_i32, [_vp, _u32, _u32]),
    "apbf_scene_kernel_time_of": (_i32, [_vp, _i32, _pd, _pu]),
    "apbf_scene_substep_totals": (_i32, [_vp, _vp, _u32, _i32]),
    "apbf_scene_diagnostics": (_i32, [_vp, _vp, _u32]),
    "apbf_scene_kernel_time": (_i32, [_vp, _pd, _pu]),
    "apbf_prefix_sum_workspace_size": (_sz, [_u32]),
    "apbf_radix_sort_workspace_size": (_sz, [_u32]),
    "apbf_neighborhood_green_workspace_size": (_sz, [_u32]),
    "apbf_prefix_sum": (_i32, [_vp, _vp, _u32, _vp, _sz, _vp]),
    "apbf_radix_sort": (_i32, [_vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp, _sz, _vp]),
    "apbf_radix_sort_noalloc": (_i32, [_vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp]),
    "apbf_calculate_position_hash": (_i32, [_vp, _vp, _u32, _pf, _pf, _u32, _u32, _vp, _vp]),
    "apbf_calculate_position_code": (_i32, [_vp, _vp, _u32, _u32, _vp, _vp]),
    "apbf_find_value_ranges": (_i32, [_vp, _vp, _u32, _u32, _vp, _vp, _vp]),
    "apbf_neighborhood_green": (_i32, [_vp, _vp, _vp, _u32, _f32, _pf, _pf, _u32, _u32, _vp, _vp, _vp, _vp,
                                       _u32, _vp, _vp, _sz, _vp]),
    "apbf_neighborhood_brute_force": (_i32, [_vp, _vp, _vp, _u32, _f32, _vp, _vp, _u32, _vp, _vp, _sz, _vp]),
    "apbf_copy_scattered_read": (_i32, [_vp, _vp, _vp, _vp, _u32, _u32, _vp]),
    "apbf_hidden_edit_workspace_size": (_sz, [_u32, _u32, _u32]),
    "apbf_apply_hidden_edit": (_i32, [_vp, _vp, _u32, _u32, _vp, _vp, _u32, _u32, _vp, _vp, _vp, _vp,
                                      _sz, _vp]),
    "apbf_delete_edit_list_workspace_size": (_sz, [_u32]),
    "apbf_delete_edit_list": (_i32, [_vp, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _sz, _vp]),
    "apbf_append_list": (_i32, [_vp, _u32, _vp, _vp, _vp, _u32, _vp, _u32, _vp]),
    "apbf_write_increasing_sequence": (_i32, [_vp, _u32, _vp, _u32, _u32, _vp, _vp, _vp]),
    "apbf_write_increasing_sequence_from_to": (_i32, [_vp, _u32, _vp, _vp, _vp, _vp]),
    "apbf_write_sequence": (_i32, [_vp, _vp, _u32, _u32, _u32, _vp]),
    "apbf_device_alloc": (_i32, [_sz, C.POINTER(_vp)]),
    "apbf_device_free": (_i32, [_vp]),
    "apbf_copy": (_i32, [_vp, _vp, _sz, _vp]),
    "apbf_stream_synchronize": (_i32, [_vp]),
    "apbf_context_create": (_i32, [C.POINTER(SceneParams), C.POINTER(_vp)]),
    "apbf_context_destroy": (_i32, [_vp]),
    "apbf_op_velocity_handling": (_i32, [_vp, _vp, _pf, _f32, _pf, _vp]),
    "apbf_op_particle_transfer": (_i32, [_vp, _vp, _vp, _f32, _vp]),
    "apbf_op_kernel_width_from_boundary_distance": (_i32, [_vp, _vp, _vp]),
    "apbf_op_neighborhood": (_i32, [_vp, _vp, _vp, _vp, _i32, _f32, _pf, _pf, _u32, _vp]),
    "apbf_op_spread_kernel_width": (_i32, [_vp, _vp, _vp, _vp]),
    "apbf_op_box_collision": (_i32, [_vp, _vp, _u32, _pf, _pf, _vp]),
    "apbf_op_sphere_collision": (_i32, [_vp, _vp, _pf, _f32, _i32, _vp]),
    "apbf_op_incompressibility": (_i32, [_vp, _vp, _vp, _vp]),
    "apbf_op_update_transfers": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "apbf_neighbor_pairs": (_i32, [_vp, _vp, _vp, _vp, _u32, _u32, _vp, _vp]),
    "apbf_neighbor_rows_from_pairs_workspace_size": (_sz, [_u32, _u32]),
    "apbf_neighbor_rows_from_pairs": (_i32, [_vp, _vp, _u32, _u32, _vp, _vp, _vp, _sz, _vp]),
    "apbf_scene_set_gravity": (_i32, [_vp, _pf]),
    "apbf_scene_set_neighborhood_method": (_i32, [_vp, _i32]),
    "apbf_scene_set_range_scale": (_i32, [_vp, _f32]),
    "apbf_scene_set_position_range": (_i32, [_vp, _pf, _pf, _u32]),
    "apbf_scene_capture": (_i32, [_vp, _f32]),
    "apbf_scene_replay": (_i32, [_vp, _u32]),
    "apbf_batch_create": (_i32, [C.POINTER(_vp), _u32, C.POINTER(_vp)]),
    "apbf_batch_destroy": (_i32, [_vp]),
    "apbf_batch_update": (_i32, [_vp, _f32]),
    "apbf_batch_capture": (_i32, [_vp, _f32]),
    "apbf_batch_replay": (_i32, [_vp, _u32]),
    "apbf_batch_synchronize": (_i32, [_vp]),
    "apbf_scene_set_launch_shape": (_i32, [_vp, _i32, _i32]),
    "apbf_scene_set_long_rows": (_i32, [_vp, _i32]),
    "apbf_scene_set_sort_shape": (_i32, [_vp, _i32, _i32]),
    "apbf_scene_count_walk_candidates": (_i32, [_vp, _i32]),
    "apbf_scene_sort_shape": (_i32, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "apbf_scene_autotune": (_i32, [_vp, _f32, _u32, _pi, _pi, _pf]),
    "apbf_mlp_create": (_i32, [C.POINTER(_vp)]),
    "apbf_mlp_destroy": (_i32, [_vp]),
    "apbf_mlp_set_params": (_i32, [_vp, _vp, _vp]),
    "apbf_mlp_get_params": (_i32, [_vp, _vp, _vp]),
    "apbf_mlp_forward": (_i32, [_vp, _vp, _vp, _u32, _vp, _vp]),
    "apbf_mlp_train": (_i32, [_vp, _vp, _vp, _vp, _u32, _u32, _u32, _f32, _vp, _vp]),
    "apbf_mlp_gradient": (_i32, [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp]),
    "apbf_scene_set_warm_start": (_i32, [_vp, _vp, _i32]),
    "apbf_scene_harvest": (_i32, [_vp, _vp, _vp, _u32, _pu]),
    "apbf_version": (C.c_char_p, []),
    "apbf_device_count": (_i32, []),
    "apbf_scene_set_sphere": (_i32, [_vp, _vp, _f32, _i32, _i32]),
    "apbf_snapshot_create": (_i32, [_vp, C.POINTER(_vp)]),
    "apbf_snapshot_save": (_i32, [_vp, _vp]),
    "apbf_snapshot_restore": (_i32, [_vp, _vp]),
    "apbf_snapshot_destroy": (_i32, [_vp]),
    "apbf_scene_save_particle_info": (_i32, [_vp, C.c_char_p]),
    "apbf_scene_save_measurements": (_i32, [_vp, C.c_char_p, _u32, _f32, _f32, _f32]),
    "apbf_scene_save_svg": (_i32, [_vp, C.c_char_p, _u32, _pf, _pf, _f32, _f32, _i32, _i32]),
    "apbf_check_math": (_i32, [_vp, _vp]),
    "apbf_check_conversions": (_i32, [_u32, C.POINTER(C.c_uint64), _pu, _vp, _vp]),
}


def lib():
    """Load libapbf_hip.so (built by apbf_amd/build.py); raise if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ApbfError(f"HIP library not built: {LIB_PATH} (run python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(PROTOTYPES)
