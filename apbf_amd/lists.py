"""Caller-owned device lists and the stage operators that run on them
(apbf_context / apbf_op_* in include/apbf.h).

The reference's subsystems are bound to lists with set_data and run their
shaders on those lists (source/list_definitions.h:9-20 for the lists,
source/pool.cpp:67-106 for the order of one substep). Here the lists are
torch device tensors owned by the caller (torch is the allocator only); the
operators are the library's kernels.

* :class:`ParticleLists` - hidden particles + fluid attributes (fluid id ==
  hidden slot), with a device length word.
* :class:`TransferLists` - transfer source / target / time left.
* :class:`NeighborLists` - CSR rows (the pair list of particle i is
  neighbors[row_offset[i]:row_offset[i+1]]); :meth:`NeighborLists.pairs`
  gives the reference's uvec2 layout.
* :class:`Context` - the temporaries of every stage, and one method per
  subsystem apply(); :func:`pool_update` runs one substep in the order of
  pool.cpp:67-106.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check


class ParticleListsC(C.Structure):
    _fields_ = [("position", C.c_void_p), ("velocity", C.c_void_p), ("inverse_mass", C.c_void_p),
                ("radius", C.c_void_p), ("pos_backup", C.c_void_p), ("transferring", C.c_void_p),
                ("target_radius", C.c_void_p), ("kernel_width", C.c_void_p), ("boundariness", C.c_void_p),
                ("boundary_distance", C.c_void_p), ("d_length", C.c_void_p), ("capacity", C.c_uint32)]


class TransferListsC(C.Structure):
    _fields_ = [("source", C.c_void_p), ("target", C.c_void_p), ("time_left", C.c_void_p),
                ("d_length", C.c_void_p), ("capacity", C.c_uint32)]


class NeighborListsC(C.Structure):
    _fields_ = [("row_offset", C.c_void_p), ("neighbors", C.c_void_p), ("d_length", C.c_void_p),
                ("d_rows", C.c_void_p), ("capacity", C.c_uint32)]


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(a: np.ndarray, device, dtype, size: int, width: int = 1) -> torch.Tensor:
    t = torch.zeros((size, width) if width > 1 else size, dtype=dtype, device=device)
    if len(a):
        t[: len(a)] = torch.from_numpy(np.ascontiguousarray(a)).to(device)
    return t


def _np(t: torch.Tensor, n: int, dtype) -> np.ndarray:
    return t[:n].cpu().numpy().view(dtype)


def _length(n: int, device) -> torch.Tensor:
    return torch.tensor([n], dtype=torch.int32, device=device)


class ParticleLists:
    F32 = ("inverse_mass", "radius", "target_radius", "kernel_width", "boundariness")
    U32 = ("transferring", "boundary_distance")

    def __init__(self, capacity: int, device="cuda"):
        self.capacity = int(capacity)
        self.device = torch.device(device)
        z = lambda dt, w=1: torch.zeros((capacity, w) if w > 1 else capacity, dtype=dt, device=self.device)
        self.position, self.pos_backup = z(torch.int32, 4), z(torch.int32, 4)
        self.velocity = z(torch.float32, 4)
        for k in self.F32:
            setattr(self, k, z(torch.float32))
        for k in self.U32:
            setattr(self, k, z(torch.int32))
        self.length = _length(0, self.device)

    @classmethod
    def from_state(cls, st: dict, capacity: int, device="cuda") -> "ParticleLists":
        n = len(st["radius"])
        if n > capacity:
            raise ValueError("state larger than the capacity")
        fi = np.asarray(st["fluid_index"], np.int64)
        if not np.array_equal(fi, np.arange(n)):
            raise ValueError("the lists keep fluid id == hidden slot")
        p = cls(capacity, device)
        d = p.device
        p.position = _dev(np.asarray(st["position"], np.int32).reshape(n, 4), d, torch.int32, capacity, 4)
        p.pos_backup = _dev(np.asarray(st["pos_backup"], np.int32).reshape(n, 4), d, torch.int32, capacity, 4)
        p.velocity = _dev(np.asarray(st["velocity"], np.float32).reshape(n, 4), d, torch.float32, capacity, 4)
        for k in cls.F32:
            setattr(p, k, _dev(np.asarray(st[k], np.float32), d, torch.float32, capacity))
        for k in cls.U32:
            setattr(p, k, _dev(np.asarray(st[k], np.uint32).view(np.int32), d, torch.int32, capacity))
        p.length = _length(n, d)
        return p

    def n(self) -> int:
        return int(self.length.item())

    def state(self) -> dict:
        n = self.n()
        st = dict(position=_np(self.position, n, np.int32), pos_backup=_np(self.pos_backup, n, np.int32),
                  velocity=_np(self.velocity, n, np.float32), fluid_index=np.arange(n, dtype=np.uint32))
        for k in self.F32:
            st[k] = _np(getattr(self, k), n, np.float32)
        for k in self.U32:
            st[k] = _np(getattr(self, k), n, np.uint32)
        return st

    def c(self) -> ParticleListsC:
        a = ParticleListsC()
        for k in ("position", "velocity", "inverse_mass", "radius", "pos_backup", "transferring", "target_radius",
                  "kernel_width", "boundariness", "boundary_distance"):
            setattr(a, k, getattr(self, k).data_ptr())
        a.d_length = self.length.data_ptr()
        a.capacity = self.capacity
        return a


class TransferLists:
    def __init__(self, capacity: int, device="cuda", st: dict | None = None):
        self.capacity = int(capacity)
        d = torch.device(device)
        nt = len(st["t_source"]) if st else 0
        self.source = _dev(np.asarray(st["t_source"], np.uint32).view(np.int32) if st else np.zeros(0, np.int32),
                           d, torch.int32, capacity)
        self.target = _dev(np.asarray(st["t_target"], np.uint32).view(np.int32) if st else np.zeros(0, np.int32),
                           d, torch.int32, capacity)
        self.time_left = _dev(np.asarray(st["t_time_left"], np.float32) if st else np.zeros(0, np.float32),
                              d, torch.float32, capacity)
        self.length = _length(nt, d)

    def state(self) -> dict:
        nt = int(self.length.item())
        return dict(t_source=_np(self.source, nt, np.uint32), t_target=_np(self.target, nt, np.uint32),
                    t_time_left=_np(self.time_left, nt, np.float32))

    def c(self) -> TransferListsC:
        a = TransferListsC()
        a.source, a.target, a.time_left = self.source.data_ptr(), self.target.data_ptr(), self.time_left.data_ptr()
        a.d_length = self.length.data_ptr()
        a.capacity = self.capacity
        return a


class NeighborLists:
    def __init__(self, particle_capacity: int, pair_capacity: int, device="cuda"):
        d = torch.device(device)
        self.particle_capacity = int(particle_capacity)
        self.capacity = int(pair_capacity)
        self.row_offset = torch.zeros(particle_capacity + 1, dtype=torch.int32, device=d)
        self.neighbors = torch.zeros(max(pair_capacity, 1), dtype=torch.int32, device=d)
        self.length = _length(0, d)
        self.rows = _length(0, d)

    def c(self) -> NeighborListsC:
        a = NeighborListsC()
        a.row_offset, a.neighbors = self.row_offset.data_ptr(), self.neighbors.data_ptr()
        a.d_length, a.d_rows = self.length.data_ptr(), self.rows.data_ptr()
        a.capacity = self.capacity
        return a

    def pairs(self) -> np.ndarray:
        """The reference's uvec2 pair list (apbf_neighbor_pairs), on the host."""
        L = _lib.lib()
        np_ = int(self.length.item())
        out = torch.zeros((max(np_, 1), 2), dtype=torch.int32, device=self.row_offset.device)
        check(L.apbf_neighbor_pairs(C.c_void_p(self.row_offset.data_ptr()), C.c_void_p(self.neighbors.data_ptr()),
                                    C.c_void_p(self.rows.data_ptr()), C.c_void_p(self.length.data_ptr()),
                                    self.particle_capacity, self.capacity, C.c_void_p(out.data_ptr()), _stream()),
              "apbf_neighbor_pairs")
        torch.cuda.current_stream().synchronize()
        return out[:np_].cpu().numpy().view(np.uint32)


class Context:
    """apbf_context: the temporaries of every stage for the capacities and
    settings of a scene parameter dict (apbf_amd.config)."""

    def __init__(self, params: dict):
        self._L = _lib.lib()
        self.params = params
        self._p = _lib.make_params(params)
        h = C.c_void_p()
        check(self._L.apbf_context_create(C.byref(self._p), C.byref(h)), "apbf_context_create")
        self._h = h
        self.last_delta_time = C.c_float(1.0)

    def close(self):
        if getattr(self, "_h", None):
            torch.cuda.synchronize()
            self._L.apbf_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- one method per subsystem apply()
    def velocity_handling(self, p: ParticleLists, dt: float, acceleration=None):
        acc = (C.c_float * 3)(*(acceleration if acceleration is not None else self.params["gravity"]))
        check(self._L.apbf_op_velocity_handling(self._h, C.byref(p.c()), C.byref(self.last_delta_time), float(dt),
                                                acc, _stream()), "apbf_op_velocity_handling")

    def particle_transfer(self, p: ParticleLists, t: TransferLists, dt: float):
        check(self._L.apbf_op_particle_transfer(self._h, C.byref(p.c()), C.byref(t.c()), float(dt), _stream()),
              "apbf_op_particle_transfer")

    def kernel_width_from_boundary_distance(self, p: ParticleLists):
        check(self._L.apbf_op_kernel_width_from_boundary_distance(self._h, C.byref(p.c()), _stream()),
              "apbf_op_kernel_width_from_boundary_distance")

    def neighborhood(self, p: ParticleLists, t: TransferLists | None, nb: NeighborLists, range_scale: float,
                     method: int = 0):
        mn = (C.c_float * 3)(*self.params["hash_min"])
        mx = (C.c_float * 3)(*self.params["hash_max"])
        tc = C.byref(t.c()) if t is not None else None
        check(self._L.apbf_op_neighborhood(self._h, C.byref(p.c()), tc, C.byref(nb.c()), int(method),
                                           float(range_scale), mn, mx, int(self.params["hash_res_log2"]), _stream()),
              "apbf_op_neighborhood")

    def spread_kernel_width(self, p: ParticleLists, nb: NeighborLists):
        check(self._L.apbf_op_spread_kernel_width(self._h, C.byref(p.c()), C.byref(nb.c()), _stream()),
              "apbf_op_spread_kernel_width")

    def box_collision(self, p: ParticleLists, boxes=None):
        boxes = self.params["boxes"] if boxes is None else boxes
        mn = (C.c_float * (4 * max(len(boxes), 1)))()
        mx = (C.c_float * (4 * max(len(boxes), 1)))()
        for i, (lo, hi) in enumerate(boxes):
            for d in range(3):
                mn[4 * i + d], mx[4 * i + d] = lo[d], hi[d]
        check(self._L.apbf_op_box_collision(self._h, C.byref(p.c()), len(boxes), mn, mx, _stream()),
              "apbf_op_box_collision")

    def sphere_collision(self, p: ParticleLists, center, radius: float, hollow: bool = True):
        c = (C.c_float * 3)(*center)
        check(self._L.apbf_op_sphere_collision(self._h, C.byref(p.c()), c, float(radius), int(hollow), _stream()),
              "apbf_op_sphere_collision")

    def incompressibility(self, p: ParticleLists, nb: NeighborLists):
        check(self._L.apbf_op_incompressibility(self._h, C.byref(p.c()), C.byref(nb.c()), _stream()),
              "apbf_op_incompressibility")

    def update_transfers(self, p: ParticleLists, t: TransferLists, nb: NeighborLists):
        check(self._L.apbf_op_update_transfers(self._h, C.byref(p.c()), C.byref(t.c()), C.byref(nb.c()), _stream()),
              "apbf_op_update_transfers")


def pool_update(ctx: Context, p: ParticleLists, t: TransferLists, nb: NeighborLists, dt: float, on_stage=None):
    """One substep on the lists in the order and with the flags of
    pool.cpp:67-106 (waterdrop.cpp:69-108). `on_stage(name)` is called after
    every stage (tests compare the lists there)."""
    prm = ctx.params
    s = prm["settings"]
    basic = bool(prm["basic_pbf"])
    hook = on_stage or (lambda name: None)
    ctx.velocity_handling(p, dt)
    hook("velocity")
    if not basic and (s["mMerge"] or s["mSplit"]):
        ctx.particle_transfer(p, t, dt)
        hook("particle_transfer")
    if not basic and s["mBaseKernelWidthOnBoundaryDistance"]:
        ctx.kernel_width_from_boundary_distance(p)
        hook("kernel_width_bd")
    spread = not basic and not s["mBaseKernelWidthOnBoundaryDistance"]
    ctx.neighborhood(p, t, nb, 1.5 if spread else 1.0)
    hook("neighborhood")
    if spread:
        ctx.spread_kernel_width(p, nb)
        hook("spread")
    for _ in range(prm["solver_iterations"]):
        ctx.box_collision(p)
        hook("box")
        if prm.get("sphere"):  # spherical_pool.cpp:86-89
            c, r, hollow = prm["sphere"]
            ctx.sphere_collision(p, c, r, hollow)
            hook("sphere")
        ctx.incompressibility(p, nb)
        hook("incompressibility")
    if not basic:
        ctx.update_transfers(p, t, nb)
        hook("update_transfers")
