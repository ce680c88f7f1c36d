"""The reference's time machine (source/time_machine.h:9-143) for one scene:
keyframes of the scene state every `keyframe_interval` substeps, kept in HBM
as device snapshots (apbf_snapshot_*), at most `max_keyframes` of them, and
the time step of every substep on the host. Stepping in the past re-simulates
from the last keyframe with the recorded time steps (the library's substeps
are deterministic, so the replay reproduces the recorded states bit for bit)
and loads the saved keyframe when it lands on one.

Use as the reference scenes do (pool.cpp:67-106)::

    dt = tm.step_forward(dt)        # None: keyframe loaded, skip the substep
    if dt is not None:
        scene.update(dt)
        tm.save_state(dt)
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import check

_MASK = 0xFFFFFFFF


class TimeMachine:
    def __init__(self, scene, max_keyframes: int = 5, keyframe_interval: int = 120):
        self._L = _lib.lib()
        self._scene = scene
        self.max_keyframes = max_keyframes
        self.keyframe_interval = keyframe_interval
        self.current_step = 0
        self.present_step = 0
        self.enabled = False
        self._keys: list = []   # apbf_snapshot handles, oldest first
        self._spare: list = []  # popped handles, reused by later keyframes
        self._dts: list = []    # time step of every substep in the history (non-sparse)
        self._dt = 1.0

    # ---- counters (time_machine.h:11-37)
    def on_keyframe(self) -> bool:
        return self.current_step % self.keyframe_interval == 0

    def in_past(self) -> bool:
        return self.current_step != self.present_step

    def keyframe_step_for(self, step: int) -> int:
        return step // self.keyframe_interval * self.keyframe_interval

    def oldest_history_keyframe_step(self) -> int:  # unsigned arithmetic, as in the reference
        return (self.keyframe_step_for(self.present_step) - (self.max_keyframes - 1) * self.keyframe_interval) & _MASK

    def history_index(self) -> int:
        return min(self.current_step, (self.current_step - self.oldest_history_keyframe_step()) & _MASK)

    def sparse_history_index(self) -> int:
        return self.history_index() // self.keyframe_interval

    def current_dt(self) -> float:
        """The time step recorded for the current step: the per-substep
        (non-sparse) history value that stepping through the past restores."""
        return self._dts[self.history_index()]

    # ---- snapshots
    def _take(self):
        if self._spare:
            h = self._spare.pop()
            check(self._L.apbf_snapshot_save(h, self._scene._h), "apbf_snapshot_save")
            return h
        h = C.c_void_p()
        check(self._L.apbf_snapshot_create(self._scene._h, C.byref(h)), "apbf_snapshot_create")
        return h

    def _drop_front(self, keep: int) -> None:
        while len(self._keys) > keep:
            self._spare.append(self._keys.pop(0))

    def _load(self) -> None:
        check(self._L.apbf_snapshot_restore(self._keys[self.sparse_history_index()], self._scene._h),
              "apbf_snapshot_restore")

    # ---- control (time_machine.h:104-136)
    def set_keyframe_interval(self, interval: int) -> "TimeMachine":
        self.keyframe_interval = interval
        self._drop_front(0)
        self._dts.clear()
        return self

    def enable(self, dt: float = 1.0) -> "TimeMachine":
        if not self.enabled:
            self._keys.append(self._take())
            self._dts.append(dt)
        self.enabled = True
        return self

    def disable(self) -> "TimeMachine":
        self._drop_front(0)
        self._dts.clear()
        self.enabled = False
        self.current_step = self.present_step = 0
        return self

    def step_forward(self, dt: float):
        """The time step to simulate with (the recorded one in the past), or
        None when a keyframe was loaded and the substep must be skipped."""
        if not self.enabled:
            return dt
        past = self.in_past()
        self.current_step += 1
        if not past:
            self.present_step += 1
        loaded = past and self.on_keyframe()
        if past:
            dt = self._dts[self.history_index()]
            if loaded:
                self._load()
        elif self.on_keyframe():
            self._drop_front(self.max_keyframes - 1)
            del self._dts[:max(0, len(self._dts) - (self.max_keyframes - 1) * self.keyframe_interval)]
        return None if loaded else dt

    def save_state(self, dt: float) -> None:
        if not self.enabled:
            return
        if len(self._keys) <= self.sparse_history_index():
            self._keys.append(self._take())
        if len(self._dts) <= self.history_index():
            self._dts.append(dt)

    def jump_back(self) -> bool:
        """Back to the previous keyframe (False when none is left)."""
        if (self.on_keyframe() and self.current_step + self.keyframe_interval * (self.max_keyframes - 1)
                <= self.present_step) or self.current_step == 0:
            return False
        self.current_step = (self.current_step - 1) // self.keyframe_interval * self.keyframe_interval
        self._load()
        return True

    def close(self) -> None:
        for h in self._keys + self._spare:
            self._L.apbf_snapshot_destroy(h)
        self._keys, self._spare = [], []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
