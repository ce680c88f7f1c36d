"""The time machine (time_machine.h:9-143) over device snapshots: jumping back
restores the keyframe state exactly, and stepping forward through the past
(re-simulating with the recorded time steps, loading the later keyframes)
reproduces every recorded substep bit for bit, because the substeps are
deterministic."""
import numpy as np
import pytest

from apbf_amd import config as cfg
from statecmp import compare

pytestmark = pytest.mark.gpu


def _state(s):
    st = s.get_state(pairs=False)
    st["pairs"] = np.zeros((0, 2), np.uint32)
    return st


def _same(a, b, where):
    bad = compare(a, b, check_pairs=False)
    assert not bad, f"{where}: " + "; ".join(bad)


def test_time_machine_replays_exactly(device):
    from apbf_amd.scene import Scene
    from apbf_amd.time_machine import TimeMachine
    p, st = cfg.small_water_drop(seed=3, active=True)
    s = Scene(p)
    s.set_state(st)
    tm = TimeMachine(s, max_keyframes=3, keyframe_interval=3).enable(1.0)
    base = p["dt"]
    dts = [base * (1.0 + 0.25 * (k % 3)) for k in range(12)]
    rec = {0: _state(s)}
    for k in range(1, 11):  # ten substeps in the present
        dt = tm.step_forward(dts[k])
        assert dt == dts[k]
        s.update(dt)
        tm.save_state(dt)
        rec[k] = _state(s)
    assert rec[10]["t_source"].size or rec[5]["t_source"].size  # transfers are part of the state

    assert tm.jump_back() and tm.current_step == 9
    _same(_state(s), rec[9], "jump to keyframe 9")
    assert tm.jump_back() and tm.current_step == 6
    assert tm.jump_back() and tm.current_step == 3
    _same(_state(s), rec[3], "jump to keyframe 3")
    assert not tm.jump_back()  # keyframe 3 is the oldest one kept

    for k in range(4, 11):  # forward through the past
        dt = tm.step_forward(base)  # the argument is ignored in the past
        if k % 3 == 0:
            assert dt is None  # keyframe loaded, no substep
        else:
            assert dt == dts[k]
            s.update(dt)
            tm.save_state(dt)
        _same(_state(s), rec[k], f"replayed substep {k}")
    assert not tm.in_past()
    dt = tm.step_forward(dts[11])
    assert dt == dts[11]
    s.update(dt)
    tm.save_state(dt)
    tm.close()
    s.close()


def test_snapshot_restore_is_exact(device):
    """A snapshot taken mid-run and restored later puts back every list; the
    next substeps then equal those after the snapshot."""
    from apbf_amd.scene import Scene
    from apbf_amd.time_machine import TimeMachine
    p, st = cfg.small_water_drop(seed=5)
    s = Scene(p)
    s.set_state(st)
    for _ in range(2):
        s.update(p["dt"])
    tm = TimeMachine(s, max_keyframes=2, keyframe_interval=100).enable(p["dt"])
    a = [_state(s)]
    for _ in range(3):
        s.update(p["dt"])
        a.append(_state(s))
    assert tm.jump_back() is False  # on keyframe 0 with nothing older: the reference refuses too
    tm._load()
    _same(_state(s), a[0], "restored")
    assert s.counts()[2] == 0  # the neighbour list is not part of a keyframe
    for k in range(1, 4):
        s.update(p["dt"])
        _same(_state(s), a[k], f"substep {k} after the restore")
    tm.close()
    s.close()


def test_time_machine_reference_sequence(device):
    """The reference's own time-machine test (source/test.cpp:585-619): two
    keyframes at most, one every 3 steps; three steps in the present, then
    jump_back and four step_forward calls with the expected returns. The
    reference's sparse list is the scene here, and its per-step value `i` is
    the recorded time step."""
    from apbf_amd.scene import Scene
    from apbf_amd.time_machine import TimeMachine
    p, st = cfg.dam_break(n_side=6)
    s = Scene(p)
    s.set_state(st)
    tm = TimeMachine(s, max_keyframes=2, keyframe_interval=3).enable(0.0)
    rec = {0: _state(s)}
    dts = {k: p["dt"] * (1.0 + 0.1 * k) for k in range(1, 5)}
    for k in (1, 2, 3):
        assert tm.step_forward(dts[k]) is not None  # no keyframe load in the present
        s.update(dts[k])
        tm.save_state(dts[k])
        rec[k] = _state(s)
    assert tm.jump_back()
    assert tm.current_step == 0 and tm.current_dt() == 0.0
    _same(_state(s), rec[0], "jump_back to keyframe 0")
    for k in (1, 2):
        dt = tm.step_forward(99.0)
        assert dt == dts[k] and tm.current_dt() == dts[k]
        s.update(dt)
        tm.save_state(dt)
        _same(_state(s), rec[k], f"past step {k}")
    assert tm.step_forward(99.0) is None  # lands on keyframe 3: loaded
    assert tm.current_dt() == dts[3]
    tm.save_state(dts[3])
    _same(_state(s), rec[3], "keyframe 3 loaded")
    assert tm.step_forward(dts[4]) == dts[4] and not tm.in_past()  # the present again
    tm.save_state(dts[4])
    tm.close()
