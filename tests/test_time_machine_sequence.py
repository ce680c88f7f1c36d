"""The time machine's step bookkeeping (apbf_amd/time_machine.py, after
source/time_machine.h:9-143) on the reference's own test sequence
(source/test.cpp:585-619), with the device snapshots replaced by copies of a
host list, so it runs without a GPU. The GPU version of the same sequence
over a scene is tests/test_gpu_time_machine.py."""
from apbf_amd.time_machine import TimeMachine


class ListTimeMachine(TimeMachine):
    """Keyframes of a host list (the reference test's gpu_list) instead of a scene."""

    def __init__(self, data: list, **kw):
        self._data = data
        self.loads = 0
        super().__init__(scene=None, **kw)

    def _take(self):
        return list(self._data)

    def _load(self):
        self.loads += 1
        self._data[:] = self._keys[self.sparse_history_index()]

    def close(self):
        self._keys, self._spare = [], []


def test_reference_time_machine_sequence():
    data = [0, 1, 2]
    tm = ListTimeMachine(data, max_keyframes=2, keyframe_interval=3).enable(0.0)
    for i in (1, 2, 3):
        assert tm.step_forward(float(i)) == float(i)       # !step_forward(): nothing loaded
        data[:] = [i, i + 1, i + 2]                         # write_sequence(list, 3, id, 1)
        tm.save_state(float(i))
    assert tm.jump_back()
    assert tm.current_dt() == 0.0 and data == [0, 1, 2]     # i == 0, keyframe 0 restored
    assert tm.step_forward(99.0) == 1.0                     # !step_forward(); i == 1
    tm.save_state(1.0)
    assert tm.step_forward(99.0) == 2.0                     # !step_forward(); i == 2
    tm.save_state(2.0)
    assert tm.step_forward(99.0) is None                    # step_forward(): keyframe loaded
    assert tm.current_dt() == 3.0 and data == [3, 4, 5]     # i == 3
    tm.save_state(3.0)
    assert tm.step_forward(4.0) == 4.0 and not tm.in_past()  # !step_forward(); the present
    tm.save_state(4.0)
    assert tm.loads == 2  # jump_back and the keyframe at step 3


def test_jump_back_stops_at_the_oldest_keyframe():
    data = [0]
    tm = ListTimeMachine(data, max_keyframes=2, keyframe_interval=3).enable(0.0)
    for i in range(1, 8):
        tm.step_forward(float(i))
        data[:] = [i]
        tm.save_state(float(i))
    # keyframes kept: steps 3 and 6 (two at most); the present is step 7
    assert tm.jump_back() and tm.current_step == 6 and data == [6]
    assert tm.jump_back() and tm.current_step == 3 and data == [3]
    assert not tm.jump_back()
