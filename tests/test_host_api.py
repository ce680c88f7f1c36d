"""The C++ host interface (include/apbf.hpp): the subsystem classes run in the
order of pool.cpp:67-106 must reproduce scene::update bit for bit."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(__file__), "cpp", "host_api_check")


def _exe():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/host_api_check is not built (run __graft_entry__.build())")
    return EXE


def test_host_api_links_and_reports_missing_device():
    """Without a GPU the driver loads the library and exits with status 2."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = subprocess.run([_exe()], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "no device" in r.stdout


@pytest.mark.gpu
def test_host_api_subsystems_match_update(device):
    r = subprocess.run([_exe()], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "basic pbf: ok" in r.stdout and "adaptive: ok" in r.stdout and "spherical pool: ok" in r.stdout
    assert "time machine: ok" in r.stdout
    assert "time machine reference sequence: ok" in r.stdout  # source/test.cpp:585-619
    assert "lists basic pbf: ok" in r.stdout and "lists adaptive: ok" in r.stdout  # list-bound subsystems
    assert "gpu_list: ok" in r.stdout
