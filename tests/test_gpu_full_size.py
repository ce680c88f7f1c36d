"""The benchmark's 1,000,137-particle adaptive water drop at full size: one
substep from its initial state against the oracle. Every list of the state
(positions, velocities, masses, radii, kernel widths, boundariness, boundary
distances, target radii, the transfers as a set and the 81 M ordered
neighbour pairs of the first substep) must match bit for bit. The oracle
takes 20-50 s of one core for the substep."""
import pytest

from apbf_amd import config as cfg
from oracle.oracle import OracleScene
from statecmp import compare

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_water_drop_full_size_substep_matches_oracle(device):
    from apbf_amd.scene import Scene
    p, st = cfg.water_drop()
    assert len(st["radius"]) == 1_000_137
    g = Scene(p)
    g.set_state(st)
    g.set_error_measurement(True)  # the benchmark's path: error pass fused into the split/merge find
    g.update(p["dt"])
    assert g.overflow_flags() == 0
    gs = g.get_state()
    g.close()
    r = OracleScene(p)
    r.set_state(st)
    r.update(p["dt"])
    rs = r.get_state()
    assert len(gs["pairs"]) > 50_000_000
    bad = compare(gs, rs, check_pairs=True)
    assert not bad, "; ".join(bad)
