"""The 8M-particle large pool (plain PBF, 200 x 100 x 100 h box, at rest
density, wave impulse) at full size. After two substeps the neighbourhood stage runs on the moved
particles; its particle order must be sorted by the oracle's position hash,
and its ordered neighbour lists must equal the oracle's neighborhood_green
over the same positions, cell tables and kernel widths, pair for pair."""
import numpy as np
import pytest

from apbf_amd import config as cfg
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_large_pool_neighbourhood_full_size(device):
    from apbf_amd.scene import Scene
    p, st = cfg.large_pool()
    n = len(st["radius"])
    assert n == 386 * 107 * 193
    g = Scene(p)
    g.set_state(st)
    del st
    for _ in range(2):
        g.update(p["dt"])
    g.stage("neighborhood", p["dt"])
    assert g.overflow_flags() == 0
    gs = g.get_state()
    g.close()
    pos = gs["position"]
    assert len(pos) == n
    assert np.all(np.isfinite(gs["velocity"]))

    # the walls keep every particle inside the box (pool.cpp:30-40)
    w = pos[:, :3].astype(np.float64) / cfg.POS_RESOLUTION
    assert w.min() > -1.0 and np.all(w.max(0) < np.array([201.0, 101.0, 101.0]))

    mn, mx, res = p["hash_min"], p["hash_max"], p["hash_res_log2"]
    hashes = orc.position_hash(pos, mn, mx, res)
    assert np.all(hashes[1:] >= hashes[:-1]), "particles are not in hash order"
    cs, ce = orc.find_value_ranges(hashes, 1 << (3 * res))
    del hashes
    pairs = gs["pairs"]
    assert len(pairs) > 25 * n
    ref = orc.neighborhood_green(pos, gs["kernel_width"], cs, ce, 1.0, mn, mx, res, len(pairs) + 1)
    assert len(ref) == len(pairs)
    rows = np.flatnonzero(np.any(ref != pairs, axis=1))
    assert rows.size == 0, f"{rows.size} pairs differ, first at {rows[:4]}: {pairs[rows[:4]]} vs {ref[rows[:4]]}"
