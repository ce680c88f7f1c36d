"""Packed batches (apbf_scene_create_packed, apbf_amd.scene.PackedBatch): the
64 scenes of the batched workload (10 k - 500 k particles each, about 8 M in
all, dam breaks and adaptive water drops) in two packed lists, every stage
one launch per list. After two substeps every scene must equal the same scene
stepped alone, bit for bit (particles, transfers as sets, neighbour lists in
order), and the four smallest scenes must equal the oracle's."""
import numpy as np
import pytest

from apbf_amd import config as cfg
from statecmp import compare

pytestmark = pytest.mark.gpu


def _packed_and_singles(cases, substeps):
    from apbf_amd.scene import PackedBatch, Scene
    dt = cases[0][0]["dt"]
    batch = PackedBatch(cases)
    for _ in range(substeps):
        batch.update(dt)
    got = batch.scene_states()
    flags = [g["scene"].overflow_flags() for g in batch.groups]
    batch.close()
    return got, flags


def test_packed_batch_equals_the_scenes_alone_full_size(device):
    from apbf_amd.scene import Scene
    cases = cfg.batch_scenes(64)
    dt = cases[0][0]["dt"]
    print(f"64 scenes, {sum(len(st['radius']) for _, st in cases)} particles", flush=True)
    got, flags = _packed_and_singles(cases, 2)
    assert flags == [0, 0]  # no shared capacity was reached
    for k, (p, st) in enumerate(cases):
        s = Scene(p)
        s.set_state(st)
        s.update(dt)
        s.update(dt)
        bad = compare(got[k], s.get_state(), check_pairs=True)
        s.close()
        assert not bad, f"scene {k} ({p['workload']}): " + "; ".join(bad)


def test_packed_batch_smallest_scenes_equal_the_oracle(device):
    from oracle.oracle import OracleScene
    cases = cfg.batch_scenes(64)
    dt = cases[0][0]["dt"]
    got, _ = _packed_and_singles(cases, 2)
    smallest = np.argsort([len(st["radius"]) for _, st in cases])[:4]
    kinds = set()
    for k in smallest:
        p, st = cases[k]
        r = OracleScene(p)
        r.set_state(st)
        r.update(dt)
        r.update(dt)
        bad = compare(got[k], r.get_state(), check_pairs=True)
        assert not bad, f"scene {k} ({p['workload']}): " + "; ".join(bad)
        kinds.add(int(p["basic_pbf"]))
    print(f"oracle-checked scenes {sorted(smallest.tolist())}, basic_pbf kinds {sorted(kinds)}", flush=True)


def test_packed_batch_small_mixed(device):
    """Eight small scenes of both kinds, four substeps, against the scenes alone."""
    from apbf_amd.scene import Scene
    cases = cfg.batch_scenes(8, n_min=2000, n_max=12000, seed=3)
    dt = cases[0][0]["dt"]
    got, flags = _packed_and_singles(cases, 4)
    assert flags == [0, 0]
    for k, (p, st) in enumerate(cases):
        s = Scene(p)
        s.set_state(st)
        for _ in range(4):
            s.update(dt)
        bad = compare(got[k], s.get_state(), check_pairs=True)
        s.close()
        assert not bad, f"scene {k}: " + "; ".join(bad)
