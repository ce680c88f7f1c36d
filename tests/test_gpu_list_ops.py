"""The stage operators over caller-owned lists (apbf_op_*, apbf_amd/lists.py)
against the oracle: one substep in the order of pool.cpp:67-106, the lists
compared bit for bit with the oracle's state after every stage (particles,
transfers as sets, and the neighbour list in the reference's uvec2 layout
through apbf_neighbor_pairs)."""
import numpy as np
import pytest

from apbf_amd import config as cfg
from statecmp import compare

pytestmark = pytest.mark.gpu


def _lists_state(P, T, NB):
    st = P.state()
    st.update(T.state())
    st["pairs"] = NB.pairs()
    return st


CASES = {
    "dam_break": lambda: cfg.dam_break(n_side=10),
    "water_drop_active": lambda: cfg.small_water_drop(seed=6, active=True),
    "pool_scene": lambda: cfg.pool_scene(),
    "spherical_pool": lambda: cfg.spherical_pool_scene(max_particles=6000),
}


@pytest.mark.parametrize("kind", sorted(CASES))
def test_list_operators_match_the_oracle_stage_by_stage(device, kind):
    from apbf_amd.lists import Context, NeighborLists, ParticleLists, TransferLists, pool_update
    from oracle.oracle import OracleScene
    p, st = CASES[kind]()
    dt = p["dt"]
    ctx = Context(p)
    ctx.last_delta_time.value = st["last_delta_time"]
    P = ParticleLists.from_state(st, p["cap_particles"], device)
    T = TransferLists(p["cap_transfers"], device, st)
    NB = NeighborLists(p["cap_particles"], p["cap_pairs"], device)
    ref = OracleScene(p)
    ref.set_state(st)
    seen = []
    for step in range(3):
        def hook(name):
            ref.stage(name, dt if name in ("velocity", "particle_transfer") else 0.0)
            bad = compare(_lists_state(P, T, NB), ref.get_state(), check_pairs=True)
            assert not bad, f"substep {step}, after {name}: " + "; ".join(bad)
            seen.append(name)
        pool_update(ctx, P, T, NB, dt, on_stage=hook)
    assert "neighborhood" in seen and "incompressibility" in seen
    if kind == "water_drop_active":
        assert {"particle_transfer", "spread", "update_transfers"} <= set(seen)
    ctx.close()


def test_list_pool_update_equals_the_scene(device):
    """Several substeps through the list operators equal the scene's fused
    update bit for bit (the scene fuses stages; the operators run them apart)."""
    from apbf_amd.lists import Context, NeighborLists, ParticleLists, TransferLists, pool_update
    from apbf_amd.scene import Scene
    p, st = cfg.small_water_drop(seed=7, active=True)
    dt = p["dt"]
    ctx = Context(p)
    ctx.last_delta_time.value = st["last_delta_time"]
    P = ParticleLists.from_state(st, p["cap_particles"], device)
    T = TransferLists(p["cap_transfers"], device, st)
    NB = NeighborLists(p["cap_particles"], p["cap_pairs"], device)
    sc = Scene(p)
    sc.set_state(st)
    for k in range(6):
        pool_update(ctx, P, T, NB, dt)
        sc.update(dt)
        bad = compare(_lists_state(P, T, NB), sc.get_state(), check_pairs=True)
        assert not bad, f"substep {k}: " + "; ".join(bad)
    ctx.close()
    sc.close()


def test_neighbor_rows_from_pairs_round_trip(device):
    """apbf_neighbor_rows_from_pairs rebuilds the CSR rows from the uvec2
    pairs in any order (an atomic-order list), keeping each particle's order."""
    import ctypes as C
    import torch
    from apbf_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(3)
    rows = 500
    counts = rng.integers(0, 9, rows)
    first = np.repeat(np.arange(rows), counts).astype(np.uint32)
    second = rng.integers(0, rows, len(first)).astype(np.uint32)
    pairs = np.stack([first, second], 1)
    shuffled = pairs[rng.permutation(len(pairs))]
    np_ = len(pairs)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).to(device)
    d_pairs = dev(shuffled.reshape(-1))
    d_np = torch.tensor([np_], dtype=torch.int32, device=device)
    row = torch.zeros(rows + 1, dtype=torch.int32, device=device)
    col = torch.zeros(np_, dtype=torch.int32, device=device)
    ws = torch.zeros(L.apbf_neighbor_rows_from_pairs_workspace_size(np_, rows), dtype=torch.uint8, device=device)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.apbf_neighbor_rows_from_pairs(C.c_void_p(d_pairs.data_ptr()), C.c_void_p(d_np.data_ptr()), np_, rows,
                                           C.c_void_p(row.data_ptr()), C.c_void_p(col.data_ptr()),
                                           C.c_void_p(ws.data_ptr()), ws.numel(), s) == 0
    torch.cuda.synchronize()
    r = row.cpu().numpy().view(np.uint32)
    c = col.cpu().numpy().view(np.uint32)
    assert np.array_equal(r, np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32))
    for i in range(rows):  # per row, the shuffled list's order
        assert np.array_equal(c[r[i]:r[i + 1]], shuffled[shuffled[:, 0] == i, 1])
    # and back to pairs: the rows in row order
    out = torch.zeros(2 * np_, dtype=torch.int32, device=device)
    d_rows = torch.tensor([rows], dtype=torch.int32, device=device)
    assert L.apbf_neighbor_pairs(C.c_void_p(row.data_ptr()), C.c_void_p(col.data_ptr()),
                                 C.c_void_p(d_rows.data_ptr()), C.c_void_p(d_np.data_ptr()), rows, np_,
                                 C.c_void_p(out.data_ptr()), s) == 0
    torch.cuda.synchronize()
    back = out.cpu().numpy().view(np.uint32).reshape(-1, 2)
    assert np.array_equal(back[:, 0], first)
