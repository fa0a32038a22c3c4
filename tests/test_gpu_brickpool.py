"""Brick pool manager on the GPU (yx_brick_pool_*; CuBrickPoolManager::allocMap,
reaction/voxtree/brickpool/BrickPool.cpp:21-62, and UpdateBrickPool,
reaction/voxtree/density.cu:240-284) against the numpy restatement of its
policy, and compositing over a pool-managed volume against the same volume
packed on the host."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_random_operations_match_policy(yx, oracle_mod, cuda):
    torch = cuda
    rng = np.random.default_rng(21)
    g, brick, cap = 16, 4, 1500
    nb, vox = g ** 3, brick ** 3
    dev = torch.device("cuda", 0)
    table = torch.full((nb,), -1, dtype=torch.int32, device=dev)
    params = torch.zeros((cap * vox, 4), dtype=torch.float32, device=dev)
    pool = yx.composite.BrickPool(table, params, g * brick, brick)
    orc = oracle_mod.BrickPoolOracle(np.full(nb, -1, np.int32), np.zeros((cap * vox, 4), np.float32), brick)
    for step in range(60):
        op = rng.choice(["alloc", "alloc", "release", "update"])
        k = int(rng.integers(1, 400))
        bricks = rng.integers(0, nb, size=k).astype(np.int32)
        bt = torch.from_numpy(bricks).to(dev)
        if op == "alloc":
            try:
                exp_slots, exp_new = orc.alloc(bricks)
                ok = True
            except ValueError:
                ok = False
            if ok:
                slots, new = pool.alloc(bt)
                torch.cuda.synchronize()
                assert new == exp_new, step
                assert np.array_equal(slots.cpu().numpy(), exp_slots), step
            else:
                with pytest.raises(yx.YoxelError):
                    pool.alloc(bt)
        elif op == "release":
            assert pool.release(bt) == orc.release(bricks), step
        else:
            packed = rng.standard_normal((k * vox, 4)).astype(np.float32)
            pool.update(bt, torch.from_numpy(packed).to(dev))
            orc.update(bricks, packed)
        torch.cuda.synchronize()
        assert np.array_equal(table.cpu().numpy(), orc.table), (step, op)
        assert pool.stats() == (int(orc.used.sum()), cap), (step, op)
    assert np.array_equal(params.cpu().numpy(), orc.params)
    # an index outside the table fails and changes nothing
    before = table.clone()
    with pytest.raises(yx.YoxelError):
        pool.alloc(torch.tensor([3, nb], dtype=torch.int32, device=dev))
    assert torch.equal(before, table)


def test_existing_table_and_bad_table(yx, cuda):
    torch = cuda
    vol = yx.composite.Volume(64, 8, occupancy=0.2, seed=3)
    table, params = vol.to(torch.device("cuda", 0))
    pool = yx.composite.BrickPool(table, params, 64, 8)
    assert pool.stats() == (vol.num_slots, vol.num_slots)
    bad = table.clone()
    live = torch.nonzero(bad >= 0).flatten()
    bad[live[1]] = bad[live[0]]  # one slot for two bricks
    with pytest.raises(yx.YoxelError):
        yx.composite.BrickPool(bad, params, 64, 8)


def test_compositing_over_pool_equals_packed_volume(yx, cuda):
    torch = cuda
    dev = torch.device("cuda", 0)
    grid, brick = 128, 8
    vol = yx.composite.Volume(grid, brick, occupancy=0.1, seed=5)
    table_ref, params_ref = vol.to(dev)
    vox = brick ** 3
    live = np.nonzero(vol.table >= 0)[0].astype(np.int32)
    rng = np.random.default_rng(4)
    order = rng.permutation(live)
    # the same bricks allocated in another order in a pool with spare slots
    cap = vol.num_slots + 37
    table = torch.full_like(table_ref, -1)
    params = torch.zeros((cap * vox, 4), dtype=torch.float32, device=dev)
    pool = yx.composite.BrickPool(table, params, grid, brick)
    bt = torch.from_numpy(order).to(dev)
    slots, new = pool.alloc(bt)
    assert new == live.size
    src = vol.params.reshape(-1, vox, 4)[vol.table[order]]
    pool.update(bt, torch.from_numpy(np.ascontiguousarray(src.reshape(-1, 4))).to(dev))
    comp = yx.composite.Compositor()
    views = torch.from_numpy(yx.composite.orbit_views(3, 60.0, 96, 80)).to(dev)
    a = comp.forward(table_ref, params_ref, grid, brick, views, 96, 80)
    b = comp.forward(table, params, grid, brick, views, 96, 80)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    # free a tenth of the bricks in both; the renders still agree
    drop = order[: live.size // 10]
    pool.release(torch.from_numpy(drop).to(dev))
    t2 = table_ref.clone()
    t2[torch.from_numpy(drop.astype(np.int64)).to(dev)] = -1
    a = comp.forward(t2, params_ref, grid, brick, views, 96, 80)
    b = comp.forward(table, params, grid, brick, views, 96, 80)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    assert pool.stats()[0] == live.size - drop.size
