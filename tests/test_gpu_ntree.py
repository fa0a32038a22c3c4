"""GPU parity of the brick n-tree: the GPU builder's arrays equal the CPU
restatement's byte for byte (from a dense volume and from the fBm), and the
DDA cast equals the restatement's per-pixel outputs exactly (RGBA8, t, first
voxel, node and voxel read counts), and a cast without the read counters
equals the counted one; at the full brick_4096 size the build is checked on
sampled bricks and the cast on every 8th row of a 1080p frame."""
import numpy as np
import pytest

from test_ntree_oracle import palette as make_palette

pytestmark = pytest.mark.gpu

KEYS = ("rgba", "t", "voxel", "nodes", "voxels")
DT = {"rgba": np.uint32, "t": np.float32, "voxel": np.int32, "nodes": np.uint32, "voxels": np.uint32}


def gpu_cast(torch, tree, view, w, h, pal):
    dpal = torch.from_numpy(pal.view(np.int32)).cuda()
    out = tree.render_tensors(view, w, h, dpal, outputs=KEYS)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy().view(DT[k]) for k, v in out.items()}


def assert_cast_equal(g, o, rows=None, what=""):
    for k in KEYS:
        a, b = g[k], o[k]
        if rows is not None:
            a, b = a[rows], b[rows]
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError(f"{what}: {k} differs at {len(bad)} elements, first {bad[:3].tolist()}: "
                                 f"gpu {a[tuple(bad[0])]} cpu {b[tuple(bad[0])]}")


def assert_uncounted_equal(torch, tree, view, w, h, pal, counted, what=""):
    """A cast without the read counters (the kernel variant that leaves them
    out) gives the same RGBA8, t and first voxel as the counted cast."""
    dpal = torch.from_numpy(pal.view(np.int32)).cuda()
    out = tree.render_tensors(view, w, h, dpal, outputs=("rgba", "t", "voxel"))
    torch.cuda.synchronize()
    for k, v in out.items():
        a = v.cpu().numpy().view(DT[k])
        assert np.array_equal(a, counted[k]), f"{what}: {k} differs without the read counters"


def views_for(yx, w, h, seed=0):
    cams = yx.workloads.orbit_cameras(3, seed=seed) + [((0.41, 0.37, 0.55), (0.4, 1.0, -0.3)),
                                                         ((0.5, 0.5, 2.5), (0.0, 1e-3, -1.0))]
    return yx.composite.camera_views(cams, 60.0, w, h)


@pytest.mark.parametrize("size,levels", [(64, 1), (128, 2), (32, 0), (128, 1), (256, 2)])
def test_dense_build_and_cast(yx, oracle_mod, cuda, size, levels):
    o = oracle_mod.NTreeOracle()
    d = o.desc(size, levels, 4, 4.0, 0.4, 300.0, size * 7 + levels)
    dense = o.density_box(d, 0, 0, 0, size, size, size)
    root, nodes, bricks = o.build_dense(dense, levels)
    t = yx.BrickTree.from_dense(dense, levels)
    info = t.info()
    assert (info["num_nodes"], info["num_bricks"]) == (len(nodes), len(bricks))
    g_root, g_nodes, g_bricks = t.download()
    assert np.array_equal(g_root, root) and np.array_equal(g_nodes, nodes) and np.array_equal(g_bricks, bricks)
    pal = make_palette(size)
    w, h = 67, 45
    launches = [(True, 16), (False, 16), (True, 1), (True, 64)]
    for i, view in enumerate(views_for(yx, w, h, seed=size)):
        c = o.render(root, nodes, bricks, size, levels, view, pal, w, h)
        for persistent, refill in launches:
            print(f"n-tree {size}/{levels} view {i} persistent {persistent} refill {refill}", flush=True)
            t.set_launch(persistent, refill)
            g = gpu_cast(cuda, t, view, w, h, pal)
            assert_cast_equal(g, c, what=f"size {size} levels {levels} view {i} launch {persistent}/{refill}")
        t.set_launch(False, 16)
        assert_uncounted_equal(cuda, t, view, w, h, pal, g, what=f"size {size} levels {levels} view {i}")
        assert (g["t"] >= 0).any()


@pytest.mark.parametrize("size,levels", [(64, 1), (128, 2), (256, 2)])
def test_fbm_build_matches_restatement(yx, oracle_mod, cuda, size, levels):
    o = oracle_mod.NTreeOracle()
    desc = dict(size=size, levels=levels, octaves=5, frequency=8.0, threshold=0.45, scale=400.0, seed=5 + size)
    d = o.desc(**desc)
    dense = o.density_box(d, 0, 0, 0, size, size, size)
    root, nodes, bricks = o.build_dense(dense, levels)
    t = yx.BrickTree.build_fbm(**desc)
    g_root, g_nodes, g_bricks = t.download()
    assert np.array_equal(g_root, root) and np.array_equal(g_nodes, nodes) and np.array_equal(g_bricks, bricks)
    t2 = yx.BrickTree.build_fbm(**desc)  # deterministic
    assert all(np.array_equal(a, b) for a, b in zip(t2.download(), (g_root, g_nodes, g_bricks)))


def brick_index(root, nodes, size, levels, bx, by, bz):
    R = (size // 8) >> (2 * levels)
    sh = 2 * levels
    idx = int(root[((bz >> sh) * R + (by >> sh)) * R + (bx >> sh)])
    for l in range(1, levels + 1):
        if idx == 0xFFFFFFFF:
            return None
        s = 2 * (levels - l)
        c = ((((bz >> s) & 3) * 4) + ((by >> s) & 3)) * 4 + ((bx >> s) & 3)
        n = nodes[idx]
        mask = int(n[0]) | (int(n[1]) << 32)
        if not (mask >> c) & 1:
            return None
        idx = int(n[2]) + bin(mask & ((1 << c) - 1)).count("1")
    return None if idx == 0xFFFFFFFF else idx


def test_brick_4096_full(yx, oracle_mod, cuda):
    """The full workload: 4096^3, 8^3 root grid, 3 node levels, ~3 % occupancy.
    Sampled bricks (allocated and empty) against the restatement's densities;
    a 1920x1080 cast against the restatement on every 8th row."""
    torch = cuda
    o = oracle_mod.NTreeOracle()
    cfg = yx.workloads.BRICK_4096
    print("building brick_4096", flush=True)
    t = yx.BrickTree.build_fbm(**cfg)
    info = t.info()
    total = (cfg["size"] // 8) ** 3
    print(f"brick_4096: {info}", flush=True)
    assert 0.02 * total < info["num_bricks"] < 0.06 * total
    root, nodes, bricks = t.download()
    d = o.desc(**cfg)
    rng = np.random.default_rng(0)
    nb = cfg["size"] // 8
    # bricks at random positions, and random allocated bricks found by walking the tree
    found = 0
    for k in range(400):
        bx, by, bz = (int(v) for v in rng.integers(0, nb, 3))
        ref = o.density_box(d, bx * 8, by * 8, bz * 8, 8, 8, 8)
        idx = brick_index(root, nodes, cfg["size"], cfg["levels"], bx, by, bz)
        assert (idx is not None) == bool(ref.any()), (bx, by, bz)
        if idx is not None:
            assert np.array_equal(bricks[idx].reshape(8, 8, 8), ref)
            found += 1
    assert found > 0
    w, h = 1920, 1080
    pal = yx.workloads.brick_palette()
    view = yx.composite.camera_views(yx.workloads.orbit_cameras(4, seed=2)[:1], 60.0, w, h)[0]
    g = gpu_cast(torch, t, view, w, h, pal)
    c = o.render(root, nodes, bricks, cfg["size"], cfg["levels"], view, pal, w, h, ystep=8, threads=16)
    rows = np.arange(0, h, 8)
    assert_cast_equal(g, c, rows=rows, what="brick_4096 1080p")
    assert_uncounted_equal(torch, t, view, w, h, pal, g, what="brick_4096 1080p")
    hit = (g["t"] >= 0).mean()
    print(f"brick_4096 frame: {hit:.3f} of the rays reach a non-empty voxel", flush=True)
    assert 0.05 < hit


def test_empty_and_edge_shapes(yx, oracle_mod, cuda):
    """An all-empty volume (no bricks: every ray misses), the largest root grid
    (16^3, 512^3 volume with one node level) from a sparse random volume, and a
    four-level tree (2048^3 from the fBm) checked on sampled bricks and rows."""
    o = oracle_mod.NTreeOracle()
    pal = make_palette(9)
    w, h = 48, 40
    # empty
    t = yx.BrickTree.from_dense(np.zeros((64, 64, 64), np.uint8), 1)
    info = t.info()
    assert info["num_bricks"] == 0 and info["num_nodes"] == 0
    for view in views_for(yx, w, h, seed=3):
        g = gpu_cast(cuda, t, view, w, h, pal)
        assert np.all(g["rgba"] == 0) and np.all(g["t"] == -1) and np.all(g["voxel"] == -1)
        assert np.all(g["nodes"] == 0) and np.all(g["voxels"] == 0)
    # 16^3 root grid: 512^3 with one level of 4^3 nodes, sparse random blobs
    rng = np.random.default_rng(5)
    dense = np.zeros((512, 512, 512), np.uint8)
    for _ in range(300):
        c = rng.integers(0, 500, 3)
        s = rng.integers(2, 12, 3)
        dense[c[2]:c[2] + s[2], c[1]:c[1] + s[1], c[0]:c[0] + s[0]] = rng.integers(1, 256)
    root, nodes, bricks = o.build_dense(dense, 1)
    assert root.size == 16 ** 3
    t = yx.BrickTree.from_dense(dense, 1)
    g_root, g_nodes, g_bricks = t.download()
    assert np.array_equal(g_root, root) and np.array_equal(g_nodes, nodes) and np.array_equal(g_bricks, bricks)
    for i, view in enumerate(views_for(yx, w, h, seed=11)):
        g = gpu_cast(cuda, t, view, w, h, pal)
        c = o.render(root, nodes, bricks, 512, 1, view, pal, w, h)
        assert_cast_equal(g, c, what=f"root 16^3 view {i}")
    # four node levels: 2048 = 8 * 4^4 * 1
    cfg = dict(size=2048, levels=4, octaves=4, frequency=6.0, threshold=0.55, scale=300.0, seed=19)
    t = yx.BrickTree.build_fbm(**cfg)
    root, nodes, bricks = t.download()
    assert root.size == 1
    d = o.desc(**cfg)
    for k in range(200):
        bx, by, bz = (int(v) for v in rng.integers(0, 256, 3))
        ref = o.density_box(d, bx * 8, by * 8, bz * 8, 8, 8, 8)
        idx = brick_index(root, nodes, 2048, 4, bx, by, bz)
        assert (idx is not None) == bool(ref.any())
        if idx is not None:
            assert np.array_equal(bricks[idx].reshape(8, 8, 8), ref)
    W, H = 320, 200
    view = yx.composite.camera_views(yx.workloads.orbit_cameras(2, seed=4)[:1], 60.0, W, H)[0]
    g = gpu_cast(cuda, t, view, W, H, pal)
    c = o.render(root, nodes, bricks, 2048, 4, view, pal, W, H, ystep=4, threads=16)
    assert_cast_equal(g, c, rows=np.arange(0, H, 4), what="four levels")
    assert (g["t"] >= 0).any()
