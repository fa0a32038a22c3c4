"""GPU parity: the HIP ray cast against the CPU restatement on the same scenes
and cameras, through the C ABI. Hit ids, voxel coordinates, t, RGBA8 and the
shadow flag are compared bit-exactly (the stated bar is bit-exact ids and
voxels, t within 1e-5 relative, colour within 1 LSB; the kernel repeats the
oracle's float operations exactly, so exact equality is asserted)."""
import numpy as np
import pytest

import svo_helpers as H

pytestmark = pytest.mark.gpu

KEYS = ("hit_id", "voxel", "t", "rgba", "shadow", "visits")


TUNINGS = [
    dict(kernel=0),
    dict(kernel=1),  # library default
    dict(kernel=2, window=0, block=512, tile_w=16, tile_h=4, refill_min=1, queues=1),
    dict(kernel=2, window=1, block=256, tile_w=64, tile_h=1, refill_min=64, queues=13),
    dict(kernel=2, window=2, block=512, tile_w=32, tile_h=32, refill_min=32, queues=64, waves_per_cu=4),
    dict(kernel=1, block=256, window=0, tile_w=1, tile_h=1, refill_min=8, queues=8, waves_per_cu=1),
    dict(kernel=1, block=128, window=2, tile_w=16, tile_h=16, refill_min=1, queues=3),
    dict(kernel=1, block=64, window=6, tile_w=32, tile_h=2, refill_min=64, queues=64, waves_per_cu=2),
    dict(kernel=1, block=256, window=4, tile_w=4, tile_h=16, refill_min=20, queues=2),
    # top tree levels staged in LDS (all levels asked: as many as fit)
    dict(kernel=1, lds_levels=3),
    dict(kernel=1, block=256, window=2, lds_levels=5, refill_min=16),
    dict(kernel=1, block=128, lds_levels=16, tile_w=16, tile_h=4),
]


def gpu_render(yx, torch, scene, pos, d, w, h, fov=60.0, light_mode=0, light=(0, 0, 1), keys=KEYS,
               tuning=None, renderer=None):
    r = renderer or yx.SVORenderer()
    if tuning:
        r.set_tuning(**tuning)
    r.SetScene(scene)
    r.SetViewPos(pos)
    r.SetViewDir(d)
    r.SetFOV(fov)
    r.SetResolution(w, h)
    r.SetLight(light, light_mode)
    out = r.render_tensors(keys)
    torch.cuda.synchronize()
    dt = {"hit_id": np.uint32, "voxel": np.int32, "t": np.float32, "rgba": np.uint32, "shadow": np.uint8,
          "visits": np.uint32}
    return {k: v.cpu().numpy().view(dt[k]) for k, v in out.items()}


def assert_same(g, o, keys=KEYS, what=""):
    for k in keys:
        a, b = g[k], o[k].reshape(g[k].shape)
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError(f"{what}: {k} differs at {len(bad)} elements, first {bad[:3].tolist()} "
                                 f"gpu={a[tuple(bad[0])]} oracle={b[tuple(bad[0])]}")


def test_plain_512_eight_views(yx, orc, cuda):
    """Config 'plain_512': depth 8, 64 spheres, 512x512, 8 orbit views,
    light at the eye (the reference CPU shader's setting)."""
    s = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    nodes, root = s.nodes(), s.GetRoot()
    for i, (pos, d) in enumerate(yx.workloads.orbit_cameras(8, seed=0)):
        print(f"plain_512 view {i}", flush=True)
        g = gpu_render(yx, cuda, s, pos, d, 512, 512)
        o = orc.render(nodes, root, pos, d, fov=60.0, w=512, h=512, threads=8)
        assert_same(g, o, what=f"view {i}")
        assert (g["hit_id"] != H.MISS).sum() > 10000


def test_tunings_render_identically(yx, orc, cuda):
    """Every launch shape of both kernels gives the oracle's outputs, on the
    plain_512 scene (light at the eye and sun + shadow) and on one full
    large_1080p frame; the queues reset themselves between launches."""
    small = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    large = yx.SVOData.build_synthetic(**yx.workloads.LARGE_1080P)
    cams = yx.workloads.orbit_cameras(8, seed=0)
    cases = [(small, cams[1], 512, 512, 0), (small, cams[6], 512, 512, 2),
             (large, yx.workloads.orbit_cameras(32, seed=0)[13], 1920, 1080, 2)]
    for s, (pos, d), w, h, mode in cases:
        o = orc.render(s.nodes(), s.GetRoot(), pos, d, fov=60.0, w=w, h=h, light_mode=mode,
                       light=yx.workloads.SUN, threads=16)
        r = yx.SVORenderer()
        for tu in TUNINGS:
            print(f"tuning {tu} on {w}x{h} mode {mode}", flush=True)
            for rep in range(2):
                g = gpu_render(yx, cuda, s, pos, d, w, h, light_mode=mode, light=yx.workloads.SUN, tuning=tu,
                               renderer=r)
                assert_same(g, o, what=f"tuning {tu} rep {rep}")


def test_plain_512_sun_shadow(yx, orc, cuda):
    s = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    nodes, root = s.nodes(), s.GetRoot()
    for i, (pos, d) in enumerate(yx.workloads.orbit_cameras(3, seed=5)):
        g = gpu_render(yx, cuda, s, pos, d, 512, 512, light_mode=2, light=yx.workloads.SUN)
        o = orc.render(nodes, root, pos, d, fov=60.0, w=512, h=512, light_mode=2, light=yx.workloads.SUN,
                       threads=8)
        assert_same(g, o, what=f"sun view {i}")
        assert (g["shadow"] == 1).any() and (g["shadow"] == 0).any()


def test_large_1080p_frames(yx, orc, cuda):
    """Config 'large_1080p' at full size: depth 11, 1920x1080 primary + shadow,
    4 of the 32 orbit frames compared pixel for pixel."""
    s = yx.SVOData.build_synthetic(**yx.workloads.LARGE_1080P)
    nodes, root = s.nodes(), s.GetRoot()
    cams = yx.workloads.orbit_cameras(32, seed=0)
    for i in (0, 9, 17, 26):
        pos, d = cams[i]
        print(f"large_1080p frame {i}", flush=True)
        g = gpu_render(yx, cuda, s, pos, d, 1920, 1080, light_mode=2, light=yx.workloads.SUN)
        o = orc.render(nodes, root, pos, d, fov=60.0, w=1920, h=1080, light_mode=2, light=yx.workloads.SUN,
                       threads=16)
        assert_same(g, o, what=f"frame {i}")


def test_edge_cases(yx, orc, cuda):
    rng = np.random.default_rng(1)
    # mixed-level leaves, shuffled node ids (exercises the canonical id map)
    occ = rng.random((16, 16, 16)) < 0.1
    leaves = H.dense_leaves(occ, rng) + [(2, 0, 0, 0, 77), (1, 1, 1, 1, 99)]
    leaves = [l for l in leaves if not (l[0] == 4 and l[1] < 4 and l[2] < 4 and l[3] < 4)]
    leaves = [l for l in leaves if not (l[0] == 4 and l[1] >= 8 and l[2] >= 8 and l[3] >= 8)]
    nodes, root = H.tree_from_leaves(leaves, order="shuffle", seed=3)
    s = yx.SVOData.from_nodes(nodes, root)
    cases = [
        ((2.0, 1.7, 1.4), (-1.0, -0.8, -0.6), 37, 29),      # ragged resolution
        ((0.3, 0.45, 0.6), (1.0, 0.2, -0.1), 64, 64),       # eye inside the cube
        ((0.5, 0.5, 3.0), (0.0, 0.0, -1.0), 33, 31),        # axis-aligned view, AdjustDir
        ((-1.0, 0.5, 0.5), (1.0, 0.0, 0.0), 1, 1),          # single pixel
    ]
    for pos, d, w, h in cases:
        for mode in (0, 2):
            o = orc.render(nodes, root, pos, d, fov=70.0, w=w, h=h, light_mode=mode, light=yx.workloads.SUN)
            for tu in (TUNINGS[0], TUNINGS[1], TUNINGS[4]):
                g = gpu_render(yx, cuda, s, pos, d, w, h, fov=70.0, light_mode=mode, light=yx.workloads.SUN,
                               tuning=tu)
                assert_same(g, o, what=f"edge {pos} {w}x{h} mode {mode} tuning {tu}")


def test_deep_and_tiny_trees(yx, orc, cuda):
    # single leaf at level 16 (maximum depth), and a depth-1 tree
    for leaves in ([(16, 40000, 20000, 30000, 5)], [(1, 1, 0, 1, 9), (1, 0, 1, 0, 3)]):
        nodes, root = H.tree_from_leaves(leaves)
        s = yx.SVOData.from_nodes(nodes, root)
        lv, x, y, z, _ = leaves[0]
        c = (np.array([x, y, z]) + 0.5) / 2.0 ** lv
        pos = (float(c[0]) - 1.0, float(c[1]) - 0.3, float(c[2]) + 0.2)
        d = (1.0, 0.3, -0.2)
        o = orc.render(nodes, root, pos, d, fov=0.01 if lv == 16 else 40.0, w=64, h=64)
        for tu in (TUNINGS[0], TUNINGS[1], TUNINGS[2]):
            g = gpu_render(yx, cuda, s, pos, d, 64, 64, fov=0.01 if lv == 16 else 40.0, tuning=tu)
            assert_same(g, o, what=f"tree level {lv} tuning {tu}")
            assert (g["hit_id"] != H.MISS).any()


def test_empty_node_chain_below_leaves(yx, orc, cuda):
    """A node chain with no leaves reaching below the deepest leaf level (legal
    in a .vox file) must not overrun the per-level stack rows."""
    leaves = [(2, 3, 3, 3, 5)]  # under root slot 7; slot 0 stays empty
    nodes, root = H.tree_from_leaves(leaves)
    # hang an 8-deep chain of empty nodes off slot 0 of the root
    extra = np.zeros((8, 10), np.uint32)
    extra[:, 2:] = H.EMPTY
    extra[:, 0] = 0xFF00
    base = nodes.shape[0]
    for k in range(7):
        extra[k, 2] = base + k + 1
        extra[k, 0] = 0xFE00
    nodes = np.concatenate([nodes, extra])
    assert nodes[root, 2] == H.EMPTY
    nodes[root, 2] = base
    nodes[root, 0] &= ~np.uint32(1 << 8)
    s = yx.SVOData.from_nodes(nodes, root)
    pos, d = (-0.5, -0.4, -0.3), (1.0, 0.9, 0.8)
    o = orc.render(nodes, root, pos, d, fov=50.0, w=48, h=48)
    for tu in TUNINGS:
        g = gpu_render(yx, cuda, s, pos, d, 48, 48, fov=50.0, tuning=tu)
        assert_same(g, o, what=f"empty chain tuning {tu}")


def test_empty_scene_all_miss(yx, cuda):
    s = yx.SVOData.build_synthetic(depth=5, num_spheres=0)
    g = gpu_render(yx, cuda, s, (2, 2, 2), (-1, -1, -1), 40, 30)
    assert np.all(g["hit_id"] == H.MISS) and np.all(g["rgba"] == 0) and np.all(g["t"] == -1)


def test_render_frame_host_buffer(yx, orc, cuda):
    """RenderFrame (cell/ppu_renderer.cpp:76-85) returns the shaded colours."""
    s = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    r = yx.SVORenderer()
    r.SetScene(s)
    r.SetViewPos((0.5, 0.5, 0.3))
    r.SetViewDir((-1, -1, -1.5))  # the reference's test view (cell/main.cpp:25-27)
    r.SetResolution(1024, 768)
    img = r.RenderFrame()
    o = orc.render(s.nodes(), s.GetRoot(), (0.5, 0.5, 0.3), (-1, -1, -1.5), fov=70.0, w=1024, h=768,
                   threads=8)
    assert np.array_equal(img, o["rgba"])


def test_golden_fixture(yx, cuda):
    """GPU output hashes equal the committed oracle fixture."""
    import golden.make_golden as mg
    gold = mg.load()
    s = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    for ent in gold["views"]:
        g = gpu_render(yx, cuda, s, ent["pos"], ent["dir"], gold["w"], gold["h"], light_mode=ent["light_mode"],
                       light=gold["light"])
        assert mg.digest(g) == ent["digest"], f"view {ent['view']}"
    small = mg.load(small=True)
    for ent in small["views"]:
        g = gpu_render(yx, cuda, s, ent["pos"], ent["dir"], small["w"], small["h"], light_mode=ent["light_mode"],
                       light=small["light"])
        assert mg.digest(g) == ent["digest"], f"64x64 view {ent['view']} mode {ent['light_mode']}"
