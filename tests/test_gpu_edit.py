"""Sphere edits with the GPU (DynamicSVO::BuildRange with SphereSource,
ore/src/main.cpp:121-127, demo/Demo.cpp:82-114; then CudaSVO::Update,
demo/SVORenderer.cpp:39-53): the scene rebuilt on the GPU after each edit
equals the CPU rebuild byte for byte; an in-place edit changes only the pages
around the ball, UpdateScene uploads exactly those, and the edited scene
renders what the CPU restatement renders."""
import time

import numpy as np
import pytest

import svo_helpers as H
from test_gpu_update import render, same_layout
from test_scene_edit import DESC, EDITS, check_pool, leaf_colors

pytestmark = pytest.mark.gpu


def changed_pages(a, b):
    """256-node pages of b whose bytes differ from a's (or are new)."""
    nb = b.shape[0]
    pages = 0
    for p in range(-(-nb // 256)):
        n0, n1 = p * 256, min(nb, p * 256 + 256)
        if n1 > a.shape[0] or not np.array_equal(a[n0:n1], b[n0:n1]):
            pages += 1
    return pages


def test_gpu_rebuild_equals_cpu_rebuild(yx, cuda):
    c = yx.SVOData.build_synthetic(**DESC)
    g = yx.SVOData.build_synthetic_gpu(**DESC)
    for k, (mode, level, pos, radius, color) in enumerate(EDITS):
        print("edit", k, flush=True)
        c.BuildRange(level, pos, mode, radius, color, how=yx.SVOData.REBUILD)
        g.BuildRange(level, pos, mode, radius, color, how=yx.SVOData.REBUILD_GPU, stream=0)
        assert c.info() == g.info()
        assert np.array_equal(c.nodes(), g.nodes())
    # in-place edits first, then a GPU rebuild: the canonical tree again
    m = yx.SVOData.build_synthetic(**DESC)
    m.BuildRanges(EDITS[:4])
    m.BuildRanges(EDITS[4:], how=yx.SVOData.REBUILD_GPU, stream=0)
    assert np.array_equal(m.nodes(), c.nodes())


def test_gpu_rebuild_of_a_large_scene(yx, cuda):
    desc = dict(depth=10, num_spheres=200, radius_min=8.0, radius_max=60.0, heightfield=True, hf_octaves=5,
                hf_amplitude=0.2, hf_base=0.3, seed=21)
    c = yx.SVOData.build_synthetic(**desc)
    g = yx.SVOData.build_synthetic_gpu(**desc)
    p = yx.SVOData.build_synthetic_gpu(**desc)
    edits = [(0, 10, (500.0, 520.0, 700.0), 90.0, 0xFF2040F0),
             (1, 8, (60.0, 60.0, 40.0), 30.0, 0),
             (0, 10, (1000.0, 3.0, 1020.0), 40.5, 0xFF00FFFF)]
    for mode, level, pos, radius, color in edits:
        t0 = time.perf_counter()
        c.BuildRange(level, pos, mode, radius, color, how=yx.SVOData.REBUILD)
        t1 = time.perf_counter()
        g.BuildRange(level, pos, mode, radius, color, how=yx.SVOData.REBUILD_GPU, stream=0)
        t2 = time.perf_counter()
        p.BuildRange(level, pos, mode, radius, color)
        t3 = time.perf_counter()
        print(f"edit: cpu rebuild {t1 - t0:.3f} s, gpu rebuild {t2 - t1:.3f} s, in place {t3 - t2:.3f} s, "
              f"{g.info()['num_nodes']} nodes", flush=True)
        assert c.info() == g.info()
        assert np.array_equal(c.nodes(), g.nodes())
    assert leaf_colors(p) == leaf_colors(g)
    check_pool(p.nodes(), p.GetRoot())


def test_edit_versions_update_and_render(yx, cuda, orc):
    s = yx.SVOData.build_synthetic_gpu(depth=8, num_spheres=24, radius_min=4.0, radius_max=22.0,
                                       heightfield=True, seed=13)
    r = same_layout(yx, s, what="initial")
    assert r.CountChangedPages(s) == (0, 0)
    # an edit far out of the domain changes no page
    s.BuildRange(8, (-100.0, 40.0, 40.0), 0, 20.0, 0xFF0000FF)
    assert r.CountChangedPages(s) == (0, 0)
    steps = [(0, 8, (200.0, 60.0, 180.0), 14.0, 0xFF3060C0, yx.SVOData.IN_PLACE),
             (1, 8, (100.0, 100.0, 60.0), 24.0, 0, yx.SVOData.IN_PLACE),
             (0, 6, (10.0, 50.0, 20.0), 2.5, 0xFFFFFFFF, yx.SVOData.IN_PLACE),
             (1, 8, (30.0, 200.0, 50.0), 10.0, 0, yx.SVOData.REBUILD_GPU),
             (0, 8, (128.0, 128.0, 128.0), 6.0, 0xFF00FF00, yx.SVOData.IN_PLACE)]
    for k, (mode, level, pos, radius, color, how) in enumerate(steps):
        before = s.nodes()
        s.BuildRange(level, pos, mode, radius, color, how=how, stream=0)
        after = s.nodes()
        exp = changed_pages(before, after)
        total = -(-after.shape[0] // 256)
        pages, nbytes = r.CountChangedPages(s)
        print(f"edit {k} ({how}): {exp} of {total} pages changed", flush=True)
        assert pages == exp and exp > 0
        if how == yx.SVOData.IN_PLACE:
            assert exp < total // 4
        same_layout(yx, s, r, f"edit {k}")
        assert r.last_update()[0] == exp
        g = render(yx, cuda, r, pos=(1.5, 1.4, 1.2), d=(-1.0, -0.9, -0.7))
        o = orc.render(s.nodes(), s.GetRoot(), (1.5, 1.4, 1.2), (-1.0, -0.9, -0.7), fov=60.0, w=96, h=64,
                       light_mode=2, light=(0.4, 0.3, 0.85))
        assert np.array_equal(g["hit_id"].view(np.uint32), o["hit_id"])
        assert np.array_equal(g["rgba"].view(np.uint32), o["rgba"])


def test_gpu_rebuild_with_volume_sources(yx, cuda):
    """IsoSource / RawSource edits: the GPU rebuild equals the host rebuild byte
    for byte, at depth 6 (the mixed sequence) and depth 9 (a 48^3 source at a
    coarser level, scaled x2), and the in-place tree holds the same voxels."""
    from test_scene_edit import MIXED, apply_mixed, blob, raw_block
    c = yx.SVOData.build_synthetic(**DESC)
    g = yx.SVOData.build_synthetic_gpu(**DESC)
    for k, e in enumerate(MIXED):
        print("mixed edit", k, flush=True)
        apply_mixed(c, [e], yx.SVOData.REBUILD)
        apply_mixed(g, [e], yx.SVOData.REBUILD_GPU)
        assert c.info() == g.info()
        assert np.array_equal(c.nodes(), g.nodes())
    desc = dict(depth=9, num_spheres=40, radius_min=6.0, radius_max=30.0, heightfield=True, seed=17)
    c = yx.SVOData.build_synthetic(**desc)
    g = yx.SVOData.build_synthetic_gpu(**desc)
    p = yx.SVOData.build_synthetic(**desc)
    vol = [dict(kind="iso", mode=0, level=8, pos=(40, 60, 70), data=blob((48, 48, 48), 9), iso=70,
                color=0xFF4080FF),
           dict(kind="raw", mode=0, level=9, pos=(300, 310, 200), data=raw_block((40, 30, 20), 5)),
           dict(kind="iso", mode=1, level=9, pos=(100, 120, 100), data=blob((64, 64, 64), 11), iso=50)]
    for e in vol:
        apply_mixed(c, [e], yx.SVOData.REBUILD)
        apply_mixed(g, [e], yx.SVOData.REBUILD_GPU)
        apply_mixed(p, [e], yx.SVOData.IN_PLACE)
        assert np.array_equal(c.nodes(), g.nodes())
    assert leaf_colors(p) == leaf_colors(c)
    check_pool(p.nodes(), p.GetRoot())


@pytest.mark.parametrize("seed", [0, 1])
def test_gpu_rebuild_random_sequences(yx, cuda, seed):
    """20 random sphere and volume edits: GPU and host rebuilds identical byte for
    byte after every fifth edit."""
    from test_scene_edit import apply_mixed, blob, raw_block
    rng = np.random.default_rng(200 + seed)
    desc = dict(depth=6, num_spheres=6, radius_min=2.0, radius_max=9.0, heightfield=True, seed=seed)
    c = yx.SVOData.build_synthetic(**desc)
    g = yx.SVOData.build_synthetic_gpu(**desc)
    batch = []
    for k in range(20):
        mode = int(rng.integers(2))
        level = int(rng.integers(4, 7))
        n = 1 << level
        kind = k % 3
        if kind == 0:
            e = (mode, level, tuple(float(v) for v in rng.uniform(-0.2 * n, 1.2 * n, 3)),
                 float(rng.uniform(0.3, 0.4 * n)), int(rng.integers(0, 2 ** 32)))
        elif kind == 1:
            e = dict(kind="iso", mode=mode, level=level, pos=tuple(int(v) for v in rng.integers(-3, n, 3)),
                     data=blob(tuple(int(v) for v in rng.integers(1, 9, 3)), k), iso=int(rng.integers(20, 200)),
                     inside=bool(rng.integers(2)), color=int(rng.integers(0, 2 ** 32)))
        else:
            e = dict(kind="raw", mode=mode, level=level, pos=tuple(int(v) for v in rng.integers(-3, n, 3)),
                     data=raw_block(tuple(int(v) for v in rng.integers(1, 9, 3)), k))
        batch.append(e)
        if len(batch) == 5:
            apply_mixed(c, batch, yx.SVOData.REBUILD)
            apply_mixed(g, batch, yx.SVOData.REBUILD_GPU)
            batch = []
            assert c.info() == g.info()
            assert np.array_equal(c.nodes(), g.nodes()), f"seed {seed} after edit {k}"


def test_tree_edits_sync_and_render(yx, cuda, orc):
    """Edits of a loaded tree with FullNode interiors (qtview.py's flow: load,
    BuildRange, UpdateScene): the GPU re-layout equals the host one, only the
    edited pages go up, and the renders equal the oracle's."""
    from test_scene_tree_edit import EDITS, check_edit, solid_scene
    nodes, root, _, _ = solid_scene(7, 5)
    s = yx.SVOData.from_nodes(nodes, root)
    r = same_layout(yx, s, what="loaded tree")
    scale = [(0, 7, (50.0, 80.0, 90.0), 9.0, 0xFF0000FF), (1, 7, (64.0, 64.0, 20.0), 14.0, 0xFF808080),
             (0, 5, (3.0, 20.0, 8.0), 2.5, 0xFF00FF00)]
    for k, e in enumerate(scale):
        before = s.nodes()
        check_edit(s, 7, e)
        after = s.nodes()
        exp = changed_pages(before, after)
        pages, _ = r.CountChangedPages(s)
        print(f"tree edit {k}: {exp} of {-(-after.shape[0] // 256)} pages", flush=True)
        assert pages == exp and exp > 0
        same_layout(yx, s, r, f"tree edit {k}")
        g = render(yx, cuda, r, pos=(1.5, 1.4, 1.2), d=(-1.0, -0.9, -0.7))
        o = orc.render(s.nodes(), s.GetRoot(), (1.5, 1.4, 1.2), (-1.0, -0.9, -0.7), fov=60.0, w=96, h=64,
                       light_mode=2, light=(0.4, 0.3, 0.85))
        assert np.array_equal(g["hit_id"].view(np.uint32), o["hit_id"])
        assert np.array_equal(g["rgba"].view(np.uint32), o["rgba"])


def test_filled_tree_renders_the_same(yx, cuda, orc):
    """Filling a loaded surface-only tree changes no render (FullNode is not
    hit), and its edits sync and render like any other."""
    syn = yx.SVOData.build_synthetic(depth=8, num_spheres=24, radius_min=4.0, radius_max=22.0, heightfield=True,
                                     seed=13)
    s = yx.SVOData.from_nodes(syn.nodes(), syn.GetRoot())
    r = same_layout(yx, s, what="surface-only tree")
    g0 = render(yx, cuda, r, pos=(1.5, 1.4, 1.2), d=(-1.0, -0.9, -0.7))
    assert s.fill_interior() > 0
    same_layout(yx, s, r, "filled tree")
    g1 = render(yx, cuda, r, pos=(1.5, 1.4, 1.2), d=(-1.0, -0.9, -0.7))
    for k in g0:
        assert np.array_equal(g0[k], g1[k]), k
    s.BuildRange(8, (128.0, 128.0, 40.0), 1, 20.0, 0xFF808080)   # a crater with solid walls
    same_layout(yx, s, r, "crater")
    g2 = render(yx, cuda, r, pos=(1.5, 1.4, 1.2), d=(-1.0, -0.9, -0.7))
    o = orc.render(s.nodes(), s.GetRoot(), (1.5, 1.4, 1.2), (-1.0, -0.9, -0.7), fov=60.0, w=96, h=64,
                   light_mode=2, light=(0.4, 0.3, 0.85))
    assert np.array_equal(g2["hit_id"].view(np.uint32), o["hit_id"])
    assert np.array_equal(g2["rgba"].view(np.uint32), o["rgba"])
