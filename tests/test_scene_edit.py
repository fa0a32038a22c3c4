"""Host side of scene edits and page versions (the node pool of DynamicSVO /
HomoStorage and DynamicSVO::CountChangedPages, reaction/report/main.tex:67-71,
ore/src/main.cpp:127-128): writes, appends, leaf counts, lazy validation, and
the pages a renderer that has never synced would upload. No GPU work."""
import numpy as np
import pytest

import svo_helpers as H


def _scene(yx):
    return yx.SVOData.build_synthetic(depth=6, num_spheres=5, radius_min=3.0, radius_max=8.0, seed=3)


def test_write_overwrite_append_and_counts(yx):
    s = _scene(yx)
    nodes = s.nodes()
    info = s.info()
    n = nodes.shape[0]
    # recolour the leaves of one node: same leaf count, same depth
    i = int(np.nonzero(nodes[:, 0] & 0xFF)[0][0])
    row = nodes[i].copy()
    for sl in range(8):
        if (int(row[0]) >> sl) & 1:
            row[2 + sl] = 0xFF00FF00
    s.write_nodes(i, row)
    assert np.array_equal(s.nodes()[i], row)
    assert s.info() == info
    # append a node under a null slot of some node: one more leaf, one level deeper at most
    par = next(k for k in range(n) if any(not ((int(nodes[k, 0]) >> sl) & 1) and nodes[k, 2 + sl] == H.EMPTY
                                         for sl in range(8)))
    sl = next(q for q in range(8) if not ((int(nodes[par, 0]) >> q) & 1) and nodes[par, 2 + q] == H.EMPTY)
    new = np.full(10, H.EMPTY, np.uint32)
    new[0] = 1 | (0xFE << 8)
    new[2] = 0xFF0000FF
    s.write_nodes(n, new)
    assert s.info()["num_nodes"] == n + 1 and s.info()["num_leaves"] == info["num_leaves"] + 1
    prow = nodes[par].copy()
    prow[2 + sl] = n
    prow[0] &= ~np.uint32(1 << (8 + sl))
    s.write_nodes(int(par), prow)
    walked = H.walk_leaves(s.nodes(), s.GetRoot())
    assert len(walked) == info["num_leaves"] + 1
    assert s.info()["depth"] >= info["depth"]


def test_bad_edits_are_reported(yx):
    s = _scene(yx)
    n = s.info()["num_nodes"]
    with pytest.raises(yx.YoxelError):
        s.write_nodes(n + 1, np.zeros(10, np.uint32))  # past the end
    bad = s.nodes()[s.GetRoot()].copy()
    sl = next(q for q in range(8) if not ((int(bad[0]) >> q) & 1))
    bad[2 + sl] = n + 100  # child id out of range
    s.write_nodes(s.GetRoot(), bad)
    with pytest.raises(yx.YoxelError) as e:
        s.info()
    assert e.value.code == -4  # YX_ERR_FORMAT
    s.set_root(H.EMPTY)
    assert s.info()["depth"] == 0


def test_pages_to_upload_without_sync(yx):
    s = _scene(yx)
    r = yx.SVORenderer()
    n = s.info()["num_nodes"]
    pages, nbytes = r.CountChangedPages(s)
    assert pages == -(-n // 256) and nbytes == 40 * n


DESC = dict(depth=6, num_spheres=6, radius_min=3.0, radius_max=9.0, heightfield=True, hf_octaves=3,
            hf_amplitude=0.2, hf_base=0.25, seed=11)

# (mode, level, pos, radius, color): GROW 0 / CLEAR 1, pos and radius in voxels of `level`
EDITS = [
    (0, 6, (20.0, 30.5, 40.25), 9.5, 0xFF0000FF),   # grow a ball in the air
    (1, 6, (12.0, 12.0, 8.0), 7.0, 0),              # carve into the ground
    (0, 4, (2.0, 10.0, 5.5), 3.25, 0xFF00FF00),     # a coarser level: scales by 4
    (0, 6, (62.0, 1.0, 63.0), 6.0, 0xFFFF0000),     # crosses the domain's edges
    (1, 5, (10.5, 15.0, 20.0), 4.0, 0),             # clears part of the first ball
    (0, 6, (20.0, 30.5, 40.25), 3.0, 0xFF123456),   # regrows inside the cleared part
    (1, 6, (-40.0, 90.0, 10.0), 5.0, 0),            # entirely outside: no change
]


def check_edited(yx, s, desc, edits):
    nodes, root = s.nodes(), s.GetRoot()
    got = H.walk_leaves(nodes, root)
    surf = H.surface(H.edited_inside(desc, edits))
    xs, ys, zs = np.nonzero(surf)
    depth = desc["depth"]
    exp = {(depth, int(x), int(y), int(z)) for x, y, z in zip(xs, ys, zs)}
    assert set(got) == exp
    cols = H.edited_colors_np(xs, ys, zs, desc, edits)
    assert [got[(depth, int(x), int(y), int(z))][0] for x, y, z in zip(xs, ys, zs)] == cols.tolist()
    for x, y, z, c in list(zip(xs, ys, zs, cols))[::97]:  # the scalar restatement agrees
        assert H.edited_color(int(x), int(y), int(z), desc, edits) == c
    assert s.info()["num_leaves"] == len(exp)
    check_pool(nodes, root)


def leaf_colors(s):
    return {k: v[0] for k, v in H.walk_leaves(s.nodes(), s.GetRoot()).items()}


def check_pool(nodes, root):
    """Reachable nodes are a tree whose null flags match their slots; every
    other row is blank (flags 0xFF00, no children)."""
    seen = np.zeros(nodes.shape[0], bool)
    stack = [] if root == H.EMPTY else [int(root)]
    while stack:
        i = stack.pop()
        assert not seen[i]
        seen[i] = True
        f = int(nodes[i, 0])
        assert f & 0xFF or (f >> 8) & 0xFF != 0xFF  # no empty node is reachable
        for c in range(8):
            ch = int(nodes[i, 2 + c])
            if (f >> c) & 1:
                assert not (f >> (8 + c)) & 1
            else:
                assert bool((f >> (8 + c)) & 1) == (ch == H.EMPTY)
                if ch != H.EMPTY:
                    stack.append(ch)
    blank = ~seen
    assert np.all(nodes[blank, 0] == 0xFF00) and np.all(nodes[blank, 2:] == H.EMPTY)


@pytest.mark.parametrize("how", ["IN_PLACE", "REBUILD"])
def test_build_range_matches_dense_definition(yx, how):
    """DynamicSVO::BuildRange(level, pos, BuildMode, SphereSource) (ore/src/main.cpp:121-127,
    demo/Demo.cpp:82-114): every prefix of an edit sequence equals the dense
    definition: the recipe's inside set, balls united (GROW) or removed (CLEAR)
    in order, surface voxels, and the last ball holding a voxel giving its colour."""
    how = getattr(yx.SVOData, how)
    s = yx.SVOData.build_synthetic(**DESC)
    for k in range(len(EDITS)):
        mode, level, pos, radius, color = EDITS[k]
        s.BuildRange(level, pos, mode, radius, color, how=how)
        check_edited(yx, s, DESC, EDITS[:k + 1])
    # a fresh scene given the same edits in one call ends identical
    t = yx.SVOData.build_synthetic(**DESC)
    t.BuildRanges(EDITS, how=how)
    assert np.array_equal(s.nodes(), t.nodes()) and s.GetRoot() == t.GetRoot()
    # the in-place tree holds the rebuilt tree's voxels and colours
    r = yx.SVOData.build_synthetic(**DESC)
    r.BuildRanges(EDITS, how=yx.SVOData.REBUILD)
    assert leaf_colors(s) == leaf_colors(r)


def node_cubes(nodes, root, depth):
    """{node id: (level, origin voxel)} of the nodes reachable from root."""
    out = {}
    stack = [(int(root), 0, (0, 0, 0))]
    while stack:
        i, lvl, o = stack.pop()
        out[i] = (lvl, o)
        size = 1 << (depth - lvl - 1)
        f = int(nodes[i, 0])
        for c in range(8):
            if not (f >> c) & 1 and nodes[i, 2 + c] != H.EMPTY:
                stack.append((int(nodes[i, 2 + c]), lvl + 1,
                              (o[0] + (c & 1) * size, o[1] + ((c >> 1) & 1) * size, o[2] + ((c >> 2) & 1) * size)))
    return out


def test_in_place_edit_touches_only_its_region(yx):
    desc = dict(depth=8, num_spheres=30, radius_min=4.0, radius_max=24.0, heightfield=True, hf_octaves=4,
                hf_amplitude=0.2, hf_base=0.3, seed=5)
    s = yx.SVOData.build_synthetic(**desc)
    rng = np.random.default_rng(3)
    for k in range(6):
        before, root0 = s.nodes(), s.GetRoot()
        cubes = node_cubes(before, root0, 8)
        mode = k % 2
        c = rng.uniform(20, 236, 3)
        rad = float(rng.uniform(2, 9))
        s.BuildRange(8, tuple(c), mode, rad, 0xFF00A0F0)
        after = s.nodes()
        lo, hi = np.floor(c - rad) - 3, np.ceil(c + rad) + 3
        changed = 0
        for i, (lvl, o) in cubes.items():
            size = 1 << (8 - lvl)
            meets = all(o[a] <= hi[a] and o[a] + size - 1 >= lo[a] for a in range(3))
            if not meets:
                assert np.array_equal(before[i], after[i]), (k, i)
            elif i < after.shape[0] and not np.array_equal(before[i], after[i]):
                changed += 1
        assert changed < 80, changed
        check_pool(after, s.GetRoot())
    r = yx.SVOData.build_synthetic(**desc)
    rng = np.random.default_rng(3)
    edits = []
    for k in range(6):
        c = rng.uniform(20, 236, 3)
        rad = float(rng.uniform(2, 9))
        edits.append((k % 2, 8, tuple(c), rad, 0xFF00A0F0))
    r.BuildRanges(edits, how=yx.SVOData.REBUILD)
    assert leaf_colors(s) == leaf_colors(r)


def test_in_place_reuses_freed_nodes(yx):
    s = yx.SVOData.build_synthetic(**DESC)
    n0 = s.info()["num_nodes"]
    sizes = []
    for _ in range(4):
        s.BuildRange(6, (40.0, 40.0, 50.0), 0, 8.0, 0xFF0000FF)   # grow a ball in the air
        sizes.append(s.info()["num_nodes"])
        s.BuildRange(6, (40.0, 40.0, 50.0), 1, 9.0)               # and clear it again
    assert sizes[0] > n0 and sizes[1:] == sizes[:1] * 3
    check_pool(s.nodes(), s.GetRoot())


def test_build_range_on_empty_and_cleared_scenes(yx):
    desc = dict(depth=5, num_spheres=0, radius_min=0.0, radius_max=0.0, heightfield=False, hf_octaves=5,
                hf_amplitude=0.2, hf_base=0.3, seed=4)
    for how in (yx.SVOData.IN_PLACE, yx.SVOData.REBUILD):
        s = yx.SVOData.build_synthetic(**desc)
        assert s.GetRoot() == H.EMPTY
        edits = [(0, 5, (16.0, 16.0, 16.0), 6.0, 0xFF804020)]
        s.BuildRange(5, (16.0, 16.0, 16.0), 0, 6.0, 0xFF804020, how=how)
        check_edited(yx, s, desc, edits)
        # clearing everything leaves an empty tree
        s.BuildRange(1, (1.0, 1.0, 1.0), 1, 2.0, how=how)
        assert s.GetRoot() == H.EMPTY and s.info()["num_leaves"] == 0
        # a single voxel
        s.BuildRange(5, (3.5, 4.5, 5.5), 0, 0.25, 0xFF010203, how=how)
        got = H.walk_leaves(s.nodes(), s.GetRoot())
        assert got == {(5, 3, 4, 5): got[(5, 3, 4, 5)]} and got[(5, 3, 4, 5)][0] == 0xFF010203
        check_pool(s.nodes(), s.GetRoot())


def test_build_range_errors(yx):
    s = yx.SVOData.build_synthetic(**DESC)
    before = s.nodes()
    for args in ((0, (1.0, 1.0, 1.0), 0, 1.0),     # level 0
                 (7, (1.0, 1.0, 1.0), 0, 1.0),     # deeper than the tree
                 (6, (1.0, 1.0, 1.0), 2, 1.0),     # unknown mode
                 (6, (1.0, 1.0, 1.0), 0, 0.0),     # radius 0
                 (6, (1.0, 1.0, 1.0), 0, -2.0),
                 (6, (float("nan"), 1.0, 1.0), 0, 1.0),
                 (6, (1.0, 1.0, 1.0), 1, float("inf")),
                 (6, (1.0, 3e7, 1.0), 0, 1.0),     # beyond 2^24 voxels
                 (1, (1.0, 1.0, 1.0), 0, 1e6)):
        with pytest.raises(yx.YoxelError):
            s.BuildRange(*args)
    with pytest.raises(yx.YoxelError):
        s.BuildRange(6, (1.0, 1.0, 1.0), 0, 1.0, how=5)   # unknown method
    # a bad edit in a batch rejects the whole batch
    with pytest.raises(yx.YoxelError):
        s.BuildRanges([(0, 6, (10.0, 10.0, 10.0), 3.0, 0xFF0000FF), (0, 9, (1.0, 1.0, 1.0), 1.0, 0)])
    assert np.array_equal(s.nodes(), before)
    # the recipe is gone once nodes were written by hand, and a scene from nodes
    # has none: such scenes are edited in place only (as trees)
    s.write_nodes(0, before[0])
    for how in (yx.SVOData.REBUILD, yx.SVOData.REBUILD_GPU):
        with pytest.raises(yx.YoxelError):
            s.BuildRange(6, (1.0, 1.0, 1.0), 0, 1.0, how=how, stream=0)
    t = yx.SVOData.from_nodes(before, 0)
    with pytest.raises(yx.YoxelError):
        t.BuildRange(6, (1.0, 1.0, 1.0), 0, 1.0, how=yx.SVOData.REBUILD)
    t.BuildRange(6, (40.0, 40.0, 50.0), 0, 3.0, 0xFF0000FF)
    assert t.info()["num_leaves"] > s.info()["num_leaves"]


def blob(shape, seed, scale=255):
    """Smooth uint8 values [z][y][x] (a sum of Gaussians)."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*[np.arange(k, dtype=np.float64) for k in shape], indexing="ij")
    v = np.zeros(shape)
    for _ in range(4):
        c = rng.uniform(0, 1, 3) * np.array(shape)
        r = rng.uniform(1.5, max(1.6, max(shape) / 2.5))
        v += np.exp(-((z - c[0]) ** 2 + (y - c[1]) ** 2 + (x - c[2]) ** 2) / (2 * r * r))
    return np.clip(v / v.max() * scale, 0, 255).astype(np.uint8)


def raw_block(shape, seed):
    rng = np.random.default_rng(seed)
    b = blob(shape, seed + 50)
    col = rng.integers(0, 2 ** 24, size=shape, dtype=np.uint64).astype(np.uint32)
    return np.where(b > 90, col | np.uint32(0xFF000000), col & np.uint32(0x00FFFFFF)).astype(np.uint32)


MIXED = [
    dict(kind="iso", mode=0, level=6, pos=(30, 20, 36), data=blob((10, 14, 12), 1), iso=100, color=0xFF2020E0),
    (0, 6, (20.0, 30.5, 40.25), 7.5, 0xFF0000FF),
    dict(kind="raw", mode=0, level=5, pos=(-2, 25, 12), data=raw_block((6, 8, 7), 2)),       # scaled x2, crosses x = 0
    dict(kind="iso", mode=1, level=6, pos=(10, 10, 4), data=blob((12, 12, 12), 3), iso=60),  # carve into the ground
    dict(kind="iso", mode=0, level=6, pos=(55, 55, 50), data=blob((12, 12, 12), 4), iso=128, inside=False,
         color=0xFF00FF00),                                                                  # below iso, past the corner
    (1, 6, (36.0, 27.0, 40.0), 4.0, 0),
]


def apply_mixed(s, edits, how):
    for e in edits:
        if isinstance(e, dict):
            if e["kind"] == "iso":
                s.BuildRangeIso(e["level"], e["pos"], e["mode"], e["data"], e["iso"], e.get("inside", True),
                                e.get("color", 0xFFC0B6C0), how=how)
            else:
                s.BuildRangeRaw(e["level"], e["pos"], e["mode"], e["data"], how=how)
        else:
            mode, level, pos, radius, color = e
            s.BuildRange(level, pos, mode, radius, color, how=how)


def check_mixed(yx, s, desc, edits):
    nodes, root = s.nodes(), s.GetRoot()
    got = H.walk_leaves(nodes, root)
    inside, _, _, _ = H.mixed_inside_and_colors(desc, edits)
    xs, ys, zs = np.nonzero(H.surface(inside))
    depth = desc["depth"]
    assert set(got) == {(depth, int(x), int(y), int(z)) for x, y, z in zip(xs, ys, zs)}
    cols = H.mixed_leaf_colors(desc, edits, xs, ys, zs)
    assert [got[(depth, int(x), int(y), int(z))][0] for x, y, z in zip(xs, ys, zs)] == cols.tolist()
    check_pool(nodes, root)


@pytest.mark.parametrize("how", ["IN_PLACE", "REBUILD"])
def test_volume_sources_match_dense_definition(yx, how):
    """BuildRange with IsoSource / RawSource (scene_gen.py:74-82,111-130): every
    prefix of a mixed sequence equals the dense definition: the source's solid
    voxels placed at pos of its level (scaled to the finest level, clipped to
    the domain) united or removed in order with the sphere edits, the last
    source holding a voxel giving its colour."""
    how = getattr(yx.SVOData, how)
    s = yx.SVOData.build_synthetic(**DESC)
    for k in range(len(MIXED)):
        apply_mixed(s, MIXED[k:k + 1], how)
        check_mixed(yx, s, DESC, MIXED[:k + 1])
    r = yx.SVOData.build_synthetic(**DESC)
    apply_mixed(r, MIXED, yx.SVOData.REBUILD)
    assert leaf_colors(s) == leaf_colors(r)


def test_volume_source_errors(yx):
    s = yx.SVOData.build_synthetic(**DESC)
    before = s.nodes()
    v = blob((4, 4, 4), 0)
    with pytest.raises(yx.YoxelError):
        s.BuildRangeIso(7, (0, 0, 0), 0, v, 10)       # deeper than the tree
    with pytest.raises(yx.YoxelError):
        s.BuildRangeIso(6, (0, 0, 0), 3, v, 10)       # unknown mode
    with pytest.raises(yx.YoxelError):
        s.BuildRangeIso(6, (0, 0, 0), 0, np.zeros((0, 4, 4), np.uint8), 10)
    with pytest.raises(yx.YoxelError):
        s.BuildRangeIso(6, (1 << 25, 0, 0), 0, v, 10)
    assert np.array_equal(s.nodes(), before)
    # a source entirely outside the domain changes nothing
    s.BuildRangeIso(6, (-10, 0, 0), 0, v, 0)
    assert np.array_equal(s.nodes(), before)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_edit_sequences(yx, seed):
    """40 random GROW / CLEAR edits (spheres at levels 3-5, iso and raw volumes,
    some past the domain's faces) on a depth-5 scene: after every edit the
    in-place tree equals the dense definition, its blank rows are free, and a
    host rebuild of the same sequence holds the same voxels and colours."""
    rng = np.random.default_rng(100 + seed)
    desc = dict(depth=5, num_spheres=4, radius_min=2.0, radius_max=6.0, heightfield=True, hf_octaves=3,
                hf_amplitude=0.2, hf_base=0.3, seed=seed)
    s = yx.SVOData.build_synthetic(**desc)
    edits = []
    for k in range(40):
        mode = int(rng.integers(2))
        kind = rng.choice(["sphere", "sphere", "iso", "raw"])
        level = int(rng.integers(3, 6))
        n = 1 << level
        if kind == "sphere":
            pos = tuple(float(v) for v in rng.uniform(-0.2 * n, 1.2 * n, 3))
            e = (mode, level, pos, float(rng.uniform(0.3, 0.4 * n)), int(rng.integers(0, 2 ** 32)))
        else:
            shape = tuple(int(v) for v in rng.integers(1, 7, 3))
            pos = tuple(int(v) for v in rng.integers(-3, n, 3))
            if kind == "iso":
                e = dict(kind="iso", mode=mode, level=level, pos=pos, data=blob(shape, int(rng.integers(1000))),
                         iso=int(rng.integers(20, 200)), inside=bool(rng.integers(2)),
                         color=int(rng.integers(0, 2 ** 32)))
            else:
                e = dict(kind="raw", mode=mode, level=level, pos=pos, data=raw_block(shape, int(rng.integers(1000))))
        edits.append(e)
        apply_mixed(s, [e], yx.SVOData.IN_PLACE)
        check_mixed(yx, s, desc, edits)
    r = yx.SVOData.build_synthetic(**desc)
    apply_mixed(r, edits, yx.SVOData.REBUILD)
    assert leaf_colors(s) == leaf_colors(r)
