"""BuildRange on scenes without a recipe (loaded .vox trees, or trees written
node by node), in place: qtview.py:26-27,144-153 loads data/scene.vox and grows
or clears spheres in it (ore/src/main.cpp:121-127). The solid of such a tree is
its leaves (a leaf slot covers its child cube) and its FullNode cubes. An edit
unions (GROW) or removes (CLEAR) the source's solid voxels; the voxels of the
edit's box plus one are encoded again at the tree's leaf level: surface voxels
as leaves, other solid voxels as FullNode, the rest empty. Leaves keep their
colour, grown voxels take the source's, voxels a CLEAR exposes the CLEAR's.

These tests restate that on dense grids: decode the tree before and after,
apply the edit to the decoded solid, and require the decoded tree after the
edit to equal the restatement voxel for voxel, with every node outside the box
keeping its bytes."""
import math

import numpy as np
import pytest

import svo_helpers as H
from test_scene_edit import blob, raw_block


def edit_box(depth, e):
    """The voxels an edit re-encodes: its box within the domain plus one."""
    n = 1 << depth
    if isinstance(e, dict):
        s = 1 << (depth - e["level"])
        size = e["data"].shape[::-1]
        lo = [p * s for p in e["pos"]]
        hi = [lo[a] + size[a] * s - 1 for a in range(3)]
        lo = [max(0, min(n, max(-1, v))) for v in lo]
        hi = [min(n - 1, min(n, max(-1, v))) for v in hi]
    else:
        (_, c, r, _), = H.edit_balls(depth, [e])
        lo = [max(0, int(max(-1.0, math.floor(c[a] - r - 1.0)))) for a in range(3)]
        hi = [min(n - 1, int(min(float(n), math.ceil(c[a] + r + 1.0)))) for a in range(3)]
    if any(lo[a] > hi[a] for a in range(3)):
        return None
    return tuple(slice(max(0, lo[a] - 1), min(n - 1, hi[a] + 1) + 1) for a in range(3))


def shape_of(depth, e):
    """(mask, colours, mode, colour of exposed voxels) of an edit."""
    n = 1 << depth
    if isinstance(e, dict):
        solid, cols = H.source_mask(e)
        m = H.place_source(depth, e["level"], e["pos"], solid)
        col = H.place_colors(depth, e["level"], e["pos"], cols)
        return m, col, e["mode"], (e.get("color", 0xFFC0B6C0) if e["kind"] == "iso" else 0)
    (mode, (cx, cy, cz), r, color), = H.edit_balls(depth, [e])
    c = np.arange(n, dtype=np.float64) + 0.5
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    m = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= r * r
    return m, np.full((n, n, n), color, np.uint32), mode, color


def expected_after(depth, state, color, e):
    m, col, mode, ecol = shape_of(depth, e)
    solid = state > 0
    new_solid = solid | m if mode == 0 else solid & ~m
    exp_s, exp_c = state.copy(), color.copy()
    box = edit_box(depth, e)
    if box is None:
        return exp_s, exp_c
    canon = H.canonical_states(new_solid)
    grown = m & (mode == 0)
    leafcol = np.where(grown, col, np.where(state == 2, color, np.uint32(ecol))).astype(np.uint32)
    exp_s[box] = canon[box]
    exp_c[box] = np.where(canon[box] == 2, leafcol[box], 0)
    return exp_s, exp_c


def node_cubes(nodes, root, depth):
    out = {}
    if root & 0x80000000:
        return out
    stack = [(int(root), 0, (0, 0, 0))]
    while stack:
        i, lvl, o = stack.pop()
        out[i] = (lvl, o)
        half = 1 << (depth - lvl - 1)
        f = int(nodes[i, 0])
        for c in range(8):
            ch = int(nodes[i, 2 + c])
            if not (f >> c) & 1 and not ch & 0x80000000:
                stack.append((ch, lvl + 1, (o[0] + (c & 1) * half, o[1] + ((c >> 1) & 1) * half,
                                            o[2] + ((c >> 2) & 1) * half)))
    return out


def check_tree(nodes, root):
    """A tree (no node reached twice), flags consistent with the slots, every
    unreachable row blank."""
    seen = np.zeros(nodes.shape[0], bool)
    stack = [] if root & 0x80000000 else [int(root)]
    while stack:
        i = stack.pop()
        assert not seen[i]
        seen[i] = True
        f = int(nodes[i, 0])
        for c in range(8):
            ch = int(nodes[i, 2 + c])
            if (f >> c) & 1:
                assert not (f >> (8 + c)) & 1
            else:
                assert bool((f >> (8 + c)) & 1) == bool(ch & 0x80000000)
                if not ch & 0x80000000:
                    stack.append(ch)
    assert np.all(nodes[~seen, 0] == 0xFF00) and np.all(nodes[~seen, 2:] == H.EMPTY)


def apply_one(s, e):
    if isinstance(e, dict):
        if e["kind"] == "iso":
            s.BuildRangeIso(e["level"], e["pos"], e["mode"], e["data"], e["iso"], e.get("inside", True),
                            e.get("color", 0xFFC0B6C0))
        else:
            s.BuildRangeRaw(e["level"], e["pos"], e["mode"], e["data"])
    else:
        mode, level, pos, radius, color = e
        s.BuildRange(level, pos, mode, radius, color)


def check_edit(s, depth, e):
    before, root0 = s.nodes(), s.GetRoot()
    st0, col0 = H.decode_tree(before, root0, depth)
    cubes = node_cubes(before, root0, depth)
    apply_one(s, e)
    after, root1 = s.nodes(), s.GetRoot()
    st1, col1 = H.decode_tree(after, root1, depth)
    exp_s, exp_c = expected_after(depth, st0, col0, e)
    bad = np.argwhere(st1 != exp_s)
    assert bad.size == 0, f"{len(bad)} voxels in the wrong state, first {bad[:3].tolist()}"
    badc = np.argwhere(col1 != exp_c)
    assert badc.size == 0, f"{len(badc)} wrong colours, first {badc[:3].tolist()}"
    box = edit_box(depth, e)
    if box is not None:
        for i, (lvl, o) in cubes.items():
            size = 1 << (depth - lvl)
            if any(o[a] > box[a].stop - 1 or o[a] + size - 1 < box[a].start for a in range(3)) and i < after.shape[0]:
                assert np.array_equal(before[i], after[i]), i
    else:
        assert np.array_equal(before, after) and root0 == root1
    check_tree(after, root1)
    assert s.info()["num_leaves"] == int(sum(bin(int(f) & 0xFF).count("1") for f in after[:, 0]))


def solid_scene(depth, seed, coarse=True):
    """A tree with FullNode interiors: spheres on a floor, surface leaves with
    hash colours, and (optionally) two coarse leaf cubes in the air."""
    n = 1 << depth
    rng = np.random.default_rng(seed)
    c = np.arange(n) + 0.5
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    solid = Z < n * 0.3
    for _ in range(5):
        p = rng.uniform(0, n, 3)
        r = rng.uniform(2, n / 4)
        solid |= (X - p[0]) ** 2 + (Y - p[1]) ** 2 + (Z - p[2]) ** 2 <= r * r
    state = H.canonical_states(solid)
    xs, ys, zs = np.nonzero(state == 2)
    color = np.zeros((n, n, n), np.uint32)
    color[xs, ys, zs] = H.leaf_colors_np(H.morton3(xs, ys, zs), seed)
    cq = {}
    if coarse:
        for lvl, col in ((depth - 2, 0xFF00C0FF), (depth - 1, 0xFF40FF40)):
            size = n >> lvl
            for _ in range(100):
                o = tuple(int(v) * size for v in rng.integers(0, n // size, 3))
                sl = tuple(slice(o[a], o[a] + size) for a in range(3))
                if not solid[sl].any():
                    cq[(lvl, *o)] = col
                    state[sl] = 2
                    color[sl] = col
                    break
    nodes, root = H.encode_tree(state, color, cq)
    return nodes, root, state, color


def test_encoder_round_trip():
    nodes, root, state, color = solid_scene(5, 3)
    st, col = H.decode_tree(nodes, root, 5)
    assert np.array_equal(st, state) and np.array_equal(col, color)
    assert (st == 1).sum() > 1000 and (st == 2).sum() > 1000


EDITS = [
    (0, 5, (12.0, 20.0, 22.0), 4.5, 0xFF0000FF),      # grow in the air, touching a coarse cube or not
    (1, 5, (16.0, 16.0, 6.0), 5.0, 0xFF808080),       # carve into the floor: exposes FullNode voxels
    (1, 4, (3.0, 3.0, 2.0), 2.0, 0xFF102030),         # a coarser level
    (0, 5, (31.0, 0.5, 12.0), 3.0, 0xFFFF0000),       # across two faces of the domain
    dict(kind="iso", mode=0, level=5, pos=(20, 4, 10), data=blob((6, 8, 7), 1), iso=90, color=0xFF00FFFF),
    dict(kind="raw", mode=1, level=5, pos=(8, 22, 4), data=raw_block((5, 6, 6), 2)),
    dict(kind="iso", mode=1, level=4, pos=(-1, 10, 2), data=blob((3, 4, 5), 3), iso=60, color=0xFFAA5500),
]


def test_edits_of_a_loaded_tree(yx, tmp_path):
    nodes, root, _, _ = solid_scene(5, 3)
    path = tmp_path / "solid.vox"
    yx.SVOData.from_nodes(nodes, root).Save(str(path))
    s = yx.SVOData()
    s.Load(str(path))
    for e in EDITS:
        check_edit(s, 5, e)
    # coarse leaves every edit left alone are still there
    st, _ = H.decode_tree(s.nodes(), s.GetRoot(), 5)
    assert (st == 1).any()
    # the edited tree saves and loads back unchanged
    s.Save(str(path))
    t = yx.SVOData()
    t.Load(str(path))
    assert np.array_equal(t.nodes(), s.nodes()) and t.GetRoot() == s.GetRoot()


@pytest.mark.parametrize("seed", [0, 1])
def test_random_edits_of_a_tree(yx, seed):
    rng = np.random.default_rng(300 + seed)
    nodes, root, _, _ = solid_scene(5, 10 + seed)
    s = yx.SVOData.from_nodes(nodes, root)
    for k in range(25):
        mode = int(rng.integers(2))
        level = int(rng.integers(3, 6))
        n = 1 << level
        kind = k % 3
        if kind == 0:
            e = (mode, level, tuple(float(v) for v in rng.uniform(-0.2 * n, 1.2 * n, 3)),
                 float(rng.uniform(0.3, 0.35 * n)), int(rng.integers(0, 2 ** 32)))
        elif kind == 1:
            e = dict(kind="iso", mode=mode, level=level, pos=tuple(int(v) for v in rng.integers(-2, n, 3)),
                     data=blob(tuple(int(v) for v in rng.integers(1, 7, 3)), k), iso=int(rng.integers(20, 200)),
                     inside=bool(rng.integers(2)), color=int(rng.integers(0, 2 ** 32)))
        else:
            e = dict(kind="raw", mode=mode, level=level, pos=tuple(int(v) for v in rng.integers(-2, n, 3)),
                     data=raw_block(tuple(int(v) for v in rng.integers(1, 7, 3)), k))
        check_edit(s, 5, e)


def test_hollow_tree(yx):
    """A surface-only tree (no FullNode): its solid is its leaves, so a CLEAR
    through the shell opens it."""
    t = yx.SVOData.build_synthetic(depth=5, num_spheres=4, radius_min=3.0, radius_max=7.0, seed=4)
    s = yx.SVOData.from_nodes(t.nodes(), t.GetRoot())
    for e in EDITS[:4]:
        check_edit(s, 5, e)


def test_tree_edit_errors(yx):
    nodes, root, _, _ = solid_scene(5, 3)
    s = yx.SVOData.from_nodes(nodes, root)
    before = s.nodes()
    with pytest.raises(yx.YoxelError):
        s.BuildRange(5, (4.0, 4.0, 4.0), 0, 2.0, how=yx.SVOData.REBUILD)   # needs a recipe
    with pytest.raises(yx.YoxelError):
        s.BuildRange(6, (4.0, 4.0, 4.0), 0, 2.0)                           # below the leaf level
    with pytest.raises(yx.YoxelError):
        s.BuildRange(1, (1.0, 1.0, 1.0), 0, 2e6)                           # beyond 2^24 voxels
    assert np.array_equal(s.nodes(), before)
    # a ball over the whole domain fills it: leaves on the domain's faces, FullNode inside
    s.BuildRange(1, (1.0, 1.0, 1.0), 0, 200.0, 0xFF123456)
    st, col = H.decode_tree(s.nodes(), s.GetRoot(), 5)
    assert np.array_equal(st, H.canonical_states(np.ones((32, 32, 32), bool)))
    assert np.all(col[st == 2] == 0xFF123456)
    check_tree(s.nodes(), s.GetRoot())
    e = yx.SVOData.from_nodes(np.zeros((0, 10), np.uint32), H.EMPTY)
    with pytest.raises(yx.YoxelError):
        e.BuildRange(1, (0.5, 0.5, 0.5), 0, 0.4)                           # no leaf level
