"""Pins for the CPU restatement (oracle/svo_oracle.c) of the reference's SVO
traversal. The reference ships no golden vectors for this path, so the oracle
is pinned by:
  * brute force: first leaf along a ray by exhaustive float64 slab tests over
    every leaf box, on dense and mixed-level trees of depth 1..5;
  * closed forms: axis rays into known leaves, InitRayDir geometry;
  * hand-derived cases of the reference's comparison and tie rules
    (cell/spu/trace_spu.cpp:53 strict '>', :75-78 argmin order) and of the
    leaf-before-distance order of RecTrace (cell/ppu_renderer.cpp:27-33);
  * planted mistakes: variants of the oracle built with ORACLE_PLANT must fail
    at least one of these pins.
"""
import math

import numpy as np
import pytest

import svo_helpers as H


def _random_dirs(rng, n, target=(0.5, 0.5, 0.5), pos=None, spread=0.45):
    """Directions from pos toward random points of the cube region."""
    tgt = np.asarray(target) + rng.uniform(-spread, spread, size=(n, 3))
    d = tgt - np.asarray(pos)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d.astype(np.float32)


def _mixed_level_leaves(rng, depth, density):
    """Dense random leaves at `depth`, with random full 2x2x2 blocks collapsed
    into one coarser leaf."""
    n = 1 << depth
    occ = rng.random((n, n, n)) < density
    leaves = []
    if depth >= 2:
        # collapse some blocks: mark coarse leaves and clear their fine cells
        m = n // 2
        coarse = rng.random((m, m, m)) < 0.08
        for (x, y, z) in zip(*np.nonzero(coarse)):
            occ[2 * x:2 * x + 2, 2 * y:2 * y + 2, 2 * z:2 * z + 2] = False
            leaves.append((depth - 1, int(x), int(y), int(z), int(rng.integers(0, 2**32))))
    leaves += H.dense_leaves(occ, rng)
    return leaves


def _check_against_brute(orc, leaves, rng, n_rays=400, order="bfs"):
    nodes, root = H.tree_from_leaves(leaves, order=order, seed=int(rng.integers(1 << 30)))
    keys = [(l[0], l[1], l[2], l[3]) for l in leaves]
    checked = 0
    for _ in range(3):
        az, el = rng.uniform(0, 2 * math.pi), rng.uniform(-1.2, 1.2)
        pos = np.array([0.5 + 1.6 * math.cos(el) * math.cos(az), 0.5 + 1.6 * math.cos(el) * math.sin(az),
                        0.5 + 1.6 * math.sin(el)], np.float32)
        dirs = H.adjust(_random_dirs(rng, n_rays, pos=pos))
        res = orc.trace_rays(nodes, root, pos, dirs)
        idx, tin, amb = H.brute_first_hit(keys, pos.astype(np.float64), dirs.astype(np.float64))
        ok = ~amb
        checked += int(ok.sum())
        got_hit = res["hit_id"] != H.MISS
        exp_hit = idx >= 0
        assert np.array_equal(got_hit[ok], exp_hit[ok]), "hit/miss differs from brute force"
        sel = np.nonzero(ok & exp_hit)[0]
        kk = np.asarray(keys)[idx[sel]]
        assert np.array_equal(res["level"][sel], kk[:, 0]), "hit level differs"
        assert np.array_equal(res["voxel"][sel], kk[:, 1:4]), "hit voxel differs"
        np.testing.assert_allclose(res["t"][sel], tin[sel], rtol=1e-5, atol=1e-6)
        # hit id decodes to the reported leaf: parent node's slot holds a leaf
        par = res["hit_id"][sel] // 8
        slot = res["hit_id"][sel] % 8
        assert np.all((nodes[par, 0] >> slot) & 1)
    assert checked > 0.9 * 3 * n_rays
    return checked


@pytest.mark.parametrize("depth", [1, 2, 3, 4, 5])
def test_brute_force_dense(orc, depth):
    rng = np.random.default_rng(100 + depth)
    n = 1 << depth
    occ = rng.random((n, n, n)) < (0.25 if depth > 2 else 0.5)
    _check_against_brute(orc, H.dense_leaves(occ, rng), rng)


@pytest.mark.parametrize("depth", [3, 4, 5])
def test_brute_force_mixed_levels_shuffled_ids(orc, depth):
    rng = np.random.default_rng(200 + depth)
    _check_against_brute(orc, _mixed_level_leaves(rng, depth, 0.06), rng, order="shuffle")


def test_sparse_surface_scene_brute_force(orc, yx):
    # the synthetic voxelizer's own output at depth 5, checked exhaustively
    s = yx.SVOData.build_synthetic(depth=5, num_spheres=6, radius_min=2.0, radius_max=6.0,
                                   heightfield=True, seed=9)
    nodes, root = s.nodes(), s.GetRoot()
    leaves = [(k[0], k[1], k[2], k[3], v[0]) for k, v in H.walk_leaves(nodes, root).items()]
    rng = np.random.default_rng(7)
    _check_against_brute(orc, leaves, rng, n_rays=300)


def test_axis_ray_closed_form(orc):
    # depth-1 tree; ray along +x at y = z = 0.25 enters the cube at x = 0, t = 1
    nodes, root = H.tree_from_leaves([(1, 1, 0, 0, 0xAABBCCDD)])
    r = orc.trace_rays(nodes, root, [[-1.0, 0.25, 0.25]], [[1.0, 0.0, 0.0]])
    # child 1 (x in [0.5,1)) is entered at x = 0.5: t = 1.5 exactly
    assert r["hit_id"][0] == root * 8 + 1
    assert r["t"][0] == np.float32(1.5)
    assert list(r["voxel"][0]) == [1, 0, 0] and r["level"][0] == 1
    # the same leaf seen along -x from the far side: t = (1 - 2)/-1 = 1
    r = orc.trace_rays(nodes, root, [[2.0, 0.25, 0.25]], [[-1.0, 0.0, 0.0]])
    assert r["hit_id"][0] == root * 8 + 1 and r["t"][0] == np.float32(1.0)
    # a ray that passes beside the leaf misses
    r = orc.trace_rays(nodes, root, [[-1.0, 0.75, 0.25]], [[1.0, 0.0, 0.0]])
    assert r["hit_id"][0] == H.MISS


def test_deep_path_closed_form(orc):
    # one leaf at level 12 in the corner voxel (0,0,0): axis ray enters at t=1
    nodes, root = H.tree_from_leaves([(12, 5, 7, 9, 1)])
    lo = np.array([5, 7, 9], np.float64) / 4096.0
    p = np.array([-1.0, lo[1] + 0.5 / 4096, lo[2] + 0.5 / 4096], np.float32)
    r = orc.trace_rays(nodes, root, [p], [[1.0, 0.0, 0.0]])
    assert list(r["voxel"][0]) == [5, 7, 9] and r["level"][0] == 12
    np.testing.assert_allclose(r["t"][0], 1.0 + lo[0], rtol=1e-6)
    # visits: the root and 11 inner nodes on the path
    assert r["visits"][0] == 12


def test_first_child_strict_comparison(orc):
    # Ray (1,1,0) from (-1,-0.5,0.25) enters at (0, 0.5, 0.25), exactly on the
    # y mid-plane: t1.x = 1, t1.y = 0.5, t2.y = 1.5, tm.y = 1 = tEnter. With the
    # strict '>' of cell/spu/trace_spu.cpp:53 the first child is the lower-y
    # one (slot 0), hit at t = 1; a '>=' would pick slot 2.
    nodes, root = H.tree_from_leaves([(1, 0, 0, 0, 10), (1, 0, 1, 0, 20)])
    r = orc.trace_rays(nodes, root, [[-1.0, -0.5, 0.25]], [[1.0, 1.0, 0.0]])
    assert r["hit_id"][0] == root * 8 + 0 and r["t"][0] == np.float32(1.0)


def test_go_next_tie_order(orc):
    # Ray (1,1,0) from (-1,-1,0.25): after FindFirstChild t2 = (1.5, 1.5, ..)
    # ties on x/y. cell/spu/trace_spu.cpp:75-78 takes x when !(t2.x > t2.y), so
    # the ray steps into slot 1 (x upper) first and hits it at t = 1.5.
    nodes, root = H.tree_from_leaves([(1, 1, 0, 0, 10), (1, 0, 1, 0, 20)])
    r = orc.trace_rays(nodes, root, [[-1.0, -1.0, 0.25]], [[1.0, 1.0, 0.0]])
    assert r["hit_id"][0] == root * 8 + 1 and r["t"][0] == np.float32(1.5)


def test_leaf_behind_eye_is_reported(orc):
    # RecTrace tests the leaf flag before any distance test
    # (cell/ppu_renderer.cpp:27-33): from inside the cube at x = 0.75 looking
    # +x, the first child in traversal order is slot 0 (x < 0.5, behind the
    # eye) and it is returned with t = max(t1) = -0.75.
    nodes, root = H.tree_from_leaves([(1, 0, 0, 0, 10), (1, 1, 0, 0, 20)])
    r = orc.trace_rays(nodes, root, [[0.75, 0.25, 0.25]], [[1.0, 0.0, 0.0]])
    assert r["hit_id"][0] == root * 8 + 0 and r["t"][0] == np.float32(-0.75)
    # the occlusion form used by shadow rays skips it and finds slot 1 at t = 0
    r = orc.trace_rays(nodes, root, [[0.75, 0.25, 0.25]], [[1.0, 0.0, 0.0]], occlusion=True)
    assert r["hit_id"][0] == root * 8 + 1 and r["t"][0] == np.float32(-0.25)


def test_adjust_dir_axis_ray(orc):
    # zero components become +1e-6 (ntree/ntree_trace.cu:38-43); an exactly
    # axis-aligned ray along -z hits the top leaf of its column
    nodes, root = H.tree_from_leaves([(2, 1, 2, 0, 1), (2, 1, 2, 3, 2)])
    r = orc.trace_rays(nodes, root, [[0.3, 0.6, 2.0]], [[0.0, 0.0, -1.0]])
    assert list(r["voxel"][0]) == [1, 2, 3] and r["t"][0] == np.float32(1.0)


def test_init_ray_dir_geometry(orc):
    fwd = np.array([-1.0, -1.0, -1.5])
    w, h, fov = 1024, 768, 70.0
    b = orc.init_ray_dir(fwd, (0, 0, 1), fov, w, h)
    dirs = orc.pixel_dirs(b, w, h)
    f = fwd / np.linalg.norm(fwd)
    # centre pixel (w/2, h/2) looks along the view direction
    np.testing.assert_allclose(dirs[h // 2, w // 2], f, atol=2e-6)
    # the left edge pixel (0, h/2) is fov/2 off-axis (horizontal fov)
    ang = math.degrees(math.acos(float(np.dot(dirs[h // 2, 0], f))))
    assert abs(ang - fov / 2) < 1e-3
    # rows go down the screen: pixel y grows against the up vector
    assert dirs[0, w // 2, 2] > dirs[h - 1, w // 2, 2]
    # columns go right: (fwd x up) direction
    right = np.cross(f, [0, 0, 1])
    assert np.dot(dirs[h // 2, w - 1] - dirs[h // 2, 0], right) > 0
    # unit length
    np.testing.assert_allclose(np.linalg.norm(dirs.reshape(-1, 3), axis=1), 1.0, atol=1e-6)


def test_render_rows_counts_and_shade(orc, yx):
    s = yx.SVOData.build_synthetic(depth=5, num_spheres=5, radius_min=2.0, radius_max=5.0, seed=3)
    pos, d = yx.workloads.orbit_cameras(2, seed=1)[0]
    out = orc.render(s.nodes(), s.GetRoot(), pos, d, fov=60, w=48, h=40, light_mode=2,
                     light=yx.workloads.SUN)
    hit = out["hit_id"] != H.MISS
    assert hit.any() and (~hit).any()
    # misses: cleared colour (0,0,0,0) as cell/ppu_renderer.cpp:54
    assert np.all(out["rgba"][~hit] == 0) and np.all(out["shadow"][~hit] == 2)
    assert np.all((out["rgba"][hit] >> 24) == 255)
    assert out["rays"] == int(hit.size + (out["shadow"] < 2).sum())
    # shaded channels never exceed the leaf colour and never drop below 0.1x
    nodes = s.nodes()
    leaf = nodes[out["hit_id"][hit] // 8, 2 + out["hit_id"][hit] % 8]
    for c in range(3):
        lc = (leaf >> (8 * c)) & 255
        sc = (out["rgba"][hit] >> (8 * c)) & 255
        assert np.all(sc <= lc) and np.all(sc + 1 >= np.floor(lc * np.float32(0.1)))


# ---- planted mistakes --------------------------------------------------------

PIN_TESTS = [test_first_child_strict_comparison, test_go_next_tie_order, test_axis_ray_closed_form]


@pytest.mark.parametrize("plant", [1, 2, 3, 4, 5])
def test_planted_mistakes_are_caught(oracle_mod, tmp_path, plant, yx):
    lib = oracle_mod.build(plant=plant, out=tmp_path / f"liboracle_plant{plant}.so")
    bad = oracle_mod.Oracle(lib)
    assert bad.plant() == plant
    failures = 0
    for t in PIN_TESTS:
        try:
            t(bad)
        except AssertionError:
            failures += 1
    for depth in (3, 4):
        try:
            rng = np.random.default_rng(300 + depth)
            n = 1 << depth
            occ = rng.random((n, n, n)) < 0.2
            _check_against_brute(bad, H.dense_leaves(occ, rng), rng, n_rays=200)
        except AssertionError:
            failures += 1
    assert failures > 0, f"planted mistake {plant} passed every pin"
