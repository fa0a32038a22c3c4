"""Pins for the shading half of the CPU restatement (oracle/svo_oracle.c,
shade_pixel): the Lambert term over the entry face and the shadow rays.

The reference's shader is missing (cell/ppu_renderer.cpp:67 calls
SimpleShader::Shade, whose source is not in the tree); the lighting model of
the thesis is Lambert/Phong over a surface normal (reaction/report/voxel.tex:
97-98). The definition restated by the oracle:
  * n = outward normal of the face through which the ray entered the hit cell;
  * k = 0.1 + 0.9 * max(0, n.L) when lit, 0.1 when a shadow ray is occluded;
    L = -dir for the headlight (light at the eye), the light direction
    otherwise;
  * channel = floor(c * k + 0.5), alpha 255;
  * the shadow ray starts at the hit point moved half a cell out of the hit
    cell through its entry face and runs toward the light; any leaf in front of
    its origin occludes.
Nothing below restates the oracle's float code. The expectations come from
float64 geometry computed independently here:
  * closed forms: axis-aligned rays onto known faces, and each of the six faces
    under lights of known angle;
  * brute force: the hit box of every ray by exhaustive float64 slab tests, its
    entry face as the argmax of the per-axis entry times, and the shadow ray's
    occlusion as an exhaustive float64 search over every leaf box from the
    stepped origin. Rays whose float64 decision sits within eps of a tie (edges,
    grazing boxes) are left out, since float32 cannot separate them;
  * planted mistakes (ORACLE_PLANT 12: light sign flipped; 13: shadow origin
    stepped into the hit voxel; 14: shadow ray cast away from the light) must
    each fail these pins.
"""
import math

import numpy as np
import pytest

import svo_helpers as H

AXES = np.eye(3)
EPS = 1e-5


def _channels(rgba):
    rgba = np.asarray(rgba, np.uint64)
    return np.stack([(rgba >> np.uint64(8 * c)) & np.uint64(255) for c in range(4)], axis=-1).astype(np.float64)


def _assert_shade(rgba, leaf, k, what):
    """Each RGB channel of `rgba` is floor(c*k + 0.5) up to float32 rounding of
    k: within 0.5 (+1e-3) of the float64 product c*k; alpha 255."""
    got = _channels(rgba)
    lc = _channels(leaf)[..., :3]
    exp = lc * np.asarray(k, np.float64)[..., None]
    err = np.abs(got[..., :3] - exp)
    bad = np.nonzero(err.max(axis=-1) > 0.5 + 1e-3)[0]
    assert bad.size == 0, (f"{what}: {bad.size} rays shaded off the Lambert closed form, first "
                           f"got {got[bad[0], :3]} expected {exp[bad[0]]}")
    assert np.all(got[..., 3] == 255), f"{what}: alpha of a hit is not 255"


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _boxes(keys):
    keys = np.asarray(keys, np.float64)
    scale = 2.0 ** (-keys[:, 0])
    return keys[:, 1:4] * scale[:, None], (keys[:, 1:4] + 1.0) * scale[:, None]


def _slabs(lo, hi, p, d):
    """Per ray and box: per-axis entry times (R, B, 3), tin and tout (R, B)."""
    ta = (lo[None] - p[:, None, :]) / d[:, None, :]
    tb = (hi[None] - p[:, None, :]) / d[:, None, :]
    te = np.minimum(ta, tb)
    return te, te.max(axis=2), np.maximum(ta, tb).min(axis=2)


def _first_hit_faces(keys, pos, dirs, eps=EPS):
    """Brute-force first hit with its entry face.
    Returns idx (-1 on a miss), tin, entry axis, ambiguous."""
    idx, tin, amb = H.brute_first_hit(keys, pos, dirs, eps)
    lo, hi = _boxes(keys)
    n = dirs.shape[0]
    axis = np.full(n, -1, np.int64)
    for i in np.nonzero(idx >= 0)[0]:
        b = idx[i]
        p, d = pos[i], dirs[i]
        te = np.minimum((lo[b] - p) / d, (hi[b] - p) / d)
        order = np.argsort(te)
        axis[i] = order[2]
        scale = max(1.0, abs(te[order[2]]))
        if te[order[2]] - te[order[1]] < eps * scale:
            amb[i] = True  # the ray enters through an edge or a corner
    return idx, tin, axis, amb


def _brute_occluded(keys, origins, light, eps=EPS):
    """Exhaustive float64 occlusion of rays from `origins` along `light`: a leaf
    box occludes when part of it lies in front of the origin (tout > 0) and the
    ray crosses it (tout > tin). Returns occluded, ambiguous."""
    lo, hi = _boxes(keys)
    d = np.repeat(np.asarray(light, np.float64).reshape(1, 3), origins.shape[0], axis=0)
    occ = np.zeros(origins.shape[0], bool)
    amb = np.zeros(origins.shape[0], bool)
    for s in range(0, origins.shape[0], 128):
        _, tin, tout = _slabs(lo, hi, origins[s:s + 128], d[s:s + 128])
        span = tout - tin
        occ[s:s + 128] = np.any((span > eps) & (tout > eps), axis=1)
        amb[s:s + 128] = np.any((np.abs(span) <= eps) & (tout > -eps), axis=1) | \
            np.any((np.abs(tout) <= eps) & (span > -eps), axis=1)
    return occ, amb


def _orbit_rays(rng, n_rays, spread=0.45):
    az, el = rng.uniform(0, 2 * math.pi), rng.uniform(-1.2, 1.2)
    pos = np.array([0.5 + 1.6 * math.cos(el) * math.cos(az), 0.5 + 1.6 * math.cos(el) * math.sin(az),
                    0.5 + 1.6 * math.sin(el)], np.float32)
    tgt = 0.5 + rng.uniform(-spread, spread, size=(n_rays, 3))
    d = H.adjust(_unit(tgt - pos).astype(np.float32))
    return pos, d


def _random_light(rng):
    while True:
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        if np.min(np.abs(v)) > 0.05:  # keep clear of the AdjustDir clamp
            return v.astype(np.float32)


# ---- closed forms -------------------------------------------------------------

def test_headlight_axis_rays_return_the_leaf_colour(orc):
    # Six leaves around the centre of a depth-2 tree, each seen along an axis
    # from outside: the headlight is the face normal itself, n.L = 1, k = 1,
    # and the shaded colour is the leaf colour exactly.
    cols = [0x10FF8001, 0x2033CC66, 0x30AA0199, 0x40FEFDFC, 0x50123456, 0x60808080]
    leaves = [(2, 0, 1, 1, cols[0]), (2, 3, 2, 1, cols[1]), (2, 1, 0, 2, cols[2]),
              (2, 2, 3, 2, cols[3]), (2, 2, 1, 0, cols[4]), (2, 1, 2, 3, cols[5])]
    nodes, root = H.tree_from_leaves(leaves)
    rays = [  # (origin, direction) onto the outer face of each leaf
        ((-1.0, 0.375, 0.375), (1, 0, 0)), ((2.0, 0.625, 0.375), (-1, 0, 0)),
        ((0.375, -1.0, 0.625), (0, 1, 0)), ((0.625, 2.0, 0.625), (0, -1, 0)),
        ((0.625, 0.375, -1.0), (0, 0, 1)), ((0.375, 0.625, 2.0), (0, 0, -1))]
    pos = np.array([r[0] for r in rays], np.float32)
    dirs = np.array([r[1] for r in rays], np.float32)
    out = orc.shade_rays(nodes, root, pos, dirs, light_mode=0)
    for i in range(6):
        assert out["hit_id"][i] != H.MISS
        assert out["rgba"][i] == (cols[i] & 0x00FFFFFF) | 0xFF000000, f"axis ray {i}"
        assert out["t"][i] == np.float32(1.0)
    # a directional light along the face normal gives the same exact colour;
    # one from behind the face leaves the ambient 0.1
    for i in range(6):
        n = -dirs[i]
        lit = orc.shade_rays(nodes, root, pos[i:i + 1], dirs[i:i + 1], light_mode=1, light=n)
        assert lit["rgba"][0] == (cols[i] & 0x00FFFFFF) | 0xFF000000
        dark = orc.shade_rays(nodes, root, pos[i:i + 1], dirs[i:i + 1], light_mode=1, light=-n)
        _assert_shade(dark["rgba"], np.array([cols[i]], np.uint32), np.array([0.1]), "back light")


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("sign", [1, -1])
def test_directional_light_on_each_face(orc, axis, sign):
    # One level-1 leaf; rays cross its face on `axis` from the `sign` side with a
    # small tilt, so they enter through that face only. The outward normal is
    # n = -sign * e_axis and k = 0.1 + 0.9 * max(0, n.L) for lights of every
    # angle, both sides of the face included.
    col = 0x00C8641E
    nodes, root = H.tree_from_leaves([(1, 1, 0, 1, col)])
    lo = np.array([0.5, 0.0, 0.5])
    centre = lo + 0.25
    rng = np.random.default_rng(40 + 3 * axis + sign)
    d = np.zeros((32, 3))
    d[:, axis] = sign
    other = [a for a in range(3) if a != axis]
    d[:, other] = rng.uniform(-0.15, 0.15, size=(32, 2))
    d = _unit(d)
    # aim at points of the face, well inside it
    face = np.repeat(centre[None], 32, axis=0)
    face[:, axis] = lo[axis] + (0.0 if sign > 0 else 0.5)
    face[:, other] += rng.uniform(-0.15, 0.15, size=(32, 2))
    pos = (face - 1.5 * d).astype(np.float32)
    dirs = d.astype(np.float32)
    n = -sign * AXES[axis]
    for trial in range(6):
        L = _random_light(rng) if trial > 1 else (n if trial == 0 else -n).astype(np.float32)
        out = orc.shade_rays(nodes, root, pos, dirs, light_mode=1, light=L)
        assert np.all(out["hit_id"] == root * 8 + 5), "rays must hit the leaf"
        ndl = float(n @ L.astype(np.float64))
        k = np.full(32, 0.1 + 0.9 * max(0.0, ndl))
        _assert_shade(out["rgba"], np.full(32, col, np.uint32), k, f"face {axis},{sign}, n.L={ndl:.3f}")
    # headlight: L = -dir, so n.L = |d_axis|
    out = orc.shade_rays(nodes, root, pos, dirs, light_mode=0)
    k = 0.1 + 0.9 * np.maximum(0.0, -(dirs.astype(np.float64) @ n))
    _assert_shade(out["rgba"], np.full(32, col, np.uint32), k, "headlight")


# ---- brute force ---------------------------------------------------------------

def _random_leaves(rng, depth, density):
    n = 1 << depth
    occ = rng.random((n, n, n)) < density
    return H.dense_leaves(occ, rng)


def _check_shading(orc, leaves, rng, light_mode, n_rays=300, views=3):
    """Shade rays against the float64 brute force. Returns (rays checked,
    lit shadow rays, occluded shadow rays)."""
    nodes, root = H.tree_from_leaves(leaves, order="shuffle", seed=int(rng.integers(1 << 30)))
    keys = [(l[0], l[1], l[2], l[3]) for l in leaves]
    pay = np.array([l[4] for l in leaves], np.uint32)
    checked = lit_n = occ_n = 0
    for _ in range(views):
        pos, dirs = _orbit_rays(rng, n_rays)
        L = _random_light(rng)
        out = orc.shade_rays(nodes, root, pos, dirs, light_mode=light_mode, light=L)
        p64 = np.repeat(pos.astype(np.float64)[None], n_rays, axis=0)
        d64 = dirs.astype(np.float64)
        idx, tin, axis, amb = _first_hit_faces(keys, p64, d64)
        hit = idx >= 0
        assert np.array_equal((out["hit_id"] != H.MISS)[~amb], hit[~amb])
        sel = np.nonzero(hit & ~amb)[0]
        # outward normal of the entry face: against the ray along that axis
        nrm = np.zeros((sel.size, 3))
        nrm[np.arange(sel.size), axis[sel]] = -np.sign(d64[sel, axis[sel]])
        Lv = -d64[sel] if light_mode == 0 else np.repeat(L.astype(np.float64)[None], sel.size, axis=0)
        ndl = np.einsum("ij,ij->i", nrm, Lv)
        keep = np.abs(ndl) > 1e-6
        lam = np.maximum(0.0, ndl)
        visible = np.ones(sel.size, bool)
        if light_mode == 2:
            casts = ndl > 1e-6
            lv = np.asarray(keys, np.float64)[idx[sel], 0]
            origin = p64[sel] + d64[sel] * tin[sel][:, None] + nrm * (2.0 ** -(lv + 1))[:, None]
            occ, samb = _brute_occluded(keys, origin[casts], L)
            visible[np.nonzero(casts)[0]] = ~occ
            keep[np.nonzero(casts)[0][samb]] = False
            exp_flag = np.where(casts, np.where(visible, 0, 1), 2)
            got_flag = out["shadow"][sel]
            bad = np.nonzero(keep & (got_flag != exp_flag))[0]
            assert bad.size == 0, (f"{bad.size} shadow flags differ from the brute force, first ray "
                                   f"{sel[bad[0]]}: got {got_flag[bad[0]]}, expected {exp_flag[bad[0]]}")
            lit_n += int((keep & casts & visible).sum())
            occ_n += int((keep & casts & ~visible).sum())
            assert np.all(out["shadow"][~hit & ~amb] == 2)
        k = np.where(visible, 0.1 + 0.9 * lam, 0.1)
        s = sel[keep]
        _assert_shade(out["rgba"][s], pay[idx[s]], k[keep], f"light mode {light_mode}")
        checked += int(s.size)
    assert checked > 0.3 * views * n_rays
    return checked, lit_n, occ_n


@pytest.mark.parametrize("depth", [3, 4, 5])
@pytest.mark.parametrize("light_mode", [0, 1])
def test_entry_face_lambert_brute_force(orc, depth, light_mode):
    rng = np.random.default_rng(500 + 10 * depth + light_mode)
    _check_shading(orc, _random_leaves(rng, depth, 0.25 if depth < 5 else 0.08), rng, light_mode)


@pytest.mark.parametrize("depth,density", [(3, 0.2), (4, 0.1), (5, 0.04)])
def test_shadow_rays_brute_force(orc, depth, density):
    rng = np.random.default_rng(600 + depth)
    _, lit, occ = _check_shading(orc, _random_leaves(rng, depth, density), rng, 2, views=4)
    # both outcomes occur often enough for the pin to bite
    assert lit > 50 and occ > 50, (lit, occ)


def test_shadow_rays_on_the_synthetic_surface(orc, yx):
    # the voxelizer's own surface-only tree at depth 5 (no interior leaves)
    s = yx.SVOData.build_synthetic(depth=5, num_spheres=6, radius_min=2.0, radius_max=6.0,
                                   heightfield=True, seed=9)
    leaves = [(k[0], k[1], k[2], k[3], v[0]) for k, v in H.walk_leaves(s.nodes(), s.GetRoot()).items()]
    rng = np.random.default_rng(61)
    _, lit, occ = _check_shading(orc, leaves, rng, 2, views=3)
    assert lit > 20 and occ > 20, (lit, occ)


# ---- planted mistakes ------------------------------------------------------------

def _shade_pins(o):
    failures = 0
    checks = [
        lambda: test_headlight_axis_rays_return_the_leaf_colour(o),
        lambda: test_directional_light_on_each_face(o, 0, 1),
        lambda: test_directional_light_on_each_face(o, 2, -1),
        lambda: _check_shading(o, _random_leaves(np.random.default_rng(71), 4, 0.25),
                               np.random.default_rng(72), 1, views=2),
        lambda: _check_shading(o, _random_leaves(np.random.default_rng(73), 4, 0.1),
                               np.random.default_rng(74), 2, views=2),
    ]
    for c in checks:
        try:
            c()
        except AssertionError:
            failures += 1
    return failures


@pytest.mark.parametrize("plant", [12, 13, 14])
def test_planted_shading_mistakes_are_caught(oracle_mod, tmp_path, plant):
    lib = oracle_mod.build(plant=plant, out=tmp_path / f"liboracle_plant{plant}.so")
    bad = oracle_mod.Oracle(lib)
    assert bad.plant() == plant
    assert _shade_pins(bad) > 0, f"planted mistake {plant} passed every shading pin"


def test_unplanted_oracle_passes_the_shading_pins(orc):
    assert _shade_pins(orc) == 0
