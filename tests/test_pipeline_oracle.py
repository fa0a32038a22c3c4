"""Pins for the CPU restatement of the shading pipeline (SVORenderer::Render,
demo/SVORenderer.cpp:95-149): LOD colours and the LOD stop (svo_oracle.c), the
depth-limited blur and Phong shading from z (ssna_oracle.c).

The reference's pipeline kernels are missing and it ships no outputs of them,
so each piece is pinned by what the mathematics fixes:
  * LOD colours: an independent numpy mean over every subtree's leaves, and a
    uniform-colour tree whose every node mean is that colour;
  * LOD stop: detail 0 reduces to the exact trace, and a float64 brute force
    of the first "terminal" cell (a leaf, or a node whose edge is smaller than
    its entry distance times detail, below no such node) along each ray;
  * frame constants: their closed forms in float64;
  * blur: the kernel table against exp(-(tx^2+ty^2)) normalised; constant and
    linear depth, steps wider than the edge threshold and far pixels are left
    as they are; a noisy surface against a float64 evaluation;
  * shading: a plane facing the camera lit from the eye, whose colour has a
    closed form per pixel, and the normal of a tilted depth ramp;
  * planted mistakes (ORACLE_PLANT 9, 10, 11) each fail one of these pins.
"""
import math

import numpy as np
import pytest

import svo_helpers as H


@pytest.fixture(scope="module")
def po(oracle_mod):
    return oracle_mod.PipelineOracle()


# ---- LOD colours ---------------------------------------------------------------

def _node_keys(nodes, root):
    """(level, x, y, z) -> node id for every node reachable from root."""
    res = {}
    stack = [(int(root), 0, 0, 0, 0)]
    while stack:
        nid, lv, x, y, z = stack.pop()
        res[(lv, x, y, z)] = nid
        n = nodes[nid]
        for s in range(8):
            if not ((int(n[0]) >> s) & 1) and not (int(n[2 + s]) & 0x80000000):
                stack.append((int(n[2 + s]), lv + 1, 2 * x + (s & 1), 2 * y + ((s >> 1) & 1), 2 * z + ((s >> 2) & 1)))
    return res


def _numpy_lod(nodes, root):
    leaves = H.walk_leaves(nodes, root)
    acc = {}
    for (lv, x, y, z), (pay, _, _) in leaves.items():
        ch = np.array([(pay >> (8 * k)) & 255 for k in range(4)], np.int64)
        for a in range(lv):
            key = (a, x >> (lv - a), y >> (lv - a), z >> (lv - a))
            s, n = acc.get(key, (np.zeros(4, np.int64), 0))
            acc[key] = (s + ch, n + 1)
    out = {}
    for key, (s, n) in acc.items():
        m = (s + n // 2) // n
        out[key] = int(m[0] | (m[1] << 8) | (m[2] << 16) | (m[3] << 24))
    return out


@pytest.mark.parametrize("depth", [2, 4])
def test_lod_colors_are_subtree_means(po, depth):
    rng = np.random.default_rng(40 + depth)
    n = 1 << depth
    occ = rng.random((n, n, n)) < 0.3
    nodes, root = H.tree_from_leaves(H.dense_leaves(occ, rng), order="shuffle", seed=depth)
    got = po.lod_colors(nodes, root)
    exp = _numpy_lod(nodes, root)
    keys = _node_keys(nodes, root)
    assert len(keys) > 1
    for key, nid in keys.items():
        assert got[nid] == exp[key], key


def test_lod_colors_uniform_tree(po, yx):
    s = yx.SVOData.build_synthetic(depth=5, num_spheres=4, radius_min=2.0, radius_max=5.0, seed=2)
    nodes = s.nodes().copy()
    root = s.GetRoot()
    for nid in range(nodes.shape[0]):
        for sl in range(8):
            if (int(nodes[nid, 0]) >> sl) & 1:
                nodes[nid, 2 + sl] = 0x80C0FF10
    got = po.lod_colors(nodes, root)
    ids = list(_node_keys(nodes, root).values())
    assert np.all(got[ids] == 0x80C0FF10)


# ---- LOD stop ------------------------------------------------------------------

def _cells(nodes, root):
    """Node cells (level, x, y, z, id, parent index) in level order, and leaf
    cells (level, x, y, z, node id, slot, parent index)."""
    ncell, leaves = [], []
    q = [(int(root), 0, 0, 0, 0, -1)]
    while q:
        nq = []
        for (nid, lv, x, y, z, par) in q:
            me = len(ncell)
            ncell.append((lv, x, y, z, nid, par))
            n = nodes[nid]
            for s in range(8):
                cx, cy, cz = 2 * x + (s & 1), 2 * y + ((s >> 1) & 1), 2 * z + ((s >> 2) & 1)
                if (int(n[0]) >> s) & 1:
                    leaves.append((lv + 1, cx, cy, cz, nid, s, me))
                elif not (int(n[2 + s]) & 0x80000000):
                    nq.append((int(n[2 + s]), lv + 1, cx, cy, cz, me))
        q = nq
    return ncell, leaves


def _slab(keys, pos, dirs):
    scale = 2.0 ** (-keys[:, 0])
    lo = keys[:, 1:4] * scale[:, None]
    hi = (keys[:, 1:4] + 1.0) * scale[:, None]
    ta = (lo[None] - pos[None, None, :]) / dirs[:, None, :]
    tb = (hi[None] - pos[None, None, :]) / dirs[:, None, :]
    return np.minimum(ta, tb).max(axis=2), np.maximum(ta, tb).min(axis=2)


def brute_lod(nodes, root, pos, dirs, detail, eps=1e-5):
    """First terminal cell per ray in float64: ('n', node id, level) for an LOD
    stop, ('l', parent id, slot, level) for a leaf, None for a miss; plus an
    ambiguity flag (near ties of boxes or of the stop rule)."""
    ncell, leaves = _cells(nodes, root)
    nk = np.array([c[:4] for c in ncell], np.float64)
    lk = np.array([c[:4] for c in leaves], np.float64)
    npar = np.array([c[5] for c in ncell])
    lpar = np.array([c[6] for c in leaves])
    pos = np.asarray(pos, np.float64)
    dirs = np.asarray(dirs, np.float64)
    ntin, ntout = _slab(nk, pos, dirs)
    ltin, ltout = _slab(lk, pos, dirs)
    ncross = ntout - ntin > eps
    edge = 2.0 ** (-nk[:, 0])
    lhs, rhs = edge[None, :], ntin * detail
    stop = ncross & (lhs < rhs) & (nk[None, :, 0] > 0)
    amb_rule = ncross & (np.abs(lhs - rhs) < 1e-4 * lhs) & (nk[None, :, 0] > 0)
    anc = np.zeros_like(stop)
    for i in range(1, len(ncell)):  # parents come first (level order)
        p = npar[i]
        anc[:, i] = anc[:, p] | stop[:, p]
    term_n = stop & ~anc
    term_l = ~(anc[:, lpar] | stop[:, lpar])
    tin_all = np.concatenate([np.where(term_n & ncross, ntin, np.inf),
                              np.where(term_l & (ltout - ltin > eps), ltin, np.inf)], axis=1)
    graze = np.concatenate([np.where(term_n & ~ncross & (ntout - ntin > -eps), ntin, np.inf),
                            np.where(term_l & (np.abs(ltout - ltin) <= eps), ltin, np.inf)], axis=1)
    best = tin_all.argmin(axis=1)
    bt = tin_all[np.arange(best.size), best]
    t2 = tin_all.copy()
    t2[np.arange(best.size), best] = np.inf
    with np.errstate(invalid="ignore"):
        amb = (t2.min(axis=1) - bt < eps) | (graze.min(axis=1) < bt + eps) | amb_rule.any(axis=1)
    res = []
    for r in range(dirs.shape[0]):
        if not np.isfinite(bt[r]):
            res.append(None)
        elif best[r] < len(ncell):
            c = ncell[best[r]]
            res.append(("n", c[4], c[0]))
        else:
            c = leaves[best[r] - len(ncell)]
            res.append(("l", c[4], c[5], c[0]))
    return res, amb, bt


def _check_lod(po, nodes, root, rng, detail, n_rays=300, views=3):
    checked = stops = 0
    for _ in range(views):
        az, el = rng.uniform(0, 2 * math.pi), rng.uniform(-1.0, 1.0)
        pos = np.array([0.5 + 1.7 * math.cos(el) * math.cos(az), 0.5 + 1.7 * math.cos(el) * math.sin(az),
                        0.5 + 1.7 * math.sin(el)], np.float32)
        tgt = 0.5 + rng.uniform(-0.45, 0.45, size=(n_rays, 3))
        d = tgt - pos.astype(np.float64)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        dirs = H.adjust(d.astype(np.float32))
        got = po.lod_trace_rays(nodes, root, pos, dirs, detail)
        exp, amb, bt = brute_lod(nodes, root, pos, dirs, float(np.float32(detail)))
        for r in range(n_rays):
            if amb[r]:
                continue
            checked += 1
            e = exp[r]
            if e is None:
                assert got["node"][r] == H.MISS, r
            elif e[0] == "n":
                stops += 1
                assert (got["node"][r], got["child"][r], got["level"][r]) == (e[1], -1, e[2]), (r, e)
                np.testing.assert_allclose(got["t"][r], bt[r], rtol=1e-5, atol=1e-6)
            else:
                assert (got["node"][r], got["child"][r], got["level"][r]) == (e[1], e[2], e[3]), (r, e)
                np.testing.assert_allclose(got["t"][r], bt[r], rtol=1e-5, atol=1e-6)
    assert checked > 0.85 * views * n_rays
    return stops


@pytest.mark.parametrize("depth,detail", [(4, 0.1), (5, 0.05), (5, 0.15)])
def test_lod_stop_brute_force(po, depth, detail):
    rng = np.random.default_rng(500 + depth)
    n = 1 << depth
    occ = rng.random((n, n, n)) < 0.15
    nodes, root = H.tree_from_leaves(H.dense_leaves(occ, rng), order="shuffle", seed=depth)
    stops = _check_lod(po, nodes, root, rng, detail)
    assert stops > 20  # the rule is exercised, not only leaf hits


def test_lod_zero_detail_is_exact_trace(po, orc, yx):
    s = yx.SVOData.build_synthetic(depth=6, num_spheres=8, radius_min=3.0, radius_max=10.0, heightfield=True, seed=4)
    nodes, root = s.nodes(), s.GetRoot()
    rng = np.random.default_rng(3)
    pos = np.array([1.8, 1.2, 1.4], np.float32)
    tgt = 0.5 + rng.uniform(-0.45, 0.45, size=(2000, 3))
    d = tgt - pos
    dirs = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    a = po.lod_trace_rays(nodes, root, pos, dirs, 0.0)
    b = orc.trace_rays(nodes, root, pos, dirs)
    hit = b["hit_id"] != H.MISS
    assert hit.sum() > 200
    assert np.array_equal(a["node"][hit] * 8 + a["child"][hit].astype(np.uint32), b["hit_id"][hit])
    assert np.array_equal(a["t"], b["t"]) and np.array_equal(a["level"], b["level"])
    assert np.all(a["node"][~hit] == H.MISS)


# ---- frame constants, blur -----------------------------------------------------

def test_params_closed_forms(po):
    fov, w, depth, coef = 70.0, 1920, 11, 1.5
    p = po.params(fov, w, depth, coef)
    half = math.radians(fov / 2)
    assert p["detail"] == pytest.approx(coef * half / w, rel=1e-6)
    assert p["pixel_ang"] == pytest.approx(math.radians(fov) / w, rel=1e-6)
    assert p["pix"] == pytest.approx(2 * math.tan(half) / w, rel=1e-6)
    assert p["vox"] == 2.0 ** -11 and p["edge"] == 4 * 2.0 ** -11
    for i, zl in enumerate(p["zlimits"]):
        assert zl == pytest.approx(5 * 2.0 ** -11 / (math.radians(fov) / w * (3 + 3 * i)), rel=1e-6)


def test_blur_kernel_table(po):
    k = po.blur_kernel().reshape(5, 5)
    t = 2.0 * (np.arange(5) - 2.0) / 2.0
    g = np.exp(-(t[:, None] ** 2 + t[None, :] ** 2))
    np.testing.assert_allclose(k, g / g.sum(), rtol=1e-6)
    assert np.array_equal(k, k.T) and np.array_equal(k, k[::-1]) and np.array_equal(k, k[:, ::-1])


def _blur64(z, zlimit, edge, k):
    h, w = z.shape
    k = k.reshape(5, 5).astype(np.float64)
    zp = np.full((h + 4, w + 4), np.inf)
    zp[2:-2, 2:-2] = z
    num = np.zeros((h, w))
    den = np.zeros((h, w))
    with np.errstate(invalid="ignore"):
        for dy in range(5):
            for dx in range(5):
                zq = zp[dy:dy + h, dx:dx + w]
                m = np.abs(zq - z) < edge
                num += np.where(m, k[dy, dx] * zq, 0.0)
                den += np.where(m, k[dy, dx], 0.0)
        out = np.where(z < zlimit, num / den, z)
    return out


def test_blur_invariants(po):
    h, w = 40, 56
    k = po.blur_kernel()
    edge, zlimit = 0.01, 1.0
    yy, xx = np.mgrid[:h, :w]
    # constant and linear depth (steps between window pixels below the edge)
    for z in (np.full((h, w), 0.5, np.float32), (0.4 + 0.0009 * xx - 0.0006 * yy).astype(np.float32)):
        out = po.blur_z(z, zlimit, edge, k)
        inner = (slice(2, -2), slice(2, -2))
        np.testing.assert_allclose(out[inner], z[inner], rtol=2e-6)
        if np.all(z == z[0, 0]):
            np.testing.assert_allclose(out, z, rtol=2e-6)  # borders too: renormalised
    # a step wider than the edge: each side keeps its own depth
    z = np.where(xx < w // 2, 0.3, 0.6).astype(np.float32)
    np.testing.assert_allclose(po.blur_z(z, zlimit, edge, k), z, rtol=2e-6)
    # pixels at or beyond zlimit, and misses, are copied bit for bit
    z = (0.9 + 0.2 * np.sin(xx / 5.0) * np.cos(yy / 7.0)).astype(np.float32)
    z[5:9, 10:30] = np.inf
    out = po.blur_z(z, zlimit, edge, k)
    far = ~(z < zlimit)
    assert np.array_equal(out[far], z[far])
    assert np.any(out[~far] != z[~far])


def test_blur_noisy_surface_float64(po):
    rng = np.random.default_rng(11)
    h, w = 48, 64
    yy, xx = np.mgrid[:h, :w]
    z = 0.6 + 0.05 * np.sin(xx / 9.0) + 0.002 * rng.standard_normal((h, w))
    z[rng.random((h, w)) < 0.03] = np.inf
    z[:, 40:] += 0.03  # an edge
    z = z.astype(np.float32)
    k = po.blur_kernel()
    for zl in (0.62, 1.0):
        out = po.blur_z(z, zl, 0.004, k)
        exp = _blur64(z.astype(np.float64), zl, 0.004, k)
        fin = np.isfinite(exp)
        assert np.array_equal(np.isfinite(out), fin)
        np.testing.assert_allclose(out[fin], exp[fin], rtol=2e-6)


# ---- shading -------------------------------------------------------------------

def _frontal_case(po, w=48, h=36, z0=0.8):
    eye = np.array([-0.5, 0.5, 0.5], np.float32)
    fwd, up = (1.0, 0.0, 0.0), (0.0, 0.0, 1.0)
    basis = po.basis(fwd, up, 70.0, w, h)
    frame = po.view_frame(fwd, up)
    pix = po.params(70.0, w, 8, 1.0)["pix"]
    z = np.full((h, w), z0, np.float32)
    rng = np.random.default_rng(5)
    albedo = (rng.integers(0, 2**24, size=(h, w)) | 0xFF000000).astype(np.uint32)
    return eye, fwd, up, basis, frame, pix, z, albedo


def test_shade_frontal_plane_closed_form(po, oracle_mod):
    eye, fwd, up, basis, frame, pix, z, albedo = _frontal_case(po)
    h, w = z.shape
    dif, spe = 0.7, 0.3
    lights = oracle_mod.make_lights([dict(pos=eye, diffuse=(dif,) * 3, specular=(spe,) * 3,
                                          attenuation=(1.0, 0.0, 0.0))])
    out = po.shade(z, albedo, basis, eye, frame, pix, lights)
    yy, xx = np.mgrid[:h, :w]
    b = basis.astype(np.float64)
    q = b[None, None, 0:3] + b[None, None, 3:6] * xx[..., None] + b[None, None, 6:9] * yy[..., None]
    ndl = q[..., 0] / np.linalg.norm(q, axis=-1)  # N = -forward, L = V = -q/|q|
    rv = np.maximum(2 * ndl * ndl - 1, 0.0)
    for c in range(3):
        a = ((albedo >> (8 * c)) & 255).astype(np.float64)
        v = np.minimum(a * (0.1 + dif * ndl) + 255 * spe * rv ** 10, 255) + 0.5
        got = ((out >> (8 * c)) & 255).astype(np.int64)
        assert np.abs(got - np.floor(v)).max() <= 1
    assert np.all(out >> 24 == 255)


def test_shade_normals_of_depth_ramp(po, oracle_mod):
    eye, fwd, up, basis, frame, pix, z, albedo = _frontal_case(po)
    h, w = z.shape
    lights = oracle_mod.make_lights([])
    flat = po.shade(z, albedo, basis, eye, frame, pix, lights, show_normals=True)
    # N = -forward = (-1, 0, 0) -> (0, 127.5, 127.5) + 0.5
    assert np.all((flat & 255) == 0) and np.all(((flat >> 8) & 255) == 128) and np.all(((flat >> 16) & 255) == 128)
    g = 0.002
    ramp = (z + g * np.arange(w)[None, :]).astype(np.float32)
    out = po.shade(ramp, albedo, basis, eye, frame, pix, lights, show_normals=True)
    # depth grows to the right: the normal facing the eye turns toward +right
    r = np.asarray(frame[0:3], np.float64)
    zc = ramp.astype(np.float64)
    nv = np.stack([-pix * g * zc, np.zeros_like(zc), pix * pix * zc * zc], -1)
    nv /= np.linalg.norm(nv, axis=-1, keepdims=True)
    F = np.asarray(frame, np.float64).reshape(3, 3)
    N = -(nv @ F)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    assert np.all(N @ r > 0)
    for c in range(3):
        exp = np.floor((N[..., c] * 0.5 + 0.5) * 255 + 0.5)
        assert np.abs(((out >> (8 * c)) & 255) - exp).max() <= 1


def test_shade_misses_and_pipeline_end_to_end(po, oracle_mod, yx):
    s = yx.SVOData.build_synthetic(depth=7, num_spheres=12, radius_min=4.0, radius_max=14.0, heightfield=True,
                                   seed=6)
    pos = (0.55, 0.48, 0.42)
    lights = oracle_mod.make_lights([dict(pos=pos, diffuse=(0.7,) * 3, specular=(0.3,) * 3,
                                          attenuation=(1.0, 0.0, 0.5))])
    out = po.pipeline(s.nodes(), s.GetRoot(), 7, pos, (0.3, 0.2, -0.25), w=96, h=64, detail_coef=2.0,
                      lights=lights)
    hit = out["t"] >= 0
    assert 0.2 < hit.mean() < 1.0
    assert np.all(out["rgba"][~hit] == 0) and np.all(out["rgba"][hit] >> 24 == 255)
    assert np.all(np.isinf(out["zb"][~hit]))
    assert np.any(out["zb"][hit] != out["z"][hit])  # near voxels were smoothed
    assert np.all(out["level"][hit] >= 1) and np.all(out["level"][hit] <= 7)


# ---- planted mistakes ----------------------------------------------------------

@pytest.mark.parametrize("plant", [9, 10, 11])
def test_pipeline_planted_mistakes_are_caught(oracle_mod, tmp_path, plant, yx):
    lib = oracle_mod.build(plant=plant, out=tmp_path / f"liboracle_plant{plant}.so")
    bad = oracle_mod.PipelineOracle(lib)
    assert bad.plant() == plant
    pins = [lambda: test_lod_stop_brute_force(bad, 4, 0.1), lambda: test_blur_invariants(bad),
            lambda: test_shade_frontal_plane_closed_form(bad, oracle_mod)]
    failures = 0
    for p in pins:
        try:
            p()
        except AssertionError:
            failures += 1
    assert failures > 0, f"planted mistake {plant} passed every pin"
