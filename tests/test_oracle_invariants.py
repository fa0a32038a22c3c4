"""More pins of the SVO oracle (oracle/svo_oracle.c), from the invariants and
the textbook routine the traversal must agree with:

* mirror symmetry: mirroring the scene along an axis (child slot c -> c ^ bit)
  and the ray with it (p -> 1 - p, d -> -d) mirrors the hit voxel and leaves
  t, the hit level and the number of node visits unchanged, bit for bit. The
  slab setup of a negative direction, (1 - p) / d, is exactly the mirrored
  ray's (-p') / d', and FindFirstChild / GoNext compare the same values;
* Amanatides-Woo: on dense grids the first leaf equals a float64 3D DDA's first
  occupied voxel for 10^5 rays, among them axis-parallel, diagonal and
  edge-grazing ones. Rays the float32 traversal cannot separate are left out:
  a step whose two nearest boundaries are within 1e-5 in t, or an entry point
  within 1e-4 voxels of a boundary of another axis.
"""
import numpy as np
import pytest

import svo_helpers as H


def mirror_nodes(nodes, axis_bits):
    """Mirror a VoxNode array: slot c moves to c ^ axis_bits."""
    out = nodes.copy()
    for c in range(8):
        m = c ^ axis_bits
        out[:, 2 + m] = nodes[:, 2 + c]
    flags = nodes[:, 0].astype(np.uint64)
    nf = flags & ~np.uint64(0xFFFF)
    for c in range(8):
        m = c ^ axis_bits
        nf |= ((flags >> np.uint64(c)) & np.uint64(1)) << np.uint64(m)
        nf |= ((flags >> np.uint64(8 + c)) & np.uint64(1)) << np.uint64(8 + m)
    out[:, 0] = nf.astype(np.uint32)
    return out


def dyadic_rays(rng, n):
    """Origins whose mirror 1 - p is exact (multiples of 1/64 in [-1, 2]) and
    directions with no zero component."""
    pos = (rng.integers(-64, 129, size=(n, 3)) / 64.0).astype(np.float32)
    tgt = rng.uniform(0.05, 0.95, size=(n, 3))
    d = tgt - pos
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    d[np.abs(d) < 1e-3] = np.float32(1e-3)
    return pos, d


@pytest.mark.parametrize("bits", [1, 2, 4, 7])
def test_mirror_symmetry(orc, bits):
    rng = np.random.default_rng(70 + bits)
    cases = []
    for depth in (3, 5):
        n = 1 << depth
        occ = rng.random((n, n, n)) < 0.15
        cases.append(H.tree_from_leaves(H.dense_leaves(occ, rng), order="shuffle", seed=depth))
    for nodes, root in cases:
        mn = mirror_nodes(nodes, bits)
        pos, d = dyadic_rays(rng, 3000)
        a = orc.trace_rays(nodes, root, pos, d)
        mpos, md = pos.copy(), d.copy()
        for ax in range(3):
            if (bits >> ax) & 1:
                mpos[:, ax] = np.float32(1.0) - pos[:, ax]
                md[:, ax] = -d[:, ax]
        b = orc.trace_rays(mn, root, mpos, md)
        hit = a["hit_id"] != H.MISS
        assert hit.sum() > 300
        assert np.array_equal(hit, b["hit_id"] != H.MISS)
        assert np.array_equal(a["t"], b["t"])
        assert np.array_equal(a["level"], b["level"])
        assert np.array_equal(a["visits"], b["visits"])
        assert np.array_equal(a["hit_id"][hit] // 8, b["hit_id"][hit] // 8)
        assert np.array_equal((a["hit_id"][hit] % 8) ^ bits, b["hit_id"][hit] % 8)
        va, vb = a["voxel"][hit], b["voxel"][hit]
        size = (1 << a["level"][hit]) - 1
        for ax in range(3):
            exp = size - va[:, ax] if (bits >> ax) & 1 else va[:, ax]
            assert np.array_equal(vb[:, ax], exp)


def dda_first(occ, pos, d, eps=1e-5):
    """Float64 Amanatides-Woo over a dense n^3 grid in the unit cube. Returns
    (voxel or -1, ambiguous) per ray."""
    n = occ.shape[0]
    m = pos.shape[0]
    pos = pos.astype(np.float64)
    d = d.astype(np.float64)
    inv = 1.0 / d
    tin = np.stack([np.minimum(-pos[:, a] * inv[:, a], (1 - pos[:, a]) * inv[:, a]) for a in range(3)], 1)
    entry_axis = np.argmax(tin, axis=1)
    t0 = tin.max(axis=1)
    t1 = np.minimum.reduce([np.maximum(-pos[:, a] * inv[:, a], (1 - pos[:, a]) * inv[:, a]) for a in range(3)])
    inside = t1 > np.maximum(t0, 0) + eps
    t0 = np.maximum(t0, 0)
    p = pos + d * (t0[:, None] + 1e-12)
    v = np.clip(np.floor(p * n).astype(np.int64), 0, n - 1)
    step = np.where(d > 0, 1, -1)
    nxt = np.where(d > 0, (v + 1) / n, v / n)
    tmax = (nxt - pos) * inv
    tdel = np.abs(1.0 / (d * n))
    res = np.full((m, 3), -1, np.int64)
    amb = np.zeros(m, bool)
    active = inside.copy()
    # the entry voxel is ambiguous when the entry point lies on a voxel boundary
    # of an axis other than the entry face's
    frac = p * n - np.floor(p * n)
    near = (frac < 1e-4) | (frac > 1 - 1e-4)
    near[np.arange(m), entry_axis] = False
    amb |= inside & np.any(near, axis=1)
    for _ in range(3 * n + 3):
        idx = np.nonzero(active)[0]
        if idx.size == 0:
            break
        vv = v[idx]
        hit = occ[vv[:, 0], vv[:, 1], vv[:, 2]]
        res[idx[hit]] = vv[hit]
        active[idx[hit]] = False
        idx = idx[~hit]
        tm = tmax[idx]
        srt = np.sort(tm, axis=1)
        amb[idx] |= (srt[:, 1] - srt[:, 0]) < eps  # the next step crosses an edge or corner
        ax = np.argmin(tm, axis=1)
        v[idx, ax] += step[idx, ax]
        tmax[idx, ax] += tdel[idx, ax]
        out = np.any((v[idx] < 0) | (v[idx] >= n), axis=1)
        active[idx[out]] = False
    return res, amb


@pytest.mark.parametrize("depth", [3, 4, 5])
def test_dda_hundred_thousand_rays(orc, depth):
    rng = np.random.default_rng(90 + depth)
    n = 1 << depth
    occ = rng.random((n, n, n)) < 0.08
    leaves = H.dense_leaves(occ, rng)
    nodes, root = H.tree_from_leaves(leaves, order="bfs")
    m = 100_000
    pos = np.empty((m, 3), np.float32)
    d = np.empty((m, 3), np.float32)
    # random rays from outside
    k = m // 2
    az, el = rng.uniform(0, 2 * np.pi, k), rng.uniform(-1.2, 1.2, k)
    pos[:k] = np.stack([0.5 + 1.7 * np.cos(el) * np.cos(az), 0.5 + 1.7 * np.cos(el) * np.sin(az),
                        0.5 + 1.7 * np.sin(el)], 1)
    tgt = rng.uniform(0.02, 0.98, size=(k, 3))
    dd = tgt - pos[:k]
    d[:k] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    # axis-parallel, face-diagonal and space-diagonal rays entering on grid lines
    j = m - k
    axes = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [0, 1, 1], [1, 0, 1], [1, 1, 1], [1, -1, 1]],
                    np.float64)
    pick = axes[rng.integers(0, len(axes), j)] * rng.choice([-1.0, 1.0], size=(j, 1))
    dirn = pick / np.linalg.norm(pick, axis=1, keepdims=True)
    start = rng.integers(0, n + 1, size=(j, 3)) / n + rng.choice([0.0, 0.5 / n], size=(j, 3))
    pos[k:] = start - dirn * 2.0
    d[k:] = dirn
    d = H.adjust(d)
    got = orc.trace_rays(nodes, root, pos, d)
    exp, amb = dda_first(occ, pos, d)
    ok = ~amb
    assert ok[:k].sum() > 0.9 * k and ok.sum() > 0.5 * m
    g_hit = got["hit_id"] != H.MISS
    e_hit = exp[:, 0] >= 0
    assert np.array_equal(g_hit[ok], e_hit[ok])
    sel = ok & e_hit
    assert np.array_equal(got["voxel"][sel], exp[sel])
    assert np.all(got["level"][sel] == depth)


@pytest.mark.parametrize("plant", [3, 4, 5])
def test_planted_mistakes_fail_these_pins(oracle_mod, tmp_path, plant):
    """The dropped octant mirror (3), a wrong midpoint (4) and a wrong slab
    setup (5) each fail the mirror or DDA pin. Plants 1 and 2 change exact
    ties only; the hand-derived tie cases of test_oracle_pins.py catch them."""
    bad = oracle_mod.Oracle(oracle_mod.build(plant=plant, out=tmp_path / f"liboracle_plant{plant}.so"))
    caught = 0
    for pin in (lambda: test_mirror_symmetry(bad, 7), lambda: test_dda_hundred_thousand_rays(bad, 4)):
        try:
            pin()
        except AssertionError:
            caught += 1
    assert caught > 0
