"""Pins for the brick n-tree restatement (oracle/ntree_oracle.c):
  * densities equal an independent numpy float32 restatement of the improved
    Perlin noise fBm (table form of the gradients);
  * the n-tree encoding, decoded in Python, holds exactly the non-empty bricks
    of the dense volume with their voxels, and one node per non-empty subtree;
  * the ray cast equals a float64 Amanatides-Woo voxel walk over the dense
    grid (every voxel the ray crosses, in order, composited front to back):
    first non-empty voxel exact, its entry t within 1e-5, RGBA8 within 1 LSB,
    on every ray that no near-tie makes ambiguous;
  * planted mistakes in the restatement (ORACLE_PLANT 7: an empty node cell
    skips its parent's cell; 8: colour not weighted by opacity) fail that pin."""
import numpy as np
import pytest

MASK = 0xFFFFFFFF


def small_desc(orc, size=64, levels=1, seed=5, threshold=0.4, frequency=4.0):
    return orc.desc(size, levels, 4, frequency, threshold, 300.0, seed)


def palette(seed=3):
    rng = np.random.default_rng(seed)
    p = rng.integers(0, 256, size=(256, 4)).astype(np.uint32)
    p[:, 3] = np.clip(30 + np.arange(256) * 2, 0, 255)
    p[0] = 0
    return (p[:, 0] | (p[:, 1] << 8) | (p[:, 2] << 16) | (p[:, 3] << 24)).astype(np.uint32)


# ---- independent float32 noise ------------------------------------------------

GRAD = [(1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1),
        (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1), (1, 1, 0), (0, -1, 1), (-1, 1, 0), (0, -1, -1)]


def _grad(h, x, y, z):
    out = np.zeros_like(x)
    for k, (gx, gy, gz) in enumerate(GRAD):
        m = (h & 15) == k
        terms = [t for g, t in ((gx, x), (gy, y), (gz, z)) if g != 0]
        signs = [g for g in (gx, gy, gz) if g != 0]
        a = terms[0][m] if signs[0] > 0 else -terms[0][m]
        b = terms[1][m] if signs[1] > 0 else -terms[1][m]
        out[m] = (a + b).astype(np.float32)
    return out


def noise3(P, x, y, z):
    f32 = np.float32
    fx, fy, fz = np.floor(x), np.floor(y), np.floor(z)
    X, Y, Z = fx.astype(np.int64) & 255, fy.astype(np.int64) & 255, fz.astype(np.int64) & 255
    xf, yf, zf = x - fx, y - fy, z - fz

    def fade(t):
        return t * t * t * (t * (t * f32(6) - f32(15)) + f32(10))

    def lerp(t, a, b):
        return a + t * (b - a)

    u, v, w = fade(xf), fade(yf), fade(zf)
    a = P[X] + Y
    aa, ab = P[a] + Z, P[a + 1] + Z
    b = P[X + 1] + Y
    ba, bb = P[b] + Z, P[b + 1] + Z
    one = f32(1)
    x1, y1, z1 = xf - one, yf - one, zf - one
    l1 = lerp(u, _grad(P[aa], xf, yf, zf), _grad(P[ba], x1, yf, zf))
    l2 = lerp(u, _grad(P[ab], xf, y1, zf), _grad(P[bb], x1, y1, zf))
    l3 = lerp(u, _grad(P[aa + 1], xf, yf, z1), _grad(P[ba + 1], x1, yf, z1))
    l4 = lerp(u, _grad(P[ab + 1], xf, y1, z1), _grad(P[bb + 1], x1, y1, z1))
    return lerp(w, lerp(v, l1, l2), lerp(v, l3, l4))


def density_numpy(d, P, x0, y0, z0, n):
    f32 = np.float32
    c = np.arange(n, dtype=np.float32)
    Z, Y, X = np.meshgrid(c + f32(z0), c + f32(y0), c + f32(x0), indexing="ij")
    inv = f32(1.0) / f32(d.size)
    fq = f32(d.frequency)
    px, py, pz = (X + f32(0.5)) * inv * fq, (Y + f32(0.5)) * inv * fq, (Z + f32(0.5)) * inv * fq
    f = np.zeros_like(px)
    amp, fr = f32(1), f32(1)
    for _ in range(d.octaves):
        f = f + amp * noise3(P, px * fr, py * fr, pz * fr)
        amp, fr = amp * f32(0.5), fr * f32(2)
    thr, sc = f32(d.threshold), f32(d.scale)
    q = np.minimum((f - thr) * sc, f32(254))
    return np.where(f < thr, 0, 1 + q.astype(np.int64)).astype(np.uint8)


def test_density_matches_numpy(oracle_mod):
    o = oracle_mod.NTreeOracle()
    d = small_desc(o, size=4096, seed=77, threshold=0.3, frequency=8.0)
    P = o.perm(77).astype(np.int64)
    for (x0, y0, z0) in [(0, 0, 0), (2000, 1333, 4080), (517, 3000, 12)]:
        a = o.density_box(d, x0, y0, z0, 16, 16, 16)
        b = density_numpy(d, P, x0, y0, z0, 16)
        assert np.array_equal(a, b)
    # the standard table reproduces Perlin's reference gradient set: 12 edges
    assert len({g for g in GRAD}) == 12


# ---- tree encoding --------------------------------------------------------------

def decode_brick(root, nodes, size, levels, bx, by, bz):
    R = (size // 8) >> (2 * levels)
    sh = 2 * levels
    idx = int(root[((bz >> sh) * R + (by >> sh)) * R + (bx >> sh)])
    for l in range(1, levels + 1):
        if idx == MASK:
            return None
        s = 2 * (levels - l)
        c = ((((bz >> s) & 3) * 4) + ((by >> s) & 3)) * 4 + ((bx >> s) & 3)
        n = nodes[idx]
        mask = int(n[0]) | (int(n[1]) << 32)
        if not (mask >> c) & 1:
            return None
        idx = int(n[2]) + bin(mask & ((1 << c) - 1)).count("1")
    return None if idx == MASK else idx


@pytest.mark.parametrize("size,levels", [(64, 1), (128, 2), (64, 0), (32, 0), (128, 1)])
def test_tree_encoding(oracle_mod, size, levels):
    o = oracle_mod.NTreeOracle()
    d = small_desc(o, size=size, levels=levels, seed=size + levels)
    dense = o.density_box(d, 0, 0, 0, size, size, size)
    root, nodes, bricks = o.build_dense(dense, levels)
    nb = size // 8
    blocks = dense.reshape(nb, 8, nb, 8, nb, 8).transpose(0, 2, 4, 1, 3, 5)  # [bz, by, bx, z, y, x]
    occ = blocks.reshape(nb, nb, nb, -1).any(axis=3)
    assert 0 < occ.sum() < occ.size
    seen = set()
    for bz in range(nb):
        for by in range(nb):
            for bx in range(nb):
                idx = decode_brick(root, nodes, size, levels, bx, by, bz)
                assert (idx is not None) == bool(occ[bz, by, bx])
                if idx is not None:
                    assert np.array_equal(bricks[idx].reshape(8, 8, 8), blocks[bz, by, bx])
                    seen.add(idx)
    assert seen == set(range(len(bricks)))
    # one node per non-empty subtree of levels 1..L
    expect = 0
    o_ = occ
    for _ in range(levels):
        n = o_.shape[0] // 4
        o_ = o_.reshape(n, 4, n, 4, n, 4).any(axis=(1, 3, 5))
        expect += int(o_.sum())
    assert len(nodes) == expect or (levels == 0 and expect == 0)


# ---- ray cast vs a float64 dense walk ---------------------------------------------

def brute_ray(dense, pal, o, d, eps=1e-7):
    """Voxel walk in float64 (Amanatides & Woo 1987). Returns (rgba float4,
    first voxel, first t, ambiguous)."""
    N = dense.shape[0]
    o = np.asarray(o, np.float64)
    d = np.asarray(d, np.float64)
    inv = 1.0 / d
    lo, hi = (0.0 - o) * inv, (1.0 - o) * inv
    te = max(np.minimum(lo, hi).max(), 0.0)
    tx = np.maximum(lo, hi).min()
    acc = np.zeros(4)
    if tx < 0 or te > tx:
        return acc, None, -1.0, False
    p = (o + te * d) * N
    near = np.abs(p - np.round(p)) < 1e-4
    if te > 0:
        near[int(np.argmax(np.minimum(lo, hi)))] = False  # the entry face itself
    amb = bool(near.any())
    v = np.clip(np.floor(p).astype(np.int64), 0, N - 1)
    step = np.where(d > 0, 1, -1)
    first, tfirst = None, -1.0
    t = te
    while True:
        dd = dense[v[2], v[1], v[0]]
        if dd:
            c = pal[dd]
            col = np.array([c & 255, (c >> 8) & 255, (c >> 16) & 255, c >> 24], np.float64) / 255.0
            col[:3] *= col[3]
            acc = acc + col * (1.0 - acc[3])
            if first is None:
                first, tfirst = tuple(int(q) for q in v), t
            if acc[3] > 0.99:
                if abs(acc[3] - 0.99) < 1e-5:
                    amb = True
                break
        planes = (v + (step > 0)) / N
        tm = (planes - o) * inv
        ax = int(np.argmin(tm))
        srt = np.sort(tm)
        if srt[1] - srt[0] < eps * max(1.0, abs(srt[0])):
            amb = True
        t = tm[ax]
        if t > tx:
            break
        v[ax] += step[ax]
        if v[ax] < 0 or v[ax] >= N:
            break
    return acc, first, tfirst, amb


def ray_dirs(view, w, h):
    b = np.asarray(view[3:], np.float32).reshape(3, 3)
    xs, ys = np.meshgrid(np.arange(w, dtype=np.float32), np.arange(h, dtype=np.float32))
    dd = b[0][None, None] + b[1][None, None] * xs[..., None] + b[2][None, None] * ys[..., None]
    n = np.sqrt((dd.astype(np.float32) ** 2).sum(-1, dtype=np.float32)).astype(np.float32)
    dd = (dd / n[..., None]).astype(np.float32)
    small = np.abs(dd) < np.float32(1e-5)
    dd[small] = np.copysign(np.float32(1e-5), dd[small])
    return dd


def check_cast(yx, ntree_orc, dense_orc, size, levels, seed, min_ok=300):
    """Render with `ntree_orc` on the tree built from a dense volume and compare
    with the float64 dense walk."""
    d = small_desc(dense_orc, size=size, levels=levels, seed=seed)
    dense = dense_orc.density_box(d, 0, 0, 0, size, size, size)
    root, nodes, bricks = dense_orc.build_dense(dense, levels)
    pal = palette(seed)
    w, h = 40, 30
    ok = bad = 0
    problems = []
    cams = yx.workloads.orbit_cameras(3, seed=seed) + [((0.37, 0.52, 0.46), (0.3, -1.0, 0.2))]
    for ci, view in enumerate(yx.composite.camera_views(cams, 55.0, w, h)):
        out = ntree_orc.render(root, nodes, bricks, size, levels, view, pal, w, h)
        dirs = ray_dirs(view, w, h)
        for y in range(0, h, 2):
            for x in range(0, w, 2):
                acc, first, tf, amb = brute_ray(dense, pal, view[:3], dirs[y, x])
                if amb:
                    continue
                got = out["rgba"][y, x]
                gch = np.array([(got >> s) & 255 for s in (0, 8, 16, 24)], np.int64)
                ech = (np.clip(acc, 0, 1) * 255.0).astype(np.int64)
                gv = tuple(int(q) for q in out["voxel"][y, x])
                fine = np.all(np.abs(gch - ech) <= 1)
                if first is None:
                    fine &= gv == (-1, -1, -1) and out["t"][y, x] == -1.0
                else:
                    fine &= gv == first and abs(out["t"][y, x] - tf) <= 1e-5 * max(1.0, abs(tf))
                if fine:
                    ok += 1
                else:
                    bad += 1
                    problems.append((ci, x, y, gch.tolist(), ech.tolist(), gv, first))
    return ok, bad, problems


@pytest.mark.parametrize("size,levels", [(64, 1), (128, 2), (32, 0)])
def test_cast_matches_dense_walk(yx, oracle_mod, size, levels):
    o = oracle_mod.NTreeOracle()
    ok, bad, problems = check_cast(yx, o, o, size, levels, seed=size + 3 * levels)
    assert bad == 0, problems[:5]
    assert ok > 1000


@pytest.mark.parametrize("plant", [7, 8])
def test_planted_cast_mistakes_are_caught(yx, oracle_mod, tmp_path, plant):
    o = oracle_mod.NTreeOracle()
    p = oracle_mod.NTreeOracle(oracle_mod.build(plant=plant, out=tmp_path / f"liboracle_plant{plant}.so"))
    assert p.plant() == plant
    ok, bad, _ = check_cast(yx, p, o, 64, 1, seed=67)
    assert bad > 0
