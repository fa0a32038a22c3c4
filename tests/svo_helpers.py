"""Independent helpers for the tests: a VoxNode tree writer from leaf boxes (any
node order), a tree walker, a float64 brute-force first-hit search, and a
numpy restatement of the synthetic voxelizer's definition."""
import numpy as np

EMPTY = 0x80000000
MISS = 0xFFFFFFFF
MASK64 = (1 << 64) - 1


# ---- trees from leaf boxes -------------------------------------------------

def tree_from_leaves(leaves, order="bfs", seed=0):
    """leaves: iterable of (level, x, y, z, payload) with level >= 1 and coords
    in [0, 2^level). Returns (nodes[N,10] uint32, root). order: 'bfs' or
    'shuffle' (random node ids, root anywhere)."""
    nodes = {}  # (level, x, y, z) -> {"leaf": {slot: payload}, "child": set(slot)}

    def node(key):
        if key not in nodes:
            nodes[key] = {"leaf": {}, "child": set()}
        return nodes[key]

    node((0, 0, 0, 0))
    for (lv, x, y, z, payload) in leaves:
        assert lv >= 1
        slot = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2)
        pkey = (lv - 1, x >> 1, y >> 1, z >> 1)
        node(pkey)["leaf"][slot] = int(payload) & 0xFFFFFFFF
        # ancestors
        cl, cx, cy, cz = pkey
        while cl > 0:
            s = (cx & 1) | ((cy & 1) << 1) | ((cz & 1) << 2)
            par = (cl - 1, cx >> 1, cy >> 1, cz >> 1)
            node(par)["child"].add(s)
            cl, cx, cy, cz = par
    for key, n in nodes.items():
        assert not (set(n["leaf"]) & n["child"]), f"leaf overlaps a subtree at {key}"
    keys = sorted(nodes, key=lambda k: (k[0], k[1], k[2], k[3]))
    if order == "shuffle":
        rng = np.random.default_rng(seed)
        perm = rng.permutation(len(keys))
        ids = {k: int(perm[i]) for i, k in enumerate(keys)}
    else:
        ids = {k: i for i, k in enumerate(keys)}
    out = np.zeros((len(keys), 10), np.uint32)
    for k, n in nodes.items():
        i = ids[k]
        flags = 0
        lv, x, y, z = k
        for s in range(8):
            if s in n["leaf"]:
                flags |= 1 << s
                out[i, 2 + s] = n["leaf"][s]
            elif s in n["child"]:
                ck = (lv + 1, 2 * x + (s & 1), 2 * y + ((s >> 1) & 1), 2 * z + ((s >> 2) & 1))
                out[i, 2 + s] = ids[ck]
            else:
                flags |= 1 << (8 + s)
                out[i, 2 + s] = EMPTY
        out[i, 0] = flags
    return out, ids[(0, 0, 0, 0)]


def walk_leaves(nodes, root):
    """All leaves of a VoxNode tree: dict (level, x, y, z) -> (payload, parent id, slot)."""
    res = {}
    stack = [(int(root), 0, 0, 0, 0)]
    while stack:
        nid, lv, x, y, z = stack.pop()
        n = nodes[nid]
        for s in range(8):
            cx, cy, cz = 2 * x + (s & 1), 2 * y + ((s >> 1) & 1), 2 * z + ((s >> 2) & 1)
            if (int(n[0]) >> s) & 1:
                res[(lv + 1, cx, cy, cz)] = (int(n[2 + s]), nid, s)
            elif not (int(n[2 + s]) & 0x80000000):
                stack.append((int(n[2 + s]), lv + 1, cx, cy, cz))
    return res


def dense_leaves(occ, rng=None):
    """Leaves at the finest level of a dense boolean grid occ[x, y, z]."""
    n = occ.shape[0]
    lv = int(np.log2(n))
    xs, ys, zs = np.nonzero(occ)
    pay = (rng.integers(0, 2**32, size=xs.size, dtype=np.uint64) if rng is not None
           else np.arange(xs.size, dtype=np.uint64) * 2654435761)
    return [(lv, int(a), int(b), int(c), int(p) & 0xFFFFFFFF) for a, b, c, p in zip(xs, ys, zs, pay)]


# ---- brute force ------------------------------------------------------------

def adjust(d):
    d = np.array(d, np.float32)
    small = np.abs(d) < np.float32(1e-6)
    d[small] = np.where(d[small] < 0, np.float32(-1e-6), np.float32(1e-6))
    return d


def brute_first_hit(leaf_keys, pos, dirs, eps=1e-5):
    """First leaf box along each ray by exhaustive float64 slab tests.
    Returns (idx, tin, ambiguous) with idx -1 for a miss. A ray is ambiguous
    when the winner is not separated by eps from another candidate or from a
    grazing box."""
    keys = np.asarray(leaf_keys, np.float64)  # (L, 4): level, x, y, z
    scale = 2.0 ** (-keys[:, 0])
    lo = keys[:, 1:4] * scale[:, None]
    hi = (keys[:, 1:4] + 1.0) * scale[:, None]
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    dirs = np.asarray(dirs, np.float64).reshape(-1, 3)
    if pos.shape[0] == 1:
        pos = np.repeat(pos, dirs.shape[0], axis=0)
    n = dirs.shape[0]
    idx = np.full(n, -1, np.int64)
    tin_best = np.full(n, np.inf)
    amb = np.zeros(n, bool)
    for s in range(0, n, 256):
        p = pos[s:s + 256, None, :]
        d = dirs[s:s + 256, None, :]
        ta = (lo[None] - p) / d
        tb = (hi[None] - p) / d
        tin = np.minimum(ta, tb).max(axis=2)
        tout = np.maximum(ta, tb).min(axis=2)
        solid = tout - tin > eps
        graze = (tout - tin > -eps) & ~solid
        tin_s = np.where(solid, tin, np.inf)
        best = tin_s.argmin(axis=1)
        bt = tin_s[np.arange(best.size), best]
        tin_s2 = tin_s.copy()
        tin_s2[np.arange(best.size), best] = np.inf
        second = tin_s2.min(axis=1)
        tg = np.where(graze, tin, np.inf).min(axis=1)
        hit = np.isfinite(bt)
        idx[s:s + 256] = np.where(hit, best, -1)
        tin_best[s:s + 256] = bt
        with np.errstate(invalid="ignore"):
            amb[s:s + 256] = (second - bt < eps) | (tg < bt + eps)
    return idx, tin_best, amb


# ---- numpy restatement of the synthetic voxelizer ----------------------------

def _mix64(z):
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9 & MASK64
    z = (z ^ (z >> 27)) * 0x94D049BB133111EB & MASK64
    return z ^ (z >> 31)


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        return _mix64(self.s)

    def uniform(self):
        return float(self.next() >> 11) * 2.0 ** -53


def perlin_perm(seed):
    rng = SplitMix64(seed)
    p = list(range(256))
    for i in range(255, 0, -1):
        j = rng.next() % (i + 1)
        p[i], p[j] = p[j], p[i]
    return np.array(p + p, np.int64)


def _fade(t):
    return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)


def _lerp(t, a, b):
    return a + t * (b - a)


def _grad2(h, x, y):
    h = h & 7
    out = np.empty_like(x)
    table = [x + y, -x + y, x - y, -x - y, x, -x, y, -y]
    for k in range(8):
        m = h == k
        out[m] = table[k][m]
    return out


def noise2(perm, x, y):
    fx, fy = np.floor(x), np.floor(y)
    xi = fx.astype(np.int64) & 255
    yi = fy.astype(np.int64) & 255
    xf, yf = x - fx, y - fy
    u, v = _fade(xf), _fade(yf)
    aa = perm[perm[xi] + yi]
    ab = perm[perm[xi] + yi + 1]
    ba = perm[perm[xi + 1] + yi]
    bb = perm[perm[xi + 1] + yi + 1]
    x1 = _lerp(u, _grad2(aa, xf, yf), _grad2(ba, xf - 1.0, yf))
    x2 = _lerp(u, _grad2(ab, xf, yf - 1.0), _grad2(bb, xf - 1.0, yf - 1.0))
    return _lerp(v, x1, x2)


def fbm2(perm, x, y, octaves):
    s = np.zeros_like(x)
    amp, freq, norm = 1.0, 1.0, 0.0
    for _ in range(octaves):
        s = s + amp * noise2(perm, x * freq, y * freq)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return s / norm


def synthetic_inside(depth, num_spheres, radius_min, radius_max, heightfield, hf_octaves, hf_amplitude,
                     hf_base, seed):
    """Dense inside grid [x, y, z] of the synthetic scene definition."""
    n = 1 << depth
    rng = SplitMix64(seed)
    c = np.arange(n, dtype=np.float64) + 0.5
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    inside = np.zeros((n, n, n), bool)
    rmin = float(np.float32(radius_min))
    rmax = float(np.float32(radius_max))
    for _ in range(num_spheres):
        cx, cy, cz = rng.uniform(), rng.uniform(), rng.uniform()
        r = rmin + (rmax - rmin) * rng.uniform()
        dx, dy, dz = X - cx * n, Y - cy * n, Z - cz * n
        inside |= dx * dx + dy * dy + dz * dz <= r * r
    if heightfield:
        perm = perlin_perm(seed ^ 0xD1B54A32D192ED03)
        u = c / float(n)
        U, V = np.meshgrid(u, u, indexing="ij")
        f = fbm2(perm, U * 4.0, V * 4.0, hf_octaves)
        h = float(n) * (float(np.float32(hf_base)) + float(np.float32(hf_amplitude)) * f)
        inside |= Z < h[:, :, None]
    return inside


def surface(inside):
    pad = np.pad(inside, 1, constant_values=False)
    interior = inside.copy()
    for ax in range(3):
        for sh in (-1, 1):
            interior &= np.roll(pad, sh, axis=ax)[1:-1, 1:-1, 1:-1]
    return inside & ~interior


def morton3(x, y, z):
    code = np.zeros(np.shape(x), np.uint64)
    for b in range(21):
        code |= ((np.asarray(x, np.uint64) >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
        code |= ((np.asarray(y, np.uint64) >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + 1)
        code |= ((np.asarray(z, np.uint64) >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + 2)
    return code


def leaf_color(code, seed):
    return _mix64(int(code) ^ ((seed * 0x9E3779B97F4A7C15) & MASK64)) & 0xFFFFFFFF


def edit_balls(depth, edits):
    """Sphere edits (mode, level, pos, radius, color) in finest-level voxels:
    pos and radius scale by 2^(depth - level) (float32 inputs, float64 maths)."""
    out = []
    for mode, level, pos, radius, color in edits:
        sc = float(2 ** (depth - level))
        c = [float(np.float32(p)) * sc for p in pos]
        out.append((mode, c, float(np.float32(radius)) * sc, color))
    return out


def edited_inside(desc, edits):
    """Dense inside grid after BuildRange edits in order: GROW (mode 0) is a
    union with the ball, CLEAR (mode 1) a difference. A voxel is in a ball when
    its centre is within the radius."""
    inside = synthetic_inside(**desc)
    n = inside.shape[0]
    c = np.arange(n, dtype=np.float64) + 0.5
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    for mode, (cx, cy, cz), r, _ in edit_balls(desc["depth"], edits):
        ball = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= r * r
        inside = inside | ball if mode == 0 else inside & ~ball
    return inside


def edited_color(x, y, z, desc, edits):
    """Leaf colour after edits: the last ball holding the voxel decides (its
    colour for a GROW, the recipe's hash for a CLEAR); else the hash."""
    p = (x + 0.5, y + 0.5, z + 0.5)
    for mode, c, r, color in reversed(edit_balls(desc["depth"], edits)):
        if sum((p[a] - c[a]) ** 2 for a in range(3)) <= r * r:
            return color if mode == 0 else leaf_color(int(morton3(x, y, z)), desc["seed"])
    return leaf_color(int(morton3(x, y, z)), desc["seed"])


def leaf_colors_np(codes, seed):
    """leaf_color over an array of Morton codes (uint64 arithmetic wraps)."""
    z = np.asarray(codes, np.uint64) ^ np.uint64((seed * 0x9E3779B97F4A7C15) & MASK64)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return ((z ^ (z >> np.uint64(31))) & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def edited_colors_np(xs, ys, zs, desc, edits):
    """edited_color over arrays of voxels: the last ball holding a voxel decides."""
    out = leaf_colors_np(morton3(xs, ys, zs), desc["seed"])
    p = [np.asarray(v, np.float64) + 0.5 for v in (xs, ys, zs)]
    for mode, c, r, color in edit_balls(desc["depth"], edits):
        inb = (p[0] - c[0]) ** 2 + (p[1] - c[1]) ** 2 + (p[2] - c[2]) ** 2 <= r * r
        if mode == 0:
            out[inb] = np.uint32(color)
        else:
            out[inb] = leaf_colors_np(morton3(xs[inb], ys[inb], zs[inb]), desc["seed"])
    return out


def place_source(depth, level, pos, solid_zyx):
    """Dense [x, y, z] mask of a volume source's solid voxels: source voxel
    (a, b, c) covers voxels pos + (a, b, c) of `level`, 2^(depth - level)
    finest voxels per axis; parts outside the domain are dropped."""
    n = 1 << depth
    s = 1 << (depth - level)
    m = np.transpose(solid_zyx, (2, 1, 0))
    m = m.repeat(s, 0).repeat(s, 1).repeat(s, 2)
    out = np.zeros((n, n, n), bool)
    lo = [p * s for p in pos]
    src = [slice(max(0, -lo[a]), max(0, min(m.shape[a], n - lo[a]))) for a in range(3)]
    dst = [slice(max(0, lo[a]), max(0, min(n, lo[a] + m.shape[a]))) for a in range(3)]
    if all(d.stop > d.start for d in dst):
        out[tuple(dst)] = m[tuple(src)]
    return out


def place_colors(depth, level, pos, col_zyx):
    """Per-voxel colours of a volume source, placed as place_source does."""
    n = 1 << depth
    s = 1 << (depth - level)
    m = np.transpose(col_zyx, (2, 1, 0))
    m = m.repeat(s, 0).repeat(s, 1).repeat(s, 2)
    out = np.zeros((n, n, n), np.uint32)
    lo = [p * s for p in pos]
    src = [slice(max(0, -lo[a]), max(0, min(m.shape[a], n - lo[a]))) for a in range(3)]
    dst = [slice(max(0, lo[a]), max(0, min(n, lo[a] + m.shape[a]))) for a in range(3)]
    if all(d.stop > d.start for d in dst):
        out[tuple(dst)] = m[tuple(src)]
    return out


def source_mask(e):
    """(solid [z][y][x], colours [z][y][x]) of a volume edit dict."""
    if e["kind"] == "iso":
        above = e["data"].astype(np.int64) >= e["iso"]
        solid = above if e.get("inside", True) else ~above
        return solid, np.full(e["data"].shape, e.get("color", 0xFFC0B6C0), np.uint32)
    d = e["data"].astype(np.uint32)
    return (d >> 24) != 0, d


def mixed_inside_and_colors(desc, edits):
    """Dense inside grid and leaf colours after mixed edits: spheres as
    (mode, level, pos, radius, color) tuples, volume sources as dicts with kind
    'iso' / 'raw', mode, level, pos, data [z][y][x] and (iso) iso, inside, color."""
    depth = desc["depth"]
    inside = synthetic_inside(**desc)
    n = inside.shape[0]
    c = np.arange(n, dtype=np.float64) + 0.5
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    last = np.full((n, n, n), -1, np.int64)    # last edit holding the voxel
    colmap = np.zeros((n, n, n), np.uint32)
    grow = np.zeros((n, n, n), bool)
    for e in edits:
        if isinstance(e, dict):
            solid, cols = source_mask(e)
            m = place_source(depth, e["level"], e["pos"], solid)
            col = place_colors(depth, e["level"], e["pos"], cols)
            mode = e["mode"]
        else:
            mode, level, pos, radius, color = e
            (_, (cx, cy, cz), r, _), = edit_balls(depth, [e])
            m = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= r * r
            col = np.full((n, n, n), color, np.uint32)
        inside = inside | m if mode == 0 else inside & ~m
        colmap[m] = col[m]
        grow[m] = mode == 0
        last[m] = 1
    return inside, colmap, grow, last >= 0


def mixed_leaf_colors(desc, edits, xs, ys, zs):
    inside, colmap, grow, touched = mixed_inside_and_colors(desc, edits)
    out = leaf_colors_np(morton3(xs, ys, zs), desc["seed"])
    sel = touched[xs, ys, zs] & grow[xs, ys, zs]
    out[sel] = colmap[xs[sel], ys[sel], zs[sel]]
    return out


FULL = 0x80000001


def encode_tree(state, color, coarse=None):
    """VoxNode tree of a dense voxel grid [x, y, z] of side 2^depth: state 0
    empty, 1 solid inside (FullNode; whole cubes collapse), 2 leaf with
    color. `coarse` {(level, ox, oy, oz): colour} puts a leaf slot at a node
    cube (origin in voxels), covering it whole. Returns (nodes, root)."""
    n = state.shape[0]
    depth = n.bit_length() - 1
    coarse = coarse or {}
    rows = []

    def slot(level, o):
        size = n >> level
        if (level, *o) in coarse:
            return ("leaf", coarse[(level, *o)])
        sub = state[o[0]:o[0] + size, o[1]:o[1] + size, o[2]:o[2] + size]
        if level == depth:
            st = int(sub[0, 0, 0])
            return ("leaf", int(color[o])) if st == 2 else (("full", FULL) if st == 1 else ("empty", EMPTY))
        if not sub.any() and not any(k[0] > level and all(o[a] <= k[1 + a] < o[a] + size for a in range(3))
                                     for k in coarse):
            return ("empty", EMPTY)
        if (sub == 1).all():
            return ("full", FULL)
        half = size // 2
        kids = [slot(level + 1, (o[0] + (c & 1) * half, o[1] + ((c >> 1) & 1) * half, o[2] + ((c >> 2) & 1) * half))
                for c in range(8)]
        row = np.full(10, EMPTY, np.uint32)
        f = 0
        for c, (kind, v) in enumerate(kids):
            row[2 + c] = v
            if kind == "leaf":
                f |= 1 << c
            elif kind in ("empty", "full"):
                f |= 1 << (8 + c)
        row[0] = f
        row[1] = 0
        rows.append(row)
        return ("node", len(rows) - 1)

    kind, v = slot(0, (0, 0, 0))
    return (np.array(rows, np.uint32).reshape(-1, 10) if rows else np.zeros((0, 10), np.uint32)), \
        (v if kind == "node" else (FULL if kind == "full" else EMPTY))


def decode_tree(nodes, root, depth):
    """(state, color) [x, y, z] of a tree: leaf slots fill their cubes with
    state 2 and their colour, FullNode cubes state 1."""
    n = 1 << depth
    state = np.zeros((n, n, n), np.uint8)
    color = np.zeros((n, n, n), np.uint32)

    def walk(idx, level, o):
        size = n >> level
        row = nodes[idx]
        f = int(row[0])
        half = size // 2
        for c in range(8):
            co = (o[0] + (c & 1) * half, o[1] + ((c >> 1) & 1) * half, o[2] + ((c >> 2) & 1) * half)
            ch = int(row[2 + c])
            sl = tuple(slice(co[a], co[a] + half) for a in range(3))
            if (f >> c) & 1:
                state[sl] = 2
                color[sl] = ch
            elif ch == FULL:
                state[sl] = 1
            elif not ch & 0x80000000:
                if level + 1 < depth:
                    walk(ch, level + 1, co)

    if root == FULL:
        state[:] = 1
    elif not root & 0x80000000:
        walk(int(root), 0, (0, 0, 0))
    return state, color


def canonical_states(solid):
    """Surface voxels (solid, a 6-neighbour empty or outside) 2, other solid 1."""
    surf = surface(solid)
    return np.where(surf, 2, np.where(solid, 1, 0)).astype(np.uint8)
