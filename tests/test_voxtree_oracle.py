"""Pins of the voxtree restatement (oracle VoxtreeOracle; reaction/voxtree/density.cu
CalcDensity :32-60, MarkBricks :131-176, PackBricks :210-221, voxtree.py:80-95).
The oracle works on unpacked volumes; these checks come at it another way:

* the reference's word-level arithmetic (pair sums of adjacent bits, even / odd
  two-bit fields summed over the four bit columns, nibbles interleaved) on
  random words gives the oracle's density words;
* per-brick loops over the 5^3 window of each brick give its states;
* closed forms: an all-solid and an all-empty volume, and single voxels, whose
  mixed bricks can be counted by hand;
* the packed list is the mixed bricks sorted by (y, x, z).
Planted mistakes in the oracle (a 4^3 window, a reversed nibble order) fail them.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def vo(oracle_mod):
    return oracle_mod.VoxtreeOracle()


def structured(B, seed):
    rng = np.random.default_rng(seed)
    c = np.arange(B) + 0.5
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    vol = Z < B * (0.3 + 0.1 * np.sin(X / B * 6.0) * np.cos(Y / B * 5.0))
    for _ in range(12):
        p = rng.uniform(0, B, 3)
        r = rng.uniform(B / 32, B / 6)
        vol |= (X - p[0]) ** 2 + (Y - p[1]) ** 2 + (Z - p[2]) ** 2 <= r * r
    vol ^= rng.random(vol.shape) < 0.002  # scattered voxels and holes
    return vol


def pairs(v):
    return ((v >> 1) & 0x55555555) + (v & 0x55555555)


def interleave(even, odd):
    r = 0
    for m in range(4):
        r |= ((even >> (4 * m)) & 0xF) << (8 * m)
        r |= ((odd >> (4 * m)) & 0xF) << (8 * m + 4)
    return r


def density_words_scalar(words, B, s, x, y):
    """The 8 density words of slice s of half-resolution column (x, y), from
    the four bit columns' words (reference arithmetic)."""
    w = words.reshape(B // 128, B, B, 4)
    out = []
    for layer in range(4):
        v = [pairs(int(w[s, 2 * y + dy, 2 * x + dx, layer])) for dy in (0, 1) for dx in (0, 1)]
        even = sum(t & 0x33333333 for t in v)
        odd = sum((t >> 2) & 0x33333333 for t in v)
        out += [interleave(even, odd), interleave(even >> 16, odd >> 16)]
    return out


def check_density_words(vo, B, seed):
    rng = np.random.default_rng(seed)
    words = rng.integers(0, 2 ** 32, size=(B // 128) * B * B * 4, dtype=np.uint64).astype(np.uint32)
    dens = vo.density(words, B).reshape(B // 2, B // 2, B // 16)
    for _ in range(300):
        s, x, y = int(rng.integers(B // 128)), int(rng.integers(B // 2)), int(rng.integers(B // 2))
        assert dens[y, x, 8 * s:8 * s + 8].tolist() == density_words_scalar(words, B, s, x, y)


def test_density_equals_word_arithmetic(vo):
    for B, seed in ((128, 1), (256, 2)):
        check_density_words(vo, B, seed)


def brute_states(counts):
    V = counts.shape[0]
    G = V // 4
    st = np.zeros((G, G, G), np.uint32)
    for bx in range(G):
        for by in range(G):
            for bz in range(G):
                w = counts[4 * bx:4 * bx + 5, 4 * by:4 * by + 5, 4 * bz:4 * bz + 5]
                full = w.shape == (5, 5, 5)  # a window past the volume holds zeros
                if full and np.all(w == 8):
                    st[bx, by, bz] = 0x80000000
                elif np.all(w == 0):
                    st[bx, by, bz] = 0x40000000
    return st


def check_mark(vo, B, seed):
    vol = structured(B, seed)
    words = vo.pack_bits(vol)
    V, G = B // 2, B // 4 // 2
    counts = vo.counts(words, B)
    state, cnt = vo.mark(vo.density(words, B), V)
    exp = brute_states(counts)
    st = state.reshape(G, G, G)  # [by][bx][bz]
    for by in range(G):
        for bx in range(G):
            col = st[by, bx]
            kinds = col & np.uint32(0xC0000000)
            assert np.array_equal(kinds, exp[bx, by])
            mixed = kinds == 0
            assert np.array_equal(col & np.uint32(0x3FFFFFFF), np.cumsum(mixed) - mixed)
            assert cnt[by * G + bx] == mixed.sum()
    return state, cnt


def test_mark_equals_per_brick_loops(vo):
    state, cnt = check_mark(vo, 128, 3)
    assert 0 < cnt.sum() < cnt.size * 16  # some mixed, some uniform bricks


def test_closed_forms(vo, oracle_mod):
    B = 128
    V, G = B // 2, B // 8
    full = vo.classify(vo.pack_bits(np.ones((B, B, B), bool)), B)
    assert np.all(vo.unpack_density(full["density"], V) == 8)
    kinds = full["brick_state"].reshape(G, G, G) & np.uint32(0xC0000000)
    # bricks whose window reaches past an upper face see zeros: mixed
    assert full["total"] == G ** 3 - (G - 1) ** 3
    assert np.all(kinds[:G - 1, :G - 1, :G - 1] == 0x80000000)
    empty = vo.classify(vo.pack_bits(np.zeros((B, B, B), bool)), B)
    assert empty["total"] == 0 and np.all(empty["brick_state"] == 0x40000000)
    assert np.all(empty["density"] == 0)
    for p in [(0, 0, 0), (8, 8, 8), (9, 17, 126), (127, 127, 127), (16, 3, 64)]:
        vol = np.zeros((B, B, B), bool)
        vol[p] = True
        r = vo.classify(vo.pack_bits(vol), B)
        c = [q // 2 for q in p]
        cnt = vo.unpack_density(r["density"], V)
        assert cnt[tuple(c)] == 1 and cnt.sum() == 1
        # count coordinate c lies in the windows of bricks c // 4 and, on a
        # multiple of 4 above 0, c // 4 - 1
        per_axis = [2 if (q % 4 == 0 and q > 0) else 1 for q in c]
        assert r["total"] == int(np.prod(per_axis))
        exp = sorted({(bx, by, bz) for bx in {c[0] // 4, (c[0] - 1) // 4 if c[0] % 4 == 0 and c[0] else c[0] // 4}
                      for by in {c[1] // 4, (c[1] - 1) // 4 if c[1] % 4 == 0 and c[1] else c[1] // 4}
                      for bz in {c[2] // 4, (c[2] - 1) // 4 if c[2] % 4 == 0 and c[2] else c[2] // 4}},
                     key=lambda t: (t[1], t[0], t[2]))
        got = [(int(v) & 0xFF, (int(v) >> 8) & 0xFF, int(v) >> 16) for v in r["packed"]]
        assert got == exp


def test_packed_order(vo):
    B = 256
    r = vo.classify(vo.pack_bits(structured(B, 5)), B)
    G = B // 8
    st = r["brick_state"].reshape(G, G, G)
    by, bx, bz = np.nonzero((st & np.uint32(0xC0000000)) == 0)
    order = np.lexsort((bz, bx, by))
    exp = bx[order] | (by[order] << 8) | (bz[order] << 16)
    assert np.array_equal(r["packed"], exp.astype(np.uint32))
    assert r["total"] == len(exp) == int(r["column_count"].sum())
    assert r["column_start"][-1] + r["column_count"][-1] == r["total"]


def test_planted_mistakes_fail_these_pins(oracle_mod):
    bad = oracle_mod.VoxtreeOracle()
    bad.window = 4
    with pytest.raises(AssertionError):
        check_mark(bad, 128, 3)
    with pytest.raises(AssertionError):
        test_closed_forms(bad, oracle_mod)
    bad = oracle_mod.VoxtreeOracle()
    bad.nibble_order = list(range(7, -1, -1))
    with pytest.raises(AssertionError):
        check_density_words(bad, 128, 1)
