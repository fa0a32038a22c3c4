"""The reference's leaf payload read as colour16 + normal16 (VoxData,
reaction/report/main.tex:38,65). The unpacking bodies are missing from the
reference (demo/SVORenderer.cpp:181-185), so the layout is this project's
definition (include/yoxel.h, YX_PAYLOAD_C16N16): parity with the reference is
unpinned; these tests pin the oracle's decode to that definition:
  * colour: RGB565 closed forms (primaries, white, a mid grey);
  * normal: an independent numpy encoder of the published method (Baptista,
    "Higher Accuracy Quantized Normals": signs in three bits, (|x|, |y|) /
    (|x|+|y|+|z|) on a 126-step triangle folded into 6 + 7 bits); axis normals
    decode exactly, 20 000 random normals round-trip within the quantisation
    error with their signs; a planted swap of u and v (ORACLE_PLANT 15) fails;
  * shading: Lambert over the stored normal against its float64 closed form.
The GPU side (tests/test_gpu_payload.py) equals the oracle bit for bit."""
import numpy as np
import pytest

import svo_helpers as H


def encode_normal(n):
    """Quantize unit normals (N, 3) to 16-bit codes."""
    n = np.asarray(n, np.float64)
    a = np.abs(n)
    s = a.sum(axis=1)
    u = np.floor(a[:, 0] / s * 126.0 + 0.5).astype(np.int64)
    v = np.floor(a[:, 1] / s * 126.0 + 0.5).astype(np.int64)
    over = u + v > 126  # rounding both up can leave the triangle
    v = np.where(over, 126 - u, v)
    fold = u >= 64
    u2 = np.where(fold, 127 - u, u)
    v2 = np.where(fold, 127 - v, v)
    signs = (n[:, 0] < 0) * 0x8000 + (n[:, 1] < 0) * 0x4000 + (n[:, 2] < 0) * 0x2000
    return (signs | (u2 << 7) | v2).astype(np.uint32)


def encode_color(rgba):
    rgba = np.asarray(rgba, np.uint32)
    r, g, b = rgba & 255, (rgba >> 8) & 255, (rgba >> 16) & 255
    return ((r >> 3) << 11 | (g >> 2) << 5 | (b >> 3)).astype(np.uint32)


def payload(rgba, normals):
    return (encode_color(rgba) | (encode_normal(normals) << 16)).astype(np.uint32)


def test_color565_closed_forms(orc):
    for c16, want in [(0xF800, (255, 0, 0)), (0x07E0, (0, 255, 0)), (0x001F, (0, 0, 255)),
                      (0xFFFF, (255, 255, 255)), (0x0000, (0, 0, 0)), (0x8410, (132, 130, 132))]:
        rgba, _ = orc.decode_c16n16(c16 | (0x0000 << 16))
        assert [(rgba >> (8 * i)) & 255 for i in range(4)] == list(want) + [255], hex(c16)


def _check_normals(o):
    for axis in range(3):
        for sgn in (1.0, -1.0):
            n = np.zeros((1, 3))
            n[0, axis] = sgn
            _, got = o.decode_c16n16(int(encode_normal(n)[0]) << 16)
            assert np.array_equal(got, n[0].astype(np.float32)), (axis, sgn, got)
    rng = np.random.default_rng(5)
    n = rng.normal(size=(20000, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    codes = encode_normal(n)
    got = np.array([o.decode_c16n16(int(c) << 16)[1] for c in codes], np.float64)
    np.testing.assert_allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-6)
    # signs exactly (components below one triangle step, about 1/126 of the
    # L1 norm, quantise to 0), direction within the quantisation error of a
    # 126-step triangle (0.93 degrees at most over these normals)
    nz = np.abs(n) > 0.02
    assert np.all(np.sign(got[nz]) == np.sign(n[nz]))
    ang = np.degrees(np.arccos(np.clip(np.sum(got * n, axis=1), -1.0, 1.0)))
    assert ang.max() < 1.2, ang.max()


def test_normal_round_trip(orc):
    _check_normals(orc)


def test_stored_normal_lambert(orc):
    # one leaf whose stored normal is tilted away from the face it is entered
    # through: the shade follows the stored normal, not the face
    rng = np.random.default_rng(9)
    for trial in range(40):
        n = rng.normal(size=(1, 3))
        n /= np.linalg.norm(n)
        col = int(rng.integers(0, 1 << 24)) | 0xFF000000
        leaf = int(payload([col], n)[0])
        nodes, root = H.tree_from_leaves([(1, 1, 0, 1, leaf)])
        pos = np.array([[-1.0, 0.2, 0.7]], np.float32)
        d = np.array([[1.0, 0.05, 0.03]], np.float32)
        d /= np.linalg.norm(d)
        rgba565, nq = orc.decode_c16n16(leaf)
        c = np.array([(rgba565 >> (8 * i)) & 255 for i in range(3)], np.float64)
        for mode, L in ((0, -d[0].astype(np.float64)), (1, np.array([0.3, -0.5, 0.81]) / np.linalg.norm([0.3, -0.5, 0.81]))):
            out = orc.shade_rays(nodes, root, pos, d, light_mode=mode, light=L.astype(np.float32), payload=1)
            assert out["hit_id"][0] == root * 8 + 5
            k = 0.1 + 0.9 * max(0.0, float(nq.astype(np.float64) @ L))
            got = np.array([(int(out["rgba"][0]) >> (8 * i)) & 255 for i in range(3)], np.float64)
            assert np.all(np.abs(got - c * k) <= 0.5 + 1e-3), (trial, mode, got, c * k)


def test_planted_normal_swap_is_caught(oracle_mod, tmp_path):
    bad = oracle_mod.Oracle(oracle_mod.build(plant=15, out=tmp_path / "liboracle_plant15.so"))
    with pytest.raises(AssertionError):
        _check_normals(bad)
