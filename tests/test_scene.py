"""Scene side: the synthetic voxelizer against a dense numpy evaluation of its
definition, the canonical breadth-first node order, .vox round trips
(SVOData::Load layout, cell/svodata.h:31-50) and load errors."""
import struct

import numpy as np
import pytest

import svo_helpers as H


CASES = [
    dict(depth=5, num_spheres=6, radius_min=2.0, radius_max=7.0, heightfield=False, hf_octaves=5,
         hf_amplitude=0.2, hf_base=0.3, seed=1),
    dict(depth=6, num_spheres=10, radius_min=3.0, radius_max=12.0, heightfield=True, hf_octaves=5,
         hf_amplitude=0.2, hf_base=0.3, seed=2),
    dict(depth=6, num_spheres=0, radius_min=0.0, radius_max=0.0, heightfield=True, hf_octaves=3,
         hf_amplitude=0.35, hf_base=0.5, seed=3),
]


@pytest.mark.parametrize("case", CASES)
def test_voxelizer_matches_dense_definition(yx, case):
    s = yx.SVOData.build_synthetic(**case)
    nodes, root = s.nodes(), s.GetRoot()
    got = H.walk_leaves(nodes, root)
    inside = H.synthetic_inside(**case)
    surf = H.surface(inside)
    xs, ys, zs = np.nonzero(surf)
    depth = case["depth"]
    exp = {(depth, int(x), int(y), int(z)) for x, y, z in zip(xs, ys, zs)}
    assert set(got) == exp
    codes = H.morton3(xs, ys, zs)
    for x, y, z, c in zip(xs[:2000], ys[:2000], zs[:2000], codes[:2000]):
        assert got[(depth, int(x), int(y), int(z))][0] == H.leaf_color(int(c), case["seed"])
    info = s.info()
    assert info["num_leaves"] == len(exp) and info["depth"] == depth


def test_breadth_first_canonical_order(yx):
    s = yx.SVOData.build_synthetic(**CASES[1])
    nodes, root = s.nodes(), s.GetRoot()
    assert root == 0
    # children ids of consecutive nodes are consecutive and increasing
    nxt = 1
    for i in range(nodes.shape[0]):
        leaf = nodes[i, 0] & 0xFF
        for c in range(8):
            if not (leaf >> c) & 1 and not (nodes[i, 2 + c] & 0x80000000):
                assert nodes[i, 2 + c] == nxt
                nxt += 1
            if not (leaf >> c) & 1 and (nodes[i, 2 + c] & 0x80000000):
                assert (nodes[i, 0] >> (8 + c)) & 1  # nullFlags
    assert nxt == nodes.shape[0]


def test_empty_scene(yx):
    s = yx.SVOData.build_synthetic(depth=4, num_spheres=0)
    assert s.GetRoot() == yx.YX_EMPTY_NODE and s.info()["num_nodes"] == 0


def test_vox_round_trip(yx, tmp_path):
    s = yx.SVOData.build_synthetic(**CASES[0])
    fn = tmp_path / "scene.vox"
    s.Save(fn)
    raw = fn.read_bytes()
    root, _, _, size = struct.unpack("<4I", raw[:16])
    assert root == s.GetRoot() and size == s.info()["num_nodes"]
    assert len(raw) == 16 + 40 * size
    t = yx.SVOData().Load(fn)
    assert np.array_equal(t.nodes(), s.nodes()) and t.GetRoot() == s.GetRoot()


def test_vox_shuffled_ids_load(yx, tmp_path):
    rng = np.random.default_rng(4)
    occ = rng.random((8, 8, 8)) < 0.3
    nodes, root = H.tree_from_leaves(H.dense_leaves(occ, rng), order="shuffle", seed=5)
    fn = tmp_path / "shuffled.vox"
    fn.write_bytes(struct.pack("<4I", root, 0, 0, nodes.shape[0]) + nodes.astype("<u4").tobytes())
    t = yx.SVOData().Load(fn)
    assert np.array_equal(t.nodes(), nodes) and t.GetRoot() == root
    assert t.info()["depth"] == 3 and t.info()["num_leaves"] == int(occ.sum())


def test_load_errors(yx, tmp_path):
    with pytest.raises(yx.YoxelError) as e:
        yx.SVOData().Load(tmp_path / "missing.vox")
    assert e.value.code == -3  # YX_ERR_IO
    short = tmp_path / "short.vox"
    short.write_bytes(struct.pack("<4I", 0, 0, 0, 5) + b"\0" * 40)
    with pytest.raises(yx.YoxelError) as e:
        yx.SVOData().Load(short)
    assert e.value.code == -3
    hdr = tmp_path / "hdr.vox"
    hdr.write_bytes(b"\1\2")
    with pytest.raises(yx.YoxelError) as e:
        yx.SVOData().Load(hdr)
    assert e.value.code == -3
    # a bare header that names 2^32 - 1 nodes (171 GB) is rejected from the
    # file length, before any allocation
    huge = tmp_path / "huge.vox"
    huge.write_bytes(struct.pack("<4I", 0, 0, 0, 0xFFFFFFFF))
    with pytest.raises(yx.YoxelError) as e:
        yx.SVOData().Load(huge)
    assert e.value.code == -3 and "short file" in str(e.value)
    # a child id beyond the node array
    bad = np.zeros((1, 10), np.uint32)
    bad[0, 2:] = 0x80000000
    bad[0, 3] = 7
    with pytest.raises(yx.YoxelError) as e:
        yx.SVOData.from_nodes(bad, 0)
    assert e.value.code == -4  # YX_ERR_FORMAT
    # a cycle
    cyc = np.zeros((2, 10), np.uint32)
    cyc[:, 2:] = 0x80000000
    cyc[0, 2] = 1
    cyc[1, 2] = 0
    with pytest.raises(yx.YoxelError) as e:
        yx.SVOData.from_nodes(cyc, 0)
    assert e.value.code == -4
    # root out of range
    with pytest.raises(yx.YoxelError) as e:
        yx.SVOData.from_nodes(cyc, 9)
    assert e.value.code == -4


def test_bad_synthetic_desc(yx):
    with pytest.raises(yx.YoxelError) as e:
        yx.SVOData.build_synthetic(depth=0)
    assert e.value.code == -7
    with pytest.raises(yx.YoxelError) as e:
        yx.SVOData.build_synthetic(depth=4, num_spheres=2, radius_min=5.0, radius_max=1.0)
    assert e.value.code == -1
