"""The decoding step of SVORenderer::DumpTraceData (demo/SVORenderer.cpp:158-192):
yx_scene_decode_hits turns a frame's hit ids into the hit leaves' colours and
normals. Pinned against the oracle's payload decode (itself pinned by closed
forms and an independent encoder in test_payload_c16n16.py), against the raw
leaf words for RGBA8, and on its error cases. The full dump of a rendered frame
runs on the GPU in test_gpu_payload.py."""
import numpy as np
import pytest

import svo_helpers as H
from test_payload_c16n16 import payload


def _scene(yx):
    s = yx.SVOData.build_synthetic(depth=5, num_spheres=6, radius_min=2.0, radius_max=6.0, seed=4)
    nodes = s.nodes().copy()
    leaf = ((nodes[:, 0:1] >> np.arange(8, dtype=np.uint32)) & 1).astype(bool)
    rng = np.random.default_rng(9)
    nrm = rng.normal(size=(int(leaf.sum()), 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    sub = nodes[:, 2:].copy()
    sub[leaf] = payload(sub[leaf], nrm)
    nodes[:, 2:] = sub
    return yx.SVOData.from_nodes(nodes, s.GetRoot()), nodes, leaf


def _leaf_ids(leaf, k, seed):
    node, slot = np.nonzero(leaf)
    pick = np.random.default_rng(seed).choice(node.size, size=k, replace=False)
    return (node[pick] * 8 + slot[pick]).astype(np.uint32)


def test_decode_c16n16_matches_oracle(yx, orc):
    s, nodes, leaf = _scene(yx)
    ids = _leaf_ids(leaf, 200, 1)
    ids = np.concatenate([ids, np.full(7, H.MISS, np.uint32)])
    color, normal = s.decode_hits(ids, yx.YX_PAYLOAD_C16N16)
    for i, h in enumerate(ids):
        if h == H.MISS:
            assert color[i] == 0 and not normal[i].any()
            continue
        want_c, want_n = orc.decode_c16n16(int(nodes[h >> 3, 2 + (h & 7)]))
        assert color[i] == want_c, i
        assert np.array_equal(normal[i], want_n), i
    # unit normals, opaque colours
    hit = ids != H.MISS
    assert np.allclose(np.linalg.norm(normal[hit], axis=1), 1.0, atol=1e-6)
    assert ((color[hit] >> 24) == 255).all()


def test_decode_rgba8_is_the_leaf_word(yx):
    s, nodes, leaf = _scene(yx)
    ids = _leaf_ids(leaf, 100, 2)
    color, normal = s.decode_hits(ids, yx.YX_PAYLOAD_RGBA8)
    assert np.array_equal(color, nodes[ids >> 3, 2 + (ids & 7)])
    assert not normal.any()


def test_decode_rejects_ids_that_name_no_leaf(yx):
    s, nodes, leaf = _scene(yx)
    with pytest.raises(yx.YoxelError, match="FORMAT"):
        s.decode_hits(np.array([nodes.shape[0] * 8], np.uint32))
    node, slot = np.nonzero(~leaf)
    with pytest.raises(yx.YoxelError, match="FORMAT"):
        s.decode_hits(np.array([node[0] * 8 + slot[0]], np.uint32))
    with pytest.raises(yx.YoxelError, match="INVALID"):
        s.decode_hits(np.array([H.MISS], np.uint32), payload=7)
    c, n = s.decode_hits(np.zeros(0, np.uint32))
    assert c.size == 0 and n.shape == (0, 3)
