"""Host side of scene edits and page versions (the node pool of DynamicSVO /
HomoStorage and DynamicSVO::CountChangedPages, reaction/report/main.tex:67-71,
ore/src/main.cpp:127-128): writes, appends, leaf counts, lazy validation, and
the pages a renderer that has never synced would upload. No GPU work."""
import numpy as np
import pytest

import svo_helpers as H


def _scene(yx):
    return yx.SVOData.build_synthetic(depth=6, num_spheres=5, radius_min=3.0, radius_max=8.0, seed=3)


def test_write_overwrite_append_and_counts(yx):
    s = _scene(yx)
    nodes = s.nodes()
    info = s.info()
    n = nodes.shape[0]
    # recolour the leaves of one node: same leaf count, same depth
    i = int(np.nonzero(nodes[:, 0] & 0xFF)[0][0])
    row = nodes[i].copy()
    for sl in range(8):
        if (int(row[0]) >> sl) & 1:
            row[2 + sl] = 0xFF00FF00
    s.write_nodes(i, row)
    assert np.array_equal(s.nodes()[i], row)
    assert s.info() == info
    # append a node under a null slot of some node: one more leaf, one level deeper at most
    par = next(k for k in range(n) if any(not ((int(nodes[k, 0]) >> sl) & 1) and nodes[k, 2 + sl] == H.EMPTY
                                         for sl in range(8)))
    sl = next(q for q in range(8) if not ((int(nodes[par, 0]) >> q) & 1) and nodes[par, 2 + q] == H.EMPTY)
    new = np.full(10, H.EMPTY, np.uint32)
    new[0] = 1 | (0xFE << 8)
    new[2] = 0xFF0000FF
    s.write_nodes(n, new)
    assert s.info()["num_nodes"] == n + 1 and s.info()["num_leaves"] == info["num_leaves"] + 1
    prow = nodes[par].copy()
    prow[2 + sl] = n
    prow[0] &= ~np.uint32(1 << (8 + sl))
    s.write_nodes(int(par), prow)
    walked = H.walk_leaves(s.nodes(), s.GetRoot())
    assert len(walked) == info["num_leaves"] + 1
    assert s.info()["depth"] >= info["depth"]


def test_bad_edits_are_reported(yx):
    s = _scene(yx)
    n = s.info()["num_nodes"]
    with pytest.raises(yx.YoxelError):
        s.write_nodes(n + 1, np.zeros(10, np.uint32))  # past the end
    bad = s.nodes()[s.GetRoot()].copy()
    sl = next(q for q in range(8) if not ((int(bad[0]) >> q) & 1))
    bad[2 + sl] = n + 100  # child id out of range
    s.write_nodes(s.GetRoot(), bad)
    with pytest.raises(yx.YoxelError) as e:
        s.info()
    assert e.value.code == -4  # YX_ERR_FORMAT
    s.set_root(H.EMPTY)
    assert s.info()["depth"] == 0


def test_pages_to_upload_without_sync(yx):
    s = _scene(yx)
    r = yx.SVORenderer()
    n = s.info()["num_nodes"]
    pages, nbytes = r.CountChangedPages(s)
    assert pages == -(-n // 256) and nbytes == 40 * n
