"""Incremental scene sync on the GPU (yx_renderer_update_scene; CudaSVO::Update,
demo/SVORenderer.cpp:39-53): the traversal records and LOD colours rebuilt on
the GPU from the page-synced device copy equal the host re-layout of
yx_renderer_set_scene_host bit for bit, only edited pages are uploaded, and a
malformed edit is rejected without disturbing the records in place."""
import numpy as np
import pytest

import svo_helpers as H

pytestmark = pytest.mark.gpu


def same_layout(yx, scene, r_upd=None, what=""):
    r_ref = yx.SVORenderer()
    r_ref.SetSceneHost(scene)
    r = r_upd or yx.SVORenderer()
    r.UpdateScene(scene, stream=0)
    a, la = r_ref.layout()
    b, lb = r.layout()
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if not np.array_equal(a, b):
        bad = np.argwhere(a != b)
        raise AssertionError(f"{what}: records differ at {len(bad)} places, first {bad[:3].tolist()}")
    assert np.array_equal(la, lb), f"{what}: LOD colours differ"
    return r


def render(yx, cuda, r, pos=(1.6, 1.3, 1.1), d=(-1.1, -0.8, -0.6), w=96, h=64):
    r.SetViewPos(pos)
    r.SetViewDir(d)
    r.SetFOV(60.0)
    r.SetResolution(w, h)
    r.SetLight((0.4, 0.3, 0.85), yx.YX_LIGHT_SUN_SHADOW)
    out = r.render_tensors(("hit_id", "t", "rgba", "shadow"))
    cuda.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def test_layout_equals_host_relayout(yx, cuda):
    rng = np.random.default_rng(9)
    scenes = [
        ("depth 5", yx.SVOData.build_synthetic(depth=5, num_spheres=4, radius_min=2.0, radius_max=5.0, seed=2)),
        ("depth 8 heightfield", yx.SVOData.build_synthetic(depth=8, num_spheres=20, radius_min=4.0,
                                                           radius_max=20.0, heightfield=True, seed=5)),
        ("large_1080p", yx.SVOData.build_synthetic_gpu(**yx.workloads.LARGE_1080P)),
        ("empty", yx.SVOData.build_synthetic(depth=5, num_spheres=0)),
    ]
    occ = rng.random((16, 16, 16)) < 0.2
    nodes, root = H.tree_from_leaves(H.dense_leaves(occ, rng), order="shuffle", seed=4)
    scenes.append(("shuffled ids", yx.SVOData.from_nodes(nodes, root)))
    leaf_root = np.full((1, 10), H.EMPTY, np.uint32)
    leaf_root[0, 0] = 1 | (0xFE << 8)
    leaf_root[0, 2] = 0xFF102030
    scenes.append(("one leaf", yx.SVOData.from_nodes(leaf_root, 0)))
    for what, s in scenes:
        print(what, s.info(), flush=True)
        r = same_layout(yx, s, what=what)
        pages, nbytes = r.last_update()
        n = s.info()["num_nodes"]
        assert pages == -(-n // 256) and nbytes == 40 * n


def test_edits_upload_only_changed_pages(yx, cuda, orc):
    s = yx.SVOData.build_synthetic(depth=7, num_spheres=12, radius_min=4.0, radius_max=14.0, heightfield=True,
                                   seed=6)
    r = same_layout(yx, s, what="initial")
    n = s.info()["num_nodes"]
    assert r.CountChangedPages(s) == (0, 0)
    r.UpdateScene(s, stream=0)
    assert r.last_update() == (0, 0)
    nodes = s.nodes()

    # 1. recolour the leaves of a few nodes in one page
    cand = np.nonzero(nodes[:, 0] & 0xFF)[0]
    page = int(cand[len(cand) // 2]) // 256
    idx = [int(i) for i in cand if int(i) // 256 == page][:10]
    assert len(idx) == 10 and (page + 1) * 256 <= n
    for i in idx:
        row = nodes[i].copy()
        for sl in range(8):
            if (int(row[0]) >> sl) & 1:
                row[2 + sl] = 0xFF0000FF
        s.write_nodes(i, row)
    assert r.CountChangedPages(s) == (1, 256 * 40)
    same_layout(yx, s, r, "recolour")
    assert r.last_update() == (1, 256 * 40)

    # 2. a new node under a null slot of node 3: appended page + node 3's page
    nodes = s.nodes()
    par = next(k for k in range(3, n) if any(not ((int(nodes[k, 0]) >> q) & 1) and nodes[k, 2 + q] == H.EMPTY
                                             for q in range(8)))
    q = next(q for q in range(8) if not ((int(nodes[par, 0]) >> q) & 1) and nodes[par, 2 + q] == H.EMPTY)
    new = np.full(10, H.EMPTY, np.uint32)
    new[0] = 0x81 | (0x7E << 8)
    new[2], new[9] = 0xFF00FF00, 0xFFFF0000
    s.write_nodes(n, new)
    prow = nodes[par].copy()
    prow[2 + q] = n
    prow[0] &= ~np.uint32(1 << (8 + q))
    s.write_nodes(int(par), prow)
    pages, _ = r.CountChangedPages(s)
    assert pages == len({par // 256, n // 256})
    same_layout(yx, s, r, "append")

    # 3. a malformed edit: the update fails, the records stay as they were
    before = r.layout()
    bad = s.nodes()[par].copy()
    bad[2 + q] = n + 1000
    s.write_nodes(int(par), bad)
    with pytest.raises(yx.YoxelError) as e:
        r.UpdateScene(s, stream=0)
    assert e.value.code == -4
    after = r.layout()
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    # 4. repaired: only that page is sent
    s.write_nodes(int(par), prow)
    assert r.CountChangedPages(s)[0] == 1
    same_layout(yx, s, r, "repaired")

    # 5. a new root (a child of the old root)
    old_root = s.GetRoot()
    rn = s.nodes()[old_root]
    child = next(int(rn[2 + k]) for k in range(8) if not ((int(rn[0]) >> k) & 1) and rn[2 + k] != H.EMPTY)
    s.set_root(child)
    assert r.CountChangedPages(s) == (0, 0)  # no page changed, only the root
    same_layout(yx, s, r, "new root")

    # 6. the renders match the CPU restatement of the edited scene
    s.set_root(old_root)
    r.UpdateScene(s, stream=0)
    g = render(yx, cuda, r)
    o = orc.render(s.nodes(), s.GetRoot(), (1.6, 1.3, 1.1), (-1.1, -0.8, -0.6), fov=60.0, w=96, h=64,
                   light_mode=2, light=(0.4, 0.3, 0.85))
    assert np.array_equal(g["hit_id"].view(np.uint32), o["hit_id"])
    assert np.array_equal(g["rgba"].view(np.uint32), o["rgba"])


def test_switching_scenes_uploads_everything(yx, cuda):
    a = yx.SVOData.build_synthetic(depth=6, num_spheres=5, radius_min=3.0, radius_max=8.0, seed=3)
    b = yx.SVOData.build_synthetic(depth=7, num_spheres=9, radius_min=3.0, radius_max=9.0, seed=4)
    r = same_layout(yx, a, what="a")
    r = same_layout(yx, b, r, "b")
    assert r.last_update()[0] == -(-b.info()["num_nodes"] // 256)
    r = same_layout(yx, a, r, "a again")
    assert r.last_update()[0] == -(-a.info()["num_nodes"] // 256)


def test_scattered_pages_in_one_staged_copy(yx, cuda):
    """Many scattered dirty pages go through one pinned staging copy and a
    scatter kernel; the device copy must equal a full upload."""
    s = yx.SVOData.build_synthetic(depth=8, num_spheres=20, radius_min=4.0, radius_max=20.0, heightfield=True,
                                   seed=5)
    r = same_layout(yx, s, what="initial")
    nodes = s.nodes()
    n = nodes.shape[0]
    cand = np.nonzero(nodes[:, 0] & 0xFF)[0]
    rng = np.random.default_rng(12)
    for rnd in range(3):
        pages = sorted(rng.choice(-(-n // 256), 60, replace=False))
        edited = set()
        for p in pages:
            in_page = cand[(cand // 256) == p]
            if in_page.size == 0:
                continue
            i = int(in_page[rng.integers(in_page.size)])
            row = s.nodes()[i].copy()
            for sl in range(8):
                if (int(row[0]) >> sl) & 1:
                    row[2 + sl] = np.uint32(0xFF000000 | (rnd * 0x10203 + p))
            s.write_nodes(i, row)
            edited.add(p)
        assert len(edited) >= 40
        assert r.CountChangedPages(s)[0] == len(edited)
        same_layout(yx, s, r, f"scattered round {rnd}")
        assert r.last_update()[0] == len(edited)
