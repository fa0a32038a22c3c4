"""GPU frames with colour16 + normal16 leaf payloads (YX_PAYLOAD_C16N16) equal
the oracle's bit for bit, for a scene saved with yx_scene_save_vox and read
back with yx_scene_load_vox; the same file read as RGBA8 also matches."""
import numpy as np
import pytest

import svo_helpers as H
from test_payload_c16n16 import payload

pytestmark = pytest.mark.gpu


def test_c16n16_scene_from_file(yx, orc, cuda, tmp_path):
    print("payload c16n16: depth-7 scene through a .vox file, 3 light modes x 2 formats", flush=True)
    s = yx.SVOData.build_synthetic(depth=7, num_spheres=20, radius_min=4.0, radius_max=14.0, heightfield=True,
                                   seed=17)
    nodes = s.nodes().copy()
    rng = np.random.default_rng(3)
    leaf = ((nodes[:, 0:1] >> np.arange(8, dtype=np.uint32)) & 1).astype(bool)
    cols = nodes[:, 2:][leaf]
    nrm = rng.normal(size=(cols.size, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    sub = nodes[:, 2:].copy()
    sub[leaf] = payload(cols, nrm)
    nodes[:, 2:] = sub
    fn = tmp_path / "c16n16.vox"
    yx.SVOData.from_nodes(nodes, s.GetRoot()).Save(fn)
    loaded = yx.SVOData()
    loaded.Load(fn)
    assert np.array_equal(loaded.nodes(), nodes)
    r = yx.SVORenderer()
    r.SetScene(loaded)
    pos, d = yx.workloads.orbit_cameras(4, seed=2)[1]
    r.SetViewPos(pos)
    r.SetViewDir(d)
    r.SetFOV(60.0)
    r.SetResolution(160, 120)
    for fmt in (yx.YX_PAYLOAD_C16N16, yx.YX_PAYLOAD_RGBA8):
        r.SetPayloadFormat(fmt)
        for mode in (0, 1, 2):
            r.SetLight(yx.workloads.SUN, mode)
            out = r.render_tensors(outputs=("hit_id", "voxel", "t", "rgba", "shadow", "visits"))
            cuda.cuda.synchronize()
            ref = orc.render(nodes, s.GetRoot(), pos, d, fov=60.0, w=160, h=120, light_mode=mode,
                             light=yx.workloads.SUN, payload=fmt)
            for k in ("hit_id", "voxel", "t", "rgba", "shadow", "visits"):
                g = out[k].cpu().numpy().view(ref[k].dtype).reshape(ref[k].shape)
                assert np.array_equal(g, ref[k]), f"{k} differs (format {fmt}, light mode {mode})"
    assert (ref["hit_id"] != H.MISS).sum() > 1000


def test_dump_trace_data(yx, orc, cuda, tmp_path):
    """SVORenderer::DumpTraceData (demo/SVORenderer.cpp:158-192): the three
    files of a C16N16 frame hold zeros (distance), the hit leaves' decoded
    colours and normals, zeros for misses; the hit ids are the oracle's."""
    print("payload c16n16: DumpTraceData of a 96x64 frame", flush=True)
    s = yx.SVOData.build_synthetic(depth=6, num_spheres=10, radius_min=3.0, radius_max=9.0, heightfield=True,
                                   seed=23)
    nodes = s.nodes().copy()
    leaf = ((nodes[:, 0:1] >> np.arange(8, dtype=np.uint32)) & 1).astype(bool)
    nrm = np.random.default_rng(4).normal(size=(int(leaf.sum()), 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    sub = nodes[:, 2:].copy()
    sub[leaf] = payload(sub[leaf], nrm)
    nodes[:, 2:] = sub
    scene = yx.SVOData.from_nodes(nodes, s.GetRoot())
    r = yx.SVORenderer()
    r.SetScene(scene)
    r.SetPayloadFormat(yx.YX_PAYLOAD_C16N16)
    pos, d = yx.workloads.orbit_cameras(4, seed=5)[2]
    r.SetViewPos(pos)
    r.SetViewDir(d)
    r.SetFOV(60.0)
    r.SetResolution(96, 64)
    base = r.DumpTraceData(str(tmp_path / "dmp_0"))
    assert base.endswith("dmp_0_96x64")
    dist = np.fromfile(base + ".dist", dtype=np.float32)
    color = np.fromfile(base + ".color", dtype=np.uint32)
    normal = np.fromfile(base + ".normal", dtype=np.float32).reshape(-1, 3)
    assert dist.shape == (96 * 64,) and not dist.any()
    ref = orc.render(nodes, s.GetRoot(), pos, d, fov=60.0, w=96, h=64, light_mode=0, payload=1)
    ids = ref["hit_id"].reshape(-1)
    assert (ids != H.MISS).sum() > 500
    for i in range(ids.size):
        if ids[i] == H.MISS:
            assert color[i] == 0 and not normal[i].any()
        else:
            c, n = orc.decode_c16n16(int(nodes[ids[i] >> 3, 2 + (ids[i] & 7)]))
            assert color[i] == c and np.array_equal(normal[i], n), i
