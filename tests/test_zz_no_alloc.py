"""The allocation-free render entry point (yx_render_noalloc) against yx_render,
through the C ABI: identical outputs on two calls in a row, the second at a
smaller resolution, and under a multi-queue tuning."""
import ctypes

import numpy as np
import pytest


def _outputs(torch, w, h):
    dev = torch.device("cuda", 0)
    return {"hit_id": torch.full((h, w), 7, dtype=torch.int32, device=dev),
            "voxel": torch.full((h, w, 3), 7, dtype=torch.int32, device=dev),
            "t": torch.full((h, w), 7.0, dtype=torch.float32, device=dev),
            "rgba": torch.full((h, w), 7, dtype=torch.int32, device=dev),
            "shadow": torch.full((h, w), 7, dtype=torch.uint8, device=dev),
            "visits": torch.full((h, w), 7, dtype=torch.int32, device=dev)}


def _render(yx, torch, r, fn, w, h):
    L = yx._lib
    o = _outputs(torch, w, h)
    s = L.Outputs(*(ctypes.c_void_p(o[k].data_ptr()) for k in ("hit_id", "voxel", "t", "rgba", "shadow",
                                                                  "visits")))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.check(getattr(L.lib(), fn)(r.handle, ctypes.byref(s), stream))
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in o.items()}


@pytest.mark.gpu
def test_noalloc_equals_render(yx, cuda):
    torch = cuda
    print("no-alloc: scene depth 7, 320x240 then 96x64, then 8 queues", flush=True)
    scene = yx.SVOData.build_synthetic(depth=7, num_spheres=24, radius_min=4.0, radius_max=16.0,
                                       heightfield=True, seed=21)
    pos, d = yx.workloads.orbit_cameras(3, seed=5)[1]
    rs = []
    for _ in range(2):  # the second renderer never calls yx_render: no counters of its own
        r = yx.SVORenderer()
        r.SetScene(scene)
        r.SetViewPos(pos)
        r.SetViewDir(d)
        r.SetFOV(60.0)
        r.SetLight(yx.workloads.SUN, yx.YX_LIGHT_SUN_SHADOW)
        rs.append(r)
    for (w, h, tune) in [(320, 240, {}), (96, 64, {}), (200, 72, {"queues": 8, "refill_min": 40})]:
        for r in rs:
            r.SetResolution(w, h)
            if tune:
                r.set_tuning(**tune)
        ref = _render(yx, torch, rs[0], "yx_render", w, h)
        got = _render(yx, torch, rs[1], "yx_render_noalloc", w, h)
        assert (ref["hit_id"] != -1).any() and (ref["shadow"] < 2).any()
        for k in ref:
            assert np.array_equal(ref[k], got[k]), f"{k} differs at {w}x{h} {tune}"
    # the scratch holds 64 queues plus the retired count: the largest tuning fits
    rs[1].set_tuning(queues=64)
    got = _render(yx, torch, rs[1], "yx_render_noalloc", 200, 72)
    assert np.array_equal(ref["rgba"], got["rgba"])


@pytest.mark.gpu
def test_noalloc_needs_a_scene(yx, cuda):
    r = yx.SVORenderer()
    L = yx._lib
    o = L.Outputs()
    rc = L.lib().yx_render_noalloc(r.handle, ctypes.byref(o), ctypes.c_void_p(0))
    assert rc == L.YX_ERR_NO_SCENE
