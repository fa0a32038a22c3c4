"""GPU parity of the shading pipeline (yx_render_pipeline: LOD trace, five
depth-limited blurs, Phong shading from z; SVORenderer::Render,
demo/SVORenderer.cpp:95-149) against the CPU restatement, through the C ABI.

The kernels repeat the restatement's float operations in its order, so t, z,
albedo, hit level and visit counts are compared bit for bit; RGBA8 is held to
the stated 1-LSB bar and its exact-match count is reported."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = {"rgba": np.uint32, "z": np.float32, "t": np.float32, "albedo": np.uint32, "level": np.uint8,
      "visits": np.uint32}


@pytest.fixture(scope="module")
def po(oracle_mod):
    return oracle_mod.PipelineOracle()


def gpu_pipeline(yx, torch, r, keys=("rgba", "z", "t", "albedo", "level", "visits")):
    out = r.pipeline_tensors(keys)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy().view(DT[k]) for k, v in out.items()}


def setup_renderer(yx, scene, pos, d, w, h, fov=70.0, coef=1.0, ssna=True, lights=(), show_normals=False,
                   renderer=None, tuning=None):
    r = renderer or yx.SVORenderer()
    if tuning:
        r.set_tuning(**tuning)
    r.SetScene(scene)
    r.SetViewPos(pos)
    r.SetViewDir(d)
    r.SetFOV(fov)
    r.SetResolution(w, h)
    r.SetDetailCoef(coef)
    r.SetSSNA(ssna)
    r.SetShowNormals(show_normals)
    for i in range(4):
        r.SetLigth(i, (0, 0, 0), enabled=False)
    for i, lt in enumerate(lights):
        r.SetLigth(i, **lt)
    return r


LIGHT_DEFAULTS = dict(diffuse=(0.7,) * 3, specular=(0.3,) * 3, attenuation=(1.0, 0.0, 0.5))  # SetLigth's


def oracle_lights(oracle_mod, lights):
    return oracle_mod.make_lights([{**LIGHT_DEFAULTS, **lt} for lt in lights])


def compare(g, o, what):
    for k in ("t", "albedo", "level", "visits"):
        a, b = g[k], o[k].reshape(g[k].shape)
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError(f"{what}: {k} differs at {len(bad)} pixels, first {bad[:3].tolist()} "
                                 f"gpu={a[tuple(bad[0])]} oracle={b[tuple(bad[0])]}")
    zb = o["zb"]
    assert np.array_equal(np.isinf(g["z"]), np.isinf(zb)), what
    fin = np.isfinite(zb)
    if not np.array_equal(g["z"][fin], zb[fin]):
        d = np.abs(g["z"][fin] - zb[fin]) / np.abs(zb[fin])
        raise AssertionError(f"{what}: blurred z differs at {(d > 0).sum()} pixels, max rel {d.max():.3g}")
    diff = 0
    for c in range(4):
        a = ((g["rgba"] >> (8 * c)) & 255).astype(np.int64)
        b = ((o["rgba"] >> (8 * c)) & 255).astype(np.int64)
        diff = max(diff, int(np.abs(a - b).max()))
    exact = float(np.mean(g["rgba"] == o["rgba"]))
    print(f"{what}: rgba exact on {exact:.6f} of pixels, max channel diff {diff}", flush=True)
    assert diff <= 1, what
    return exact


def test_pipeline_small_scenes(yx, po, oracle_mod, cuda):
    s = yx.SVOData.build_synthetic(depth=7, num_spheres=12, radius_min=4.0, radius_max=14.0, heightfield=True,
                                   seed=6)
    nodes, root = s.nodes(), s.GetRoot()
    lod = po.lod_colors(nodes, root)
    cases = [
        # (pos, dir, w, h, coef, ssna, lights, show_normals)
        ((0.55, 0.48, 0.42), (0.3, 0.2, -0.25), 96, 64, 2.0, True,
         [dict(pos=(0.55, 0.48, 0.42))], False),
        ((1.6, 1.3, 1.1), (-1.1, -0.8, -0.6), 160, 120, 1.0, True,
         [dict(pos=(1.6, 1.3, 1.1)), dict(pos=(0.5, 0.5, 2.0), diffuse=(0.2, 0.4, 0.9), specular=(0.5, 0.5, 0.5),
                                          attenuation=(1.0, 10.0, 400.0))], False),
        ((0.5, 0.2, 0.6), (0.0, 1.0, -0.3), 131, 77, 0.0, True, [dict(pos=(0.5, 0.2, 0.6))], False),
        ((0.5, 0.2, 0.6), (0.0, 1.0, -0.3), 131, 77, 8.0, False, [dict(pos=(0.5, 0.2, 0.6))], False),
        ((0.3, 0.7, 0.5), (0.6, -0.2, -0.1), 64, 96, 1.0, True, [], True),
        ((0.5, 0.5, 0.5), (1.0, 0.0, 0.0), 33, 17, 1.0, True, [dict(pos=(0.4, 0.5, 0.5))], False),  # inside
        ((1.6, 1.3, 1.1), (-1.1, -0.8, -0.6), 160, 120, 16.0, True, [dict(pos=(1.6, 1.3, 1.1))], False),
        ((2.5, 2.0, 1.5), (-2.0, -1.5, -1.0), 96, 64, 4.0, True,
         [dict(pos=(2.5, 2.0, 1.5)), dict(pos=(0.2, 0.9, 1.4), attenuation=(0.5, 0.0, 0.1))], False),
    ]
    r = yx.SVORenderer()
    for i, (pos, d, w, h, coef, ssna, lights, shown) in enumerate(cases):
        print(f"case {i}: {w}x{h} coef {coef} ssna {ssna} lights {len(lights)} normals {shown}", flush=True)
        setup_renderer(yx, s, pos, d, w, h, coef=coef, ssna=ssna, lights=lights, show_normals=shown, renderer=r)
        g = gpu_pipeline(yx, cuda, r)
        o = po.pipeline(nodes, root, 7, pos, d, w=w, h=h, detail_coef=coef, ssna=ssna,
                        lights=oracle_lights(oracle_mod, lights), show_normals=shown, lod=lod)
        compare(g, o, f"case {i}")
        hit = o["t"] >= 0
        if i in (3, 6, 7):
            assert np.any(o["level"][hit] < 7), "no LOD stop exercised"
        if i in (0, 1, 6):
            assert np.any(o["zb"][hit] != o["z"][hit]), "no blur exercised"


@pytest.mark.parametrize("tuning", [dict(block=64, tile_w=8, tile_h=8), dict(block=256, tile_w=16, tile_h=4,
                                                                             refill_min=8, queues=8)])
def test_large_1080p_frames(yx, po, oracle_mod, cuda, tuning):
    wl = yx.workloads
    s = yx.SVOData.build_synthetic_gpu(**wl.LARGE_1080P)
    nodes, root = s.nodes(), s.GetRoot()
    lod = po.lod_colors(nodes, root)
    cams = wl.orbit_cameras(32, seed=0)
    r = yx.SVORenderer()
    near = (0.62, 0.41, 0.46)
    views = [(cams[0][0], cams[0][1], 1.0), (cams[13][0], cams[13][1], 4.0), (near, (-0.2, 0.3, -0.35), 1.0)]
    for k, (pos, d, coef) in enumerate(views):
        lights = [dict(pos=tuple(float(v) for v in pos))]
        print(f"1080p view {k} coef {coef} tuning {tuning}", flush=True)
        setup_renderer(yx, s, pos, d, 1920, 1080, coef=coef, lights=lights, renderer=r, tuning=tuning)
        g = gpu_pipeline(yx, cuda, r)
        o = po.pipeline(nodes, root, 11, pos, d, w=1920, h=1080, detail_coef=coef,
                        lights=oracle_lights(oracle_mod, lights), lod=lod, threads=16)
        compare(g, o, f"1080p view {k}")
        hit = o["t"] >= 0
        assert hit.mean() > 0.1
        if k == 1:
            assert np.any(o["level"][hit] < 11)
        if k == 2:
            assert np.mean(o["zb"][hit] != o["z"][hit]) > 0.05


def test_empty_scene_tiny_frames_and_resize(yx, po, oracle_mod, cuda):
    empty = yx.SVOData.build_synthetic(depth=5, num_spheres=0)
    r = setup_renderer(yx, empty, (1.5, 0.5, 0.5), (-1, 0, 0), 7, 3, lights=[dict(pos=(1.5, 0.5, 0.5))])
    g = gpu_pipeline(yx, cuda, r)
    assert np.all(g["rgba"] == 0) and np.all(np.isinf(g["z"])) and np.all(g["t"] == -1)
    assert np.all(g["level"] == 0) and np.all(g["albedo"] == 0)
    s = yx.SVOData.build_synthetic(depth=6, num_spheres=6, radius_min=3.0, radius_max=9.0, seed=12)
    nodes, root = s.nodes(), s.GetRoot()
    pos, d = (1.4, 1.2, 0.9), (-0.9, -0.7, -0.4)
    for (w, h) in ((1, 1), (5, 200), (300, 2), (64, 64), (200, 150), (17, 9)):
        setup_renderer(yx, s, pos, d, w, h, coef=3.0, lights=[dict(pos=pos)], renderer=r)
        g = gpu_pipeline(yx, cuda, r)
        o = po.pipeline(nodes, root, 6, pos, d, w=w, h=h, detail_coef=3.0,
                        lights=oracle_lights(oracle_mod, [dict(pos=pos)]))
        compare(g, o, f"{w}x{h}")


def test_null_outputs_and_repeat(yx, po, cuda):
    s = yx.SVOData.build_synthetic(depth=6, num_spheres=6, radius_min=3.0, radius_max=9.0, seed=12)
    r = setup_renderer(yx, s, (0.6, 0.4, 0.5), (0.2, 0.5, -0.3), 120, 80, lights=[dict(pos=(0.6, 0.4, 0.5))])
    a = gpu_pipeline(yx, cuda, r)
    b = gpu_pipeline(yx, cuda, r, keys=("rgba",))
    c = gpu_pipeline(yx, cuda, r, keys=("z", "level"))
    assert np.array_equal(a["rgba"], b["rgba"])
    assert np.array_equal(a["z"], c["z"]) and np.array_equal(a["level"], c["level"])


def test_pipeline_golden_fixture(yx, cuda):
    """GPU pipeline outputs hash to the committed oracle fixture."""
    import golden.make_golden_pipeline as mp
    gold = mp.load()
    s = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    r = yx.SVORenderer()
    for ent in gold["views"]:
        lt = {"pos": tuple(ent["pos"]), **{k: tuple(v) for k, v in gold["light"].items()}}
        setup_renderer(yx, s, ent["pos"], ent["dir"], gold["w"], gold["h"], fov=gold["fov"], coef=ent["coef"],
                       lights=[lt], renderer=r)
        g = gpu_pipeline(yx, cuda, r, keys=("rgba", "z", "t", "albedo", "level"))
        g["zb"] = g.pop("z")
        assert mp.digest(g) == ent["digest"], f"view {ent['view']} coef {ent['coef']}"
