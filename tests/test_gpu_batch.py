"""GPU parity of batched rendering (yx_render_batch): one persistent launch over
many (scene, camera) jobs gives, job by job, exactly the outputs of rendering
each job alone with yx_render, and the oracle's outputs; graph replays,
camera updates and every launch shape leave the outputs unchanged; the
autotuners return a valid launch shape that renders identically."""
import json

import numpy as np
import pytest

import svo_helpers as H
from test_gpu_parity import KEYS, TUNINGS, assert_same

pytestmark = pytest.mark.gpu

DT = {"hit_id": np.uint32, "voxel": np.int32, "t": np.float32, "rgba": np.uint32, "shadow": np.uint8,
      "visits": np.uint32}

BATCH_TUNINGS = [
    dict(),  # library default
    dict(tile_w=16, tile_h=4, refill_min=8, queues=8),
    dict(block=128, window=2, tile_w=4, tile_h=16, refill_min=48, queues=3),
    dict(block=256, window=4, tile_w=32, tile_h=8, refill_min=1, queues=64, waves_per_cu=2),
    dict(block=64, window=0, tile_w=1, tile_h=1, refill_min=64, queues=1, waves_per_cu=1),
]


def small_batch(yx):
    """Six scenes (depth 3..9, one empty, one with mixed-level leaves and
    shuffled node ids) and 14 jobs with ragged resolutions and all light modes."""
    rng = np.random.default_rng(4)
    occ = rng.random((16, 16, 16)) < 0.08
    leaves = H.dense_leaves(occ, rng) + [(2, 0, 0, 0, 77), (1, 1, 1, 1, 99)]
    leaves = [l for l in leaves if not (l[0] == 4 and l[1] < 4 and l[2] < 4 and l[3] < 4)]
    leaves = [l for l in leaves if not (l[0] == 4 and l[1] >= 8 and l[2] >= 8 and l[3] >= 8)]
    nodes, root = H.tree_from_leaves(leaves, order="shuffle", seed=5)
    scenes = [
        yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512),
        yx.SVOData.from_nodes(nodes, root),
        yx.SVOData.build_synthetic(depth=5, num_spheres=0),  # empty
        yx.SVOData.build_synthetic(depth=9, num_spheres=40, radius_min=6, radius_max=30, heightfield=True,
                                   hf_octaves=4, hf_amplitude=0.2, hf_base=0.3, seed=21),
        yx.SVOData.build_synthetic(depth=3, num_spheres=3, radius_min=1, radius_max=3, seed=2),
        yx.SVOData.build_synthetic(depth=7, num_spheres=10, radius_min=4, radius_max=12, seed=8),
    ]
    sun = yx.workloads.SUN
    cams = yx.workloads.orbit_cameras(5, seed=9)
    jobs = []
    res = [(37, 29), (64, 48), (1, 1), (128, 96), (33, 70), (96, 16), (200, 120)]
    k = 0
    for si in (0, 1, 3, 2, 5, 0, 4, 3, 1, 0, 5, 3, 2, 1):
        pos, d = cams[k % 5]
        w, h = res[k % len(res)]
        mode = k % 3
        jobs.append(dict(scene=si, width=w, height=h, pos=pos, dir=d, fov=50.0 + 3 * k, light_mode=mode,
                         light=sun))
        k += 1
    # eye inside the cube and an axis-aligned view
    jobs.append(dict(scene=1, width=40, height=40, pos=(0.3, 0.45, 0.6), dir=(1.0, 0.2, -0.1), fov=70.0,
                     light_mode=2, light=sun))
    jobs.append(dict(scene=0, width=31, height=33, pos=(0.5, 0.5, 3.0), dir=(0.0, 0.0, -1.0), fov=40.0,
                     light_mode=0, light=sun))
    return scenes, jobs


def to_np(out):
    return {k: v.cpu().numpy().view(DT[k]) for k, v in out.items()}


def job_view(arr, key, off, w, h):
    n = w * h
    if key == "voxel":
        return arr[off:off + n].reshape(h, w, 3)
    return arr[off:off + n].reshape(h, w)


def render_alone(yx, torch, scenes, jobs, keys=KEYS):
    """Per-scene plain launches: one SVORenderer per scene, one Render per job."""
    rs = {}
    res = []
    for j in jobs:
        r = rs.get(j["scene"])
        if r is None:
            r = rs[j["scene"]] = yx.SVORenderer()
            r.SetScene(scenes[j["scene"]])
        r.SetViewPos(j["pos"])
        r.SetViewDir(j["dir"])
        r.SetViewUp(j.get("up", (0.0, 0.0, 1.0)))
        r.SetFOV(j["fov"])
        r.SetResolution(j["width"], j["height"])
        r.SetLight(j["light"], j["light_mode"])
        res.append(r.render_tensors(keys))
    torch.cuda.synchronize()
    return [to_np(o) for o in res]


def check_batch(b, got, alone, keys=KEYS, what=""):
    for i, ((off, w, h), ref) in enumerate(zip(b.job_slices(), alone)):
        g = {k: job_view(got[k], k, off, w, h) for k in keys}
        assert_same(g, ref, keys, what=f"{what} job {i} ({w}x{h})")


def test_small_batch_matches_alone_and_oracle(yx, orc, cuda):
    scenes, jobs = small_batch(yx)
    b = yx.BatchRenderer()
    b.set_scenes(scenes)
    total = b.set_jobs(jobs)
    assert total == sum(j["width"] * j["height"] for j in jobs)
    alone = render_alone(yx, cuda, scenes, jobs)
    for i, j in enumerate(jobs):
        s = scenes[j["scene"]]
        o = orc.render(s.nodes(), s.GetRoot(), j["pos"], j["dir"], fov=j["fov"], w=j["width"], h=j["height"],
                       light_mode=j["light_mode"], light=j["light"], threads=8)
        assert_same(alone[i], o, what=f"alone job {i}")
    for tu in BATCH_TUNINGS:
        b.set_tuning(**{**dict(block=64, window=0, tile_w=8, tile_h=8, refill_min=32, queues=1,
                               waves_per_cu=0), **tu})
        for use_graph in (True, False):
            b.set_graph(use_graph)
            out = b.alloc_outputs(KEYS)
            for rep in range(3):
                print(f"small batch tuning {tu} graph {use_graph} rep {rep}", flush=True)
                for v in out.values():
                    v.fill_(0x5A)  # eager work between replays; stale values must not survive
                b.render(**out)
                cuda.cuda.synchronize()
                check_batch(b, to_np(out), alone, what=f"tuning {tu} graph {use_graph} rep {rep}")


def test_camera_update_keeps_graph_and_results(yx, cuda):
    """Changing only cameras and lights between frames replays the captured
    graph with the new job table; outputs follow the new cameras."""
    scenes, jobs = small_batch(yx)
    b = yx.BatchRenderer()
    b.set_scenes(scenes)
    b.set_jobs(jobs)
    out = b.alloc_outputs(KEYS)
    b.render(**out)
    cams = yx.workloads.orbit_cameras(7, seed=31)
    for frame in range(3):
        moved = []
        for i, j in enumerate(jobs):
            pos, d = cams[(i + frame) % 7]
            moved.append(dict(j, pos=pos, dir=d, light_mode=(j["light_mode"] + frame) % 3))
        b.set_jobs(moved)
        b.render(**out)
        cuda.cuda.synchronize()
        check_batch(b, to_np(out), render_alone(yx, cuda, scenes, moved), what=f"frame {frame}")


def test_batched_48_matches_per_scene_launches(yx, orc, cuda):
    """The full batched workload (48 scenes of depth 6-11, 416 jobs, 348.5 M
    pixels): every job's outputs equal its own plain launch; the 256x256 jobs
    on scenes up to depth 9 are also checked against the oracle."""
    torch = cuda
    recipes = yx.workloads.batched_recipes()
    jobs = yx.workloads.batch_jobs(recipes)
    print("building 48 scenes", flush=True)
    scenes = [yx.SVOData.build_synthetic(**d) for d, _ in recipes]
    b = yx.BatchRenderer()
    b.set_scenes(scenes)
    total = b.set_jobs(jobs)
    assert len(jobs) == 416 and total == 348_521_472
    keys = ("hit_id", "voxel", "t", "rgba", "shadow")
    out = b.alloc_outputs(keys)
    b.render(**out)
    torch.cuda.synchronize()
    print("batch rendered", flush=True)
    rs = {}
    checked = 0
    for i, ((off, w, h), j) in enumerate(zip(b.job_slices(), jobs)):
        r = rs.get(j["scene"])
        if r is None:
            rs.clear()  # one scene's renderer at a time (jobs are grouped by scene)
            r = rs[j["scene"]] = yx.SVORenderer()
            r.SetScene(scenes[j["scene"]])
        r.SetViewPos(j["pos"])
        r.SetViewDir(j["dir"])
        r.SetFOV(j["fov"])
        r.SetResolution(w, h)
        r.SetLight(j["light"], j["light_mode"])
        ref = r.render_tensors(keys)
        n = w * h
        for k in keys:
            got = out[k][off:off + n].reshape(ref[k].shape)
            assert torch.equal(got, ref[k]), f"job {i} ({w}x{h}, scene {j['scene']}): {k} differs"
        d = recipes[j["scene"]][0]["depth"]
        if (w, h) == (256, 256) and d <= 9 and checked < 8:
            s = scenes[j["scene"]]
            o = orc.render(s.nodes(), s.GetRoot(), j["pos"], j["dir"], fov=j["fov"], w=w, h=h,
                           light_mode=j["light_mode"], light=j["light"], threads=16)
            assert_same({k: v.cpu().numpy().view(DT[k]) for k, v in ref.items()}, o, keys, what=f"job {i}")
            checked += 1
    assert checked >= 4


def test_batch_autotune_and_cache(yx, cuda, tmp_path):
    scenes, jobs = small_batch(yx)
    b = yx.BatchRenderer()
    b.set_scenes(scenes)
    b.set_jobs(jobs)
    alone = render_alone(yx, cuda, scenes, jobs)
    out = b.alloc_outputs(KEYS)
    cache = tmp_path / "tune.json"
    best, trials = b.autotune(out, reps=1, cache=cache)
    assert len(trials) >= 15 and all(ms > 0 for _, ms in trials)
    assert best == b.get_tuning() and best["kernel"] == 1
    # coordinate descent only moves on a 1 % gain: no trial beats the result by more
    best_ms = [ms for t, ms in trials if t == best][0]
    assert best_ms <= min(ms for _, ms in trials) / 0.99 * 1.0001
    stored = json.loads(cache.read_text())
    assert list(stored.values()) == [best]
    b.set_tuning(tile_w=1, tile_h=1)
    best2, trials2 = b.autotune(out, reps=1, cache=cache)  # from the cache: no timing
    assert best2 == best and trials2 == []
    b.render(**out)
    cuda.cuda.synchronize()
    check_batch(b, to_np(out), alone, what="autotuned")


def test_renderer_autotune(yx, orc, cuda):
    s = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    pos, d = yx.workloads.orbit_cameras(8, seed=0)[3]
    r = yx.SVORenderer()
    r.SetScene(s)
    r.SetViewPos(pos)
    r.SetViewDir(d)
    r.SetFOV(60.0)
    r.SetResolution(512, 512)
    r.SetLight(yx.workloads.SUN, 2)
    out = r.render_tensors(KEYS)
    best, trials = r.autotune(out, reps=1)
    assert len(trials) >= 15 and best == r.get_tuning()
    r.Render(**out)
    cuda.cuda.synchronize()
    o = orc.render(s.nodes(), s.GetRoot(), pos, d, fov=60.0, w=512, h=512, light_mode=2, light=yx.workloads.SUN,
                   threads=8)
    assert_same(to_np(out), o, what=f"autotuned {best}")


def test_batch_c16n16_payloads(yx, orc, cuda):
    """yx_batch_set_payload: scenes whose leaves hold colour16 + normal16 render
    in one batch as each job does alone with the oracle in that format, in
    every light mode; switching the format back re-captures the graph and
    reads the same bytes as RGBA8."""
    from test_payload_c16n16 import payload
    scenes, jobs = small_batch(yx)
    conv = []
    for k, s in enumerate(scenes):
        nodes = s.nodes().copy()
        leaf = ((nodes[:, 0:1] >> np.arange(8, dtype=np.uint32)) & 1).astype(bool)
        if leaf.any():
            nrm = np.random.default_rng(40 + k).normal(size=(int(leaf.sum()), 3))
            nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
            sub = nodes[:, 2:].copy()
            sub[leaf] = payload(sub[leaf], nrm)
            nodes[:, 2:] = sub
        conv.append(yx.SVOData.from_nodes(nodes, s.GetRoot()))
    b = yx.BatchRenderer()
    b.set_scenes(conv)
    b.set_jobs(jobs)
    for fmt in (yx.YX_PAYLOAD_C16N16, yx.YX_PAYLOAD_RGBA8, yx.YX_PAYLOAD_C16N16):
        b.set_payload(fmt)
        print(f"batch payload {fmt}", flush=True)
        out = b.render_tensors(KEYS)
        cuda.cuda.synchronize()
        got = to_np(out)
        for i, ((off, w, h), j) in enumerate(zip(b.job_slices(), jobs)):
            s = conv[j["scene"]]
            o = orc.render(s.nodes(), s.GetRoot(), j["pos"], j["dir"], fov=j["fov"], w=w, h=h,
                           light_mode=j["light_mode"], light=j["light"], payload=fmt, threads=8)
            g = {k: job_view(got[k], k, off, w, h) for k in KEYS}
            assert_same(g, o, KEYS, what=f"payload {fmt} job {i} ({w}x{h})")
    with pytest.raises(yx.YoxelError):
        b.set_payload(5)
