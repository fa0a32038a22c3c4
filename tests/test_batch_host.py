"""Host-side logic of the batch and autotuner interfaces (no GPU): argument
checks, tuning validation, and the batched workload recipe."""
import ctypes

import pytest


def test_batch_argument_checks(yx):
    L = yx.lib()
    from yoxel_voxel_amd import _lib
    b = yx.BatchRenderer()
    # no scenes yet: any job names a missing scene
    with pytest.raises(yx.YoxelError, match="scene 0 not in the batch"):
        b.set_jobs([dict(scene=0, width=4, height=4, pos=(2, 2, 2), dir=(-1, -1, -1))])
    assert b.set_jobs([]) == 0
    b.render(stream=0)  # nothing to do: returns before touching a device
    for bad, msg in [(dict(kernel=0), "persistent stack kernel"), (dict(block=512), "block"),
                     (dict(tile_w=3), "powers of two"), (dict(refill_min=0), "refill_min"),
                     (dict(queues=65), "queues"), (dict(window=6), "window"), (dict(waves_per_cu=33), "waves")]:
        with pytest.raises(yx.YoxelError, match=msg):
            b.set_tuning(**bad)
    b.set_tuning(tile_w=16, tile_h=4, window=2, block=128)
    t = b.get_tuning()
    assert (t["tile_w"], t["tile_h"], t["window"], t["block"]) == (16, 4, 2, 128)
    assert b.scene_bytes() == 0
    o = ctypes.c_uint64()
    assert L.yx_batch_job_offset(b.handle, 0, ctypes.byref(o)) == _lib.YX_ERR_INVALID_ARG
    assert L.yx_batch_autotune(None, None, 1, None, 0, None, None) == _lib.YX_ERR_INVALID_ARG
    assert L.yx_renderer_autotune(None, None, 1, None, 0, None, None) == _lib.YX_ERR_INVALID_ARG
    r = yx.SVORenderer()
    out = _lib.Outputs()
    assert L.yx_renderer_autotune(r.handle, ctypes.byref(out), 1, None, 0, None, None) == _lib.YX_ERR_NO_SCENE


def test_batched_recipes(yx):
    W = yx.workloads
    rec = W.batched_recipes()
    assert len(rec) == 48
    depths = [d["depth"] for d, _ in rec]
    assert sorted(set(depths)) == list(range(6, 12)) and all(depths.count(k) == 8 for k in range(6, 12))
    assert sum(d["heightfield"] for d, _ in rec) == 24
    cams = [c for _, cs in rec for c in cs]
    assert all(1 <= len(cs) <= 16 for _, cs in rec)
    assert {(c["width"], c["height"]) for c in cams} == set(W.RESOLUTIONS)
    assert len({d["seed"] for d, _ in rec}) == 48
    jobs = W.batch_jobs(rec)
    assert len(jobs) == len(cams) == 416
    assert sum(j["width"] * j["height"] for j in jobs) == 348_521_472
    assert W.batched_recipes() == rec  # deterministic


def test_ntree_argument_checks(yx):
    """Shape and recipe checks of the n-tree entry points run before any
    device call."""
    import numpy as np
    for size, levels, msg in [(100, 1, "power of two"), (64, 3, "8 \\* 4"), (4, 0, "power of two"),
                              (64, 5, "levels"), (512, 0, "root grid")]:
        with pytest.raises(yx.YoxelError, match=msg):
            yx.BrickTree.build_fbm(size=size, levels=levels, octaves=4, frequency=4.0, threshold=0.3,
                                   scale=100.0, seed=1, stream=0)
    with pytest.raises(yx.YoxelError, match="octaves"):
        yx.BrickTree.build_fbm(size=64, levels=1, octaves=0, frequency=4.0, threshold=0.3, scale=100.0, seed=1,
                               stream=0)
    with pytest.raises(yx.YoxelError, match="power of two"):
        yx.BrickTree.from_dense(np.zeros((24, 24, 24), np.uint8), 0)


def test_gpu_scene_builder_argument_checks(yx):
    for kw, msg in [(dict(depth=0), "depth"), (dict(depth=13), "depth"),
                    (dict(depth=5, num_spheres=-1), "num_spheres"),
                    (dict(depth=5, heightfield=True, hf_octaves=0), "hf_octaves")]:
        with pytest.raises(yx.YoxelError, match=msg):
            yx.SVOData.build_synthetic_gpu(stream=0, **kw)
