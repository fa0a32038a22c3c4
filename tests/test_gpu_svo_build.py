"""GPU scene builder (yx_scene_build_synthetic_gpu) against the CPU builder:
the node arrays, root and counts are identical byte for byte, for the
workload scenes, a spread of the batched recipes, tiny and empty scenes."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def same_scene(yx, desc):
    t0 = time.perf_counter()
    c = yx.SVOData.build_synthetic(**desc)
    t1 = time.perf_counter()
    g = yx.SVOData.build_synthetic_gpu(**desc)
    t2 = time.perf_counter()
    ci, gi = c.info(), g.info()
    assert ci == gi, (ci, gi)
    if ci["num_nodes"]:
        assert np.array_equal(c.nodes(), g.nodes())
    return ci, t1 - t0, t2 - t1


def test_workload_scenes(yx, cuda):
    for name in ("PLAIN_512", "LARGE_1080P"):
        info, tc, tg = same_scene(yx, getattr(yx.workloads, name))
        print(f"{name}: {info['num_nodes']} nodes, cpu {tc:.3f} s, gpu {tg:.3f} s", flush=True)
        assert info["num_leaves"] > 0


def test_batched_recipes_and_small_scenes(yx, cuda):
    recipes = yx.workloads.batched_recipes()
    for i in (0, 3, 7, 10, 14, 23, 29, 41):
        desc = recipes[i][0]
        print(f"recipe {i}: depth {desc['depth']}", flush=True)
        same_scene(yx, desc)
    for desc in (dict(depth=1, num_spheres=1, radius_min=0.5, radius_max=0.9, seed=3),
                 dict(depth=3, num_spheres=5, radius_min=1.0, radius_max=3.0, heightfield=True, hf_octaves=2,
                      seed=4),
                 dict(depth=5, num_spheres=0),
                 dict(depth=6, num_spheres=0, heightfield=True, hf_octaves=3, hf_amplitude=0.3, hf_base=0.1,
                      seed=8)):
        same_scene(yx, desc)
    info, _, _ = same_scene(yx, dict(depth=5, num_spheres=0))
    assert info["num_leaves"] == 0 and info["root"] == yx.YX_EMPTY_NODE
