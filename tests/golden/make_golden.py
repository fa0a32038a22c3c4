"""Regression fixtures of the oracle on the plain_512 scene (depth-8 octree of
64 spheres, 8 orbit views): at 512x512 (light at the eye, and sun + shadow rays
for two views) and at 64x64 (both light modes for every view). SHA-256 digests
of every per-pixel output, plus the digest of the synthetic scene's node array.
Run this file to regenerate plain_512.json and plain_64.json."""
import hashlib
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
FIXTURE = HERE / "plain_512.json"
FIXTURE_64 = HERE / "plain_64.json"
KEYS = ("hit_id", "voxel", "t", "rgba", "shadow")


def digest(out):
    h = hashlib.sha256()
    for k in KEYS:
        a = np.ascontiguousarray(out[k])
        h.update(k.encode())
        h.update(a.astype(a.dtype.newbyteorder("<")).tobytes())
    return h.hexdigest()


def node_digest(nodes):
    return hashlib.sha256(np.ascontiguousarray(nodes, "<u4").tobytes()).hexdigest()


def generate(w=512, modes_for=lambda i: (0, 2) if i < 2 else (0,)):
    sys.path.insert(0, str(HERE.parent.parent))
    import __graft_entry__ as entry
    yx = entry.load_package()
    orc = entry.load_oracle().Oracle()
    s = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    nodes, root = s.nodes(), s.GetRoot()
    h = w
    views = []
    for i, (pos, d) in enumerate(yx.workloads.orbit_cameras(8, seed=0)):
        for mode in modes_for(i):
            o = orc.render(nodes, root, pos, d, fov=60.0, w=w, h=h, light_mode=mode, light=yx.workloads.SUN,
                           threads=8)
            views.append({"view": i, "pos": [float(x) for x in pos], "dir": [float(x) for x in d],
                          "light_mode": mode, "digest": digest(o),
                          "hits": int((o["hit_id"] != 0xFFFFFFFF).sum()), "rays": o["rays"]})
    return {"workload": "plain_512" if w == 512 else f"plain_512 scene at {w}x{h}", "w": w, "h": h, "fov": 60.0,
            "light": [float(x) for x in yx.workloads.SUN], "nodes_digest": node_digest(nodes),
            "num_nodes": int(nodes.shape[0]), "views": views}


def load(small=False):
    return json.loads((FIXTURE_64 if small else FIXTURE).read_text())


if __name__ == "__main__":
    if "--small-only" not in sys.argv:
        FIXTURE.write_text(json.dumps(generate(), indent=1) + "\n")
        print(f"wrote {FIXTURE}")
    FIXTURE_64.write_text(json.dumps(generate(64, lambda i: (0, 1, 2)), indent=1) + "\n")
    print(f"wrote {FIXTURE_64}")
