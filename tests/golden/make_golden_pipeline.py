"""Regression fixture of the shading-pipeline restatement (LOD trace, five
depth-limited blurs, Phong shade) on the plain_512 scene: 4 orbit views at
256x256, fov 60, detail coef 1, 3 and 8, SSNA on, one light at the eye with the
demo's parameters. SHA-256 digests of t, blurred z, colour, hit level and
RGBA8. Run this file to regenerate pipeline_256.json."""
import hashlib
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
FIXTURE = HERE / "pipeline_256.json"
KEYS = ("t", "zb", "albedo", "level", "rgba")
LIGHT = dict(diffuse=(0.7, 0.7, 0.7), specular=(0.3, 0.3, 0.3), attenuation=(1.0, 0.0, 0.5))


def digest(out):
    h = hashlib.sha256()
    for k in KEYS:
        a = np.ascontiguousarray(out[k])
        h.update(k.encode())
        h.update(a.astype(a.dtype.newbyteorder("<")).tobytes())
    return h.hexdigest()


def generate():
    sys.path.insert(0, str(HERE.parent.parent))
    import __graft_entry__ as entry
    yx = entry.load_package()
    om = entry.load_oracle()
    po = om.PipelineOracle()
    s = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    nodes, root = s.nodes(), s.GetRoot()
    lod = po.lod_colors(nodes, root)
    views = []
    for i, (pos, d) in enumerate(yx.workloads.orbit_cameras(4, seed=1)):
        for coef in (1.0, 3.0, 8.0):
            lights = om.make_lights([dict(pos=pos, **LIGHT)])
            o = po.pipeline(nodes, root, 8, pos, d, fov=60.0, w=256, h=256, detail_coef=coef, lights=lights,
                            lod=lod, threads=8)
            hit = o["t"] >= 0
            views.append({"view": i, "pos": [float(x) for x in pos], "dir": [float(x) for x in d], "coef": coef,
                          "digest": digest(o), "hits": int(hit.sum()),
                          "lod_stops": int((hit & (o["level"] < 8)).sum()),
                          "blurred": int((hit & (o["zb"] != o["z"])).sum())})
    return {"workload": "plain_512 pipeline", "w": 256, "h": 256, "fov": 60.0, "depth": 8, "light": LIGHT,
            "views": views}


def load():
    return json.loads(FIXTURE.read_text())


if __name__ == "__main__":
    FIXTURE.write_text(json.dumps(generate(), indent=1) + "\n")
    print(f"wrote {FIXTURE}")
