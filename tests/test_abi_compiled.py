"""The C ABI from compiled C++ (tests/cabi/abi_check.cpp, built with g++
against include/yoxel.h):
  * CPU: sizeof and every field offset of each yx_* struct agree with the
    Python ctypes mirrors, and the program lists every field the header
    declares;
  * GPU: the reference's driver sequence (testRenderer, cell/main.cpp:21-40)
    through an ISVORenderer that forwards to the C ABI - RenderFrame is NULL
    without a scene, and the 1024x768 frame of a scene read from a .vox file
    equals the oracle's bit for bit."""
import ctypes
import json
import pathlib
import re
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
HERE = pathlib.Path(__file__).resolve().parent


def _binary():
    b = HERE / "cabi" / "abi_check"
    lib = ROOT / "yoxel-voxel_amd" / "lib" / "libyoxel.so"
    if not b.exists() or b.stat().st_mtime < lib.stat().st_mtime:
        import importlib.util
        spec = importlib.util.spec_from_file_location("cabi_build", HERE / "cabi" / "build.py")
        m = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(m)
        m.build()
    return b


def _header_fields():
    """struct name -> field names, from the header's struct bodies."""
    text = (ROOT / "include" / "yoxel.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"typedef struct (\w+) \{(.*?)\} \1;", text, flags=re.S):
        names = []
        for decl in m.group(2).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            # "float pos[3], dir[3]" / "int32_t width, height" / "const void* data"
            parts = [p.strip() for p in decl.split(",")]
            first = re.findall(r"(\w+)\s*(?:\[\d+\])?$", parts[0])
            names.append(first[0])
            for p in parts[1:]:
                names.append(re.findall(r"(\w+)", p)[0])
        out[m.group(1)] = names
    return out


MIRRORS = {"yx_scene_desc": "SceneDesc", "yx_scene_info": "SceneInfo", "yx_sphere_edit": "SphereEdit",
           "yx_volume_source": "VolumeSource", "yx_outputs": "Outputs", "yx_tuning": "Tuning",
           "yx_tune_trial": "TuneTrial", "yx_light_params": "LightParams", "yx_pipeline_outputs": "PipelineOutputs",
           "yx_batch_job": "BatchJob", "yx_volume_desc": "VolumeDesc", "yx_volume_info": "VolumeInfo",
           "yx_volume_dev": "VolumeDev", "yx_ntree_desc": "NTreeDesc", "yx_ntree_info": "NTreeInfo",
           "yx_ntree_outputs": "NTreeOutputs"}


def test_struct_layouts(yx):
    out = subprocess.run([str(_binary()), "layout"], capture_output=True, text=True, check=True).stdout
    compiled = {d["struct"]: d for d in map(json.loads, out.strip().splitlines())}
    header = _header_fields()
    # the program lists every struct and field the header declares
    assert set(compiled) == set(header), set(compiled) ^ set(header)
    for name, fields in header.items():
        assert list(compiled[name]["fields"]) == fields, name
    # VoxNode is the reference's 40-byte record
    assert compiled["yx_vox_node"]["size"] == 40 and compiled["yx_vox_node"]["fields"]["child"] == 8
    # every ctypes mirror has the compiled size and offsets
    assert set(MIRRORS) == set(header) - {"yx_vox_node"}
    for cname, pyname in MIRRORS.items():
        cls = getattr(yx._lib, pyname)
        c = compiled[cname]
        assert ctypes.sizeof(cls) == c["size"], f"{pyname}: {ctypes.sizeof(cls)} != {c['size']}"
        py_fields = [f[0] for f in cls._fields_]
        assert py_fields == list(c["fields"]), f"{pyname} fields {py_fields} != {list(c['fields'])}"
        for f in py_fields:
            assert getattr(cls, f).offset == c["fields"][f], f"{pyname}.{f}"


@pytest.mark.gpu
def test_reference_driver_sequence(yx, orc, cuda, tmp_path):
    print("compiled ISVORenderer: 1024x768 frame from a .vox file", flush=True)
    scene = yx.SVOData.build_synthetic(**yx.workloads.PLAIN_512)
    fn = tmp_path / "scene.vox"
    scene.Save(fn)
    frame = tmp_path / "frame.raw"
    r = subprocess.run([str(_binary()), "render", str(fn), str(frame)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    state = json.loads(r.stdout)
    assert state["null_without_scene"] is True
    assert state["default_resolution"] == [640, 480] and state["resolution"] == [1024, 768]
    got = np.fromfile(frame, dtype=np.uint32).reshape(768, 1024)
    # the same scene read back through yx_scene_load_vox
    loaded = yx.SVOData()
    loaded.Load(fn)
    assert np.array_equal(loaded.nodes(), scene.nodes()) and loaded.GetRoot() == scene.GetRoot()
    ref = orc.render(loaded.nodes(), loaded.GetRoot(), (0.5, 0.5, 0.3), (-1.0, -1.0, -1.5), up=(0, 0, 1),
                     fov=70.0, w=1024, h=768, light_mode=0, threads=8)
    hits = int((ref["hit_id"] != 0xFFFFFFFF).sum())
    assert hits > 1000
    assert np.array_equal(got, ref["rgba"]), f"{int((got != ref['rgba']).sum())} pixels differ"
