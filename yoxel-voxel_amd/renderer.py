"""Host-side mirror of the reference renderer API over the C ABI.

SVOData      - cell/svodata.h:22-55 (Load, GetRoot, operator[])
SVORenderer  - ISVORenderer (cell/svorenderer.h:5-31): SetScene, SetViewPos,
               SetViewDir, SetViewUp, SetResolution, GetResolution, SetFOV,
               RenderFrame; plus the GPU renderer's Render(dst)
               (demo/SVORenderer.h:8-39) and the light setter, the GPU
               pipeline's settings and RenderPipeline
               (demo/SVORenderer.cpp:95-149), and the incremental scene sync
               UpdateScene / CountChangedPages (CudaSVO::Update).
Error behaviour follows the reference: RenderFrame returns None when no scene
is set (cell/ppu_renderer.cpp:78-79); other failures raise YoxelError.
"""
import ctypes
import json
import os
import pathlib

import numpy as np

from . import _lib
from ._lib import YoxelError, check, f3, lib


class SVOData:
    """Octree in the reference VoxNode layout (10 u32 per node)."""

    def __init__(self):
        self._h = ctypes.c_void_p()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.yx_scene_free(h)
            self._h = ctypes.c_void_p()

    def _replace(self, h):
        if self._h.value:
            lib().yx_scene_free(self._h)
        self._h = h

    # -- constructors ------------------------------------------------------
    def Load(self, fn):
        h = ctypes.c_void_p()
        check(lib().yx_scene_load_vox(str(fn).encode(), ctypes.byref(h)))
        self._replace(h)
        return self

    def Save(self, fn):
        check(lib().yx_scene_save_vox(self._h, str(fn).encode()))

    @classmethod
    def from_nodes(cls, nodes, root):
        a = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 10)
        s = cls()
        h = ctypes.c_void_p()
        check(lib().yx_scene_from_nodes(a.ctypes.data_as(ctypes.c_void_p), a.shape[0], int(root),
                                        ctypes.byref(h)))
        s._h = h
        return s

    @classmethod
    def build_synthetic(cls, depth, num_spheres=0, radius_min=0.0, radius_max=0.0, heightfield=False,
                        hf_octaves=5, hf_amplitude=0.2, hf_base=0.3, seed=0):
        d = _lib.SceneDesc(depth, num_spheres, radius_min, radius_max, 1 if heightfield else 0,
                           hf_octaves, hf_amplitude, hf_base, seed)
        s = cls()
        h = ctypes.c_void_p()
        check(lib().yx_scene_build_synthetic(ctypes.byref(d), ctypes.byref(h)))
        s._h = h
        return s

    @classmethod
    def build_synthetic_gpu(cls, depth, num_spheres=0, radius_min=0.0, radius_max=0.0, heightfield=False,
                            hf_octaves=5, hf_amplitude=0.2, hf_base=0.3, seed=0, stream=None):
        """The same scene as build_synthetic (identical nodes), built on the GPU."""
        d = _lib.SceneDesc(depth, num_spheres, radius_min, radius_max, 1 if heightfield else 0,
                           hf_octaves, hf_amplitude, hf_base, seed)
        s = cls()
        h = ctypes.c_void_p()
        check(lib().yx_scene_build_synthetic_gpu(ctypes.byref(d), ctypes.byref(h),
                                                 ctypes.c_void_p(_stream(stream))))
        s._h = h
        return s

    # -- edits (the node pool of DynamicSVO, reaction/report/main.tex:67-71) --
    def write_nodes(self, first, nodes):
        """Write VoxNode records (n x 10 u32) from index `first` (at most the
        current size); their 256-node pages get a new version."""
        a = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 10)
        check(lib().yx_scene_write_nodes(self._h, int(first), a.ctypes.data_as(ctypes.c_void_p), a.shape[0]))

    def set_root(self, root):
        check(lib().yx_scene_set_root(self._h, int(root)))

    GROW = _lib.YX_BUILD_GROW
    CLEAR = _lib.YX_BUILD_CLEAR
    IN_PLACE = _lib.YX_RANGE_IN_PLACE
    REBUILD = _lib.YX_RANGE_REBUILD
    REBUILD_GPU = _lib.YX_RANGE_REBUILD_GPU

    def BuildRange(self, level, pos, mode, radius, color=0xFFC0B6C0, how=IN_PLACE, stream=None):
        """DynamicSVO::BuildRange(level, pos, mode, SphereSource(radius, color))
        (ore/src/main.cpp:121-127, demo/Demo.cpp:82-114) on a synthetic scene:
        GROW unions, CLEAR subtracts a ball of `radius` voxels of `level` at
        `pos` (voxels of `level`). `how`: IN_PLACE edits the node array keeping
        untouched nodes' ids, REBUILD / REBUILD_GPU rebuild the canonical tree
        on the host / the GPU."""
        self.BuildRanges([(mode, level, pos, radius, color)], how, stream)

    def BuildRanges(self, edits, how=IN_PLACE, stream=None):
        """Several edits [(mode, level, pos, radius, color)] in one call, applied
        in order (Demo::DoEdit makes 100 per frame, demo/Demo.cpp:82-114)."""
        arr = (_lib.SphereEdit * max(1, len(edits)))()
        for k, (mode, level, pos, radius, color) in enumerate(edits):
            arr[k] = _lib.SphereEdit(int(mode), int(level), (ctypes.c_float * 3)(*pos), float(radius), int(color))
        st = ctypes.c_void_p(_stream(stream) if how == _lib.YX_RANGE_REBUILD_GPU else 0)
        check(lib().yx_scene_build_range(self._h, arr, len(edits), int(how), st))

    def fill_interior(self):
        """Mark the empty space enclosed by a surface-only tree as FullNode (see
        yx_scene_fill_interior); returns how many empty cubes became solid."""
        k = ctypes.c_uint64()
        check(lib().yx_scene_fill_interior(self._h, ctypes.byref(k)))
        return k.value

    def BuildRangeIso(self, level, pos, mode, values, iso, inside=True, color=0xFFC0B6C0, how=IN_PLACE,
                      stream=None):
        """BuildRange with an IsoSource (scene_gen.py:111-130): `values` uint8
        [z][y][x]; solid where value >= iso (inside) or < iso; voxel (0, 0, 0)
        lands at integer `pos` of `level`."""
        v = np.ascontiguousarray(values, dtype=np.uint8)
        self._volume(level, pos, mode, v, _lib.YX_SOURCE_ISO, int(iso), 1 if inside else 0, color, how, stream)

    def BuildRangeRaw(self, level, pos, mode, rgba, how=IN_PLACE, stream=None):
        """BuildRange with a RawSource (scene_gen.py:74-82): `rgba` uint32 [z][y][x]
        (or uint8 [z][y][x][4]) RGBA8; solid where alpha != 0."""
        v = np.ascontiguousarray(rgba)
        if v.dtype == np.uint8 and v.ndim == 4 and v.shape[3] == 4:
            v = np.ascontiguousarray(v).view(np.uint32)[..., 0]
        v = np.ascontiguousarray(v, dtype=np.uint32)
        self._volume(level, pos, mode, v, _lib.YX_SOURCE_RAW, 0, 1, 0, how, stream)

    def _volume(self, level, pos, mode, v, kind, iso, inside, color, how, stream):
        if v.ndim != 3:
            raise ValueError("a source volume is [z][y][x]")
        src = _lib.VolumeSource(int(mode), int(level), (ctypes.c_int32 * 3)(*[int(p) for p in pos]),
                                (ctypes.c_int32 * 3)(v.shape[2], v.shape[1], v.shape[0]), int(kind), int(iso),
                                int(inside), int(color), ctypes.c_void_p(v.ctypes.data))
        st = ctypes.c_void_p(_stream(stream) if how == _lib.YX_RANGE_REBUILD_GPU else 0)
        check(lib().yx_scene_build_range_volume(self._h, ctypes.byref(src), int(how), st))

    def decode_hits(self, hit_id, payload=_lib.YX_PAYLOAD_RGBA8):
        """Colour (uint32 RGBA8) and normal (float32 x 3) of the leaf each
        hit_id names, zeros for misses (the decoding of
        SVORenderer::DumpTraceData, demo/SVORenderer.cpp:158-192)."""
        ids = np.ascontiguousarray(np.asarray(hit_id).reshape(-1), dtype=np.uint32)
        color = np.zeros(ids.shape[0], dtype=np.uint32)
        normal = np.zeros((ids.shape[0], 3), dtype=np.float32)
        check(lib().yx_scene_decode_hits(self._h, ids.ctypes.data_as(ctypes.c_void_p), ids.shape[0], int(payload),
                                         color.ctypes.data_as(ctypes.c_void_p),
                                         normal.ctypes.data_as(ctypes.c_void_p)))
        return color, normal

    def info(self):
        i = _lib.SceneInfo()
        check(lib().yx_scene_get_info(self._h, ctypes.byref(i)))
        return {"root": i.root, "num_nodes": i.num_nodes, "num_leaves": i.num_leaves,
                "depth": i.depth}

    def GetRoot(self):
        return self.info()["root"]

    def nodes(self):
        """Copy of the node array, shape (num_nodes, 10) uint32."""
        p = ctypes.c_void_p()
        n = ctypes.c_uint32()
        r = ctypes.c_uint32()
        check(lib().yx_scene_nodes(self._h, ctypes.byref(p), ctypes.byref(n), ctypes.byref(r)))
        if n.value == 0:
            return np.zeros((0, 10), np.uint32)
        buf = (ctypes.c_uint32 * (n.value * 10)).from_address(p.value)
        return np.frombuffer(buf, dtype=np.uint32).reshape(n.value, 10).copy()

    def __getitem__(self, node_id):
        return self.nodes()[node_id]

    @property
    def handle(self):
        return self._h


class SVORenderer:
    """ISVORenderer / SVORenderer over libyoxel."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        check(lib().yx_renderer_create(ctypes.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.yx_renderer_destroy(h)
            self._h = ctypes.c_void_p()

    def SetScene(self, svo):
        """Upload `svo` and lay it out on the GPU (waits for it)."""
        check(lib().yx_renderer_set_scene(self._h, svo.handle))
        self._svo = svo

    def SetSceneHost(self, svo):
        """The same layout computed on the host (bit-identical, slower)."""
        check(lib().yx_renderer_set_scene_host(self._h, svo.handle))
        self._svo = svo

    def SetViewPos(self, pos):
        check(lib().yx_renderer_set_view_pos(self._h, f3(pos)))

    def SetViewDir(self, d):
        check(lib().yx_renderer_set_view_dir(self._h, f3(d)))

    def SetViewUp(self, up):
        check(lib().yx_renderer_set_view_up(self._h, f3(up)))

    def SetResolution(self, width, height):
        check(lib().yx_renderer_set_resolution(self._h, int(width), int(height)))

    SetViewSize = SetResolution  # demo/SVORenderer.h:20

    def GetResolution(self):
        w = ctypes.c_int32()
        h = ctypes.c_int32()
        check(lib().yx_renderer_get_resolution(self._h, ctypes.byref(w), ctypes.byref(h)))
        return (w.value, h.value)

    def SetFOV(self, fov):
        check(lib().yx_renderer_set_fov(self._h, float(fov)))

    def SetLight(self, direction, mode=_lib.YX_LIGHT_SUN_SHADOW):
        check(lib().yx_renderer_set_light(self._h, f3(direction), int(mode)))

    def SetPayloadFormat(self, fmt):
        """YX_PAYLOAD_RGBA8 (default) or YX_PAYLOAD_C16N16 (the reference's
        colour16 + normal16 VoxData, layout in include/yoxel.h)."""
        check(lib().yx_renderer_set_payload(self._h, int(fmt)))
        self._payload = int(fmt)

    def DumpTraceData(self, fnbase):
        """SVORenderer::DumpTraceData (demo/SVORenderer.cpp:158-192): trace
        the current view and write `{fnbase}_{w}x{h}.dist` (float32 per pixel,
        zeros as in the reference), `.color` (RGBA8 per pixel) and `.normal`
        (3 float32 per pixel) of the hit leaves, zeros for misses. The frame is
        the one the current camera renders (rendering is deterministic), so
        it equals the last frame when the camera has not moved since."""
        svo = getattr(self, "_svo", None)
        if svo is None:
            raise YoxelError(_lib.YX_ERR_NO_SCENE, "DumpTraceData needs a scene set with SetScene")
        import torch
        hit = self.render_tensors(("hit_id",))["hit_id"]
        torch.cuda.current_stream().synchronize()
        ids = hit.cpu().numpy().view(np.uint32)
        color, normal = svo.decode_hits(ids, getattr(self, "_payload", _lib.YX_PAYLOAD_RGBA8))
        w, h = self.GetResolution()
        base = f"{fnbase}_{w}x{h}"
        np.zeros(w * h, dtype=np.float32).tofile(base + ".dist")
        color.tofile(base + ".color")
        normal.tofile(base + ".normal")
        return base

    def get_tuning(self):
        t = _lib.Tuning()
        check(lib().yx_renderer_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        """Update launch-shape fields (see yx_tuning in include/yoxel.h)."""
        t = _lib.Tuning()
        check(lib().yx_renderer_get_tuning(self._h, ctypes.byref(t)))
        for k, v in kw.items():
            if not hasattr(t, k):
                raise KeyError(k)
            setattr(t, k, int(v))
        check(lib().yx_renderer_set_tuning(self._h, ctypes.byref(t)))

    def ray_basis(self):
        b = (ctypes.c_float * 9)()
        check(lib().yx_renderer_ray_basis(self._h, b))
        return np.array(b[:], dtype=np.float32)

    def RenderFrame(self):
        """Shaded RGBA8 frame as (h, w) uint32, or None without a scene."""
        p = ctypes.c_void_p()
        rc = lib().yx_render_frame(self._h, ctypes.byref(p))
        if rc == _lib.YX_ERR_NO_SCENE:
            return None
        check(rc)
        w, h = self.GetResolution()
        buf = (ctypes.c_uint32 * (w * h)).from_address(p.value)
        return np.frombuffer(buf, dtype=np.uint32).reshape(h, w).copy()

    def Render(self, hit_id=None, voxel=None, t=None, rgba=None, shadow=None, visits=None, stream=None):
        """Render into caller-owned device buffers (torch CUDA tensors or raw
        device addresses), asynchronously on `stream` (a torch stream, a raw
        hipStream_t address, or None for the current torch stream)."""
        o = _lib.Outputs(_addr(hit_id), _addr(voxel), _addr(t), _addr(rgba), _addr(shadow),
                         _addr(visits))
        check(lib().yx_render(self._h, ctypes.byref(o), ctypes.c_void_p(_stream(stream))))

    def render_tensors(self, outputs=("hit_id", "voxel", "t", "rgba", "shadow"), stream=None):
        """Allocate torch output tensors on the current device, render, return them."""
        import torch
        w, h = self.GetResolution()
        dev = torch.device("cuda", torch.cuda.current_device())
        spec = {"hit_id": ((h, w), torch.int32), "voxel": ((h, w, 3), torch.int32),
                "t": ((h, w), torch.float32), "rgba": ((h, w), torch.int32),
                "shadow": ((h, w), torch.uint8), "visits": ((h, w), torch.int32)}
        out = {k: torch.empty(spec[k][0], dtype=spec[k][1], device=dev) for k in outputs}
        self.Render(stream=stream, **out)
        return out

    # ---- incremental sync (CudaSVO::Update, qtview.py:69-73) ----------------

    def UpdateScene(self, svo, stream=None):
        """Upload the pages of `svo` edited since the last sync (all of them for
        another scene) and rebuild the traversal layout on the GPU."""
        check(lib().yx_renderer_update_scene(self._h, svo.handle, ctypes.c_void_p(_stream(stream))))
        self._svo = svo

    def CountChangedPages(self, svo):
        """(pages, bytes) the next UpdateScene would upload
        (DynamicSVO::CountChangedPages / CountTransfrerSize)."""
        p = ctypes.c_uint32()
        b = ctypes.c_uint64()
        check(lib().yx_renderer_count_changed_pages(self._h, svo.handle, ctypes.byref(p), ctypes.byref(b)))
        return p.value, b.value

    def last_update(self):
        p = ctypes.c_uint32()
        b = ctypes.c_uint64()
        check(lib().yx_renderer_last_update(self._h, ctypes.byref(p), ctypes.byref(b)))
        return p.value, b.value

    def layout(self):
        """The device traversal layout: (records (n, 2) uint32, LOD colours (n,) uint32)."""
        n = ctypes.c_uint64()
        check(lib().yx_renderer_layout(self._h, None, None, 0, ctypes.byref(n)))
        recs = np.zeros((n.value, 2), np.uint32)
        lod = np.zeros(n.value, np.uint32)
        if n.value:
            check(lib().yx_renderer_layout(self._h, recs.ctypes.data_as(ctypes.c_void_p),
                                           lod.ctypes.data_as(ctypes.c_void_p), n.value, ctypes.byref(n)))
        return recs, lod

    # ---- the GPU pipeline of SVORenderer::Render (demo/SVORenderer.cpp:95-149)

    def SetDetailCoef(self, coef):
        check(lib().yx_renderer_set_detail_coef(self._h, float(coef)))

    def GetDetailCoef(self):
        c = ctypes.c_float()
        check(lib().yx_renderer_get_detail_coef(self._h, ctypes.byref(c)))
        return c.value

    def SetSSNA(self, enable):
        check(lib().yx_renderer_set_ssna(self._h, 1 if enable else 0))

    def GetSSNA(self):
        e = ctypes.c_int32()
        check(lib().yx_renderer_get_ssna(self._h, ctypes.byref(e)))
        return bool(e.value)

    def SetShowNormals(self, enable):
        check(lib().yx_renderer_set_show_normals(self._h, 1 if enable else 0))

    def SetLigth(self, i, pos, diffuse=(0.7, 0.7, 0.7), specular=(0.3, 0.3, 0.3), attenuation=(1.0, 0.0, 0.5),
                 enabled=True):
        """SVORenderer::SetLigth(i, LightParams) (demo/SVORenderer.h:34), the
        reference's spelling; defaults are the demo's eye light
        (demo/Demo.cpp:141-146)."""
        lp = _lib.LightParams(1 if enabled else 0, (ctypes.c_float * 3)(*pos), (ctypes.c_float * 3)(*diffuse),
                              (ctypes.c_float * 3)(*specular), (ctypes.c_float * 3)(*attenuation))
        check(lib().yx_renderer_set_point_light(self._h, int(i), ctypes.byref(lp)))

    SetPointLight = SetLigth

    def RenderPipeline(self, rgba=None, z=None, t=None, albedo=None, level=None, visits=None, stream=None):
        """SVORenderer::Render(void*): LOD trace, SSNA depth blur, Phong shade,
        into caller-owned device buffers, asynchronously on `stream`."""
        o = _lib.PipelineOutputs(_addr(rgba), _addr(z), _addr(t), _addr(albedo), _addr(level), _addr(visits))
        check(lib().yx_render_pipeline(self._h, ctypes.byref(o), ctypes.c_void_p(_stream(stream))))

    def pipeline_tensors(self, outputs=("rgba", "z", "t", "albedo", "level"), stream=None):
        """Allocate torch output tensors, run the pipeline, return them."""
        import torch
        w, h = self.GetResolution()
        dev = torch.device("cuda", torch.cuda.current_device())
        spec = {"rgba": torch.int32, "z": torch.float32, "t": torch.float32, "albedo": torch.int32,
                "level": torch.uint8, "visits": torch.int32}
        out = {k: torch.empty((h, w), dtype=spec[k], device=dev) for k in outputs}
        self.RenderPipeline(stream=stream, **out)
        return out

    def scene_bytes(self):
        b = ctypes.c_uint64()
        check(lib().yx_renderer_scene_bytes(self._h, ctypes.byref(b)))
        return b.value

    def autotune(self, out, reps=3, cache=None, signature="", stream=None):
        """Pick the launch shape for the current scene and camera with the
        library's autotuner (yx_renderer_autotune); `out` is a dict of output
        tensors as for Render. With `cache` (a JSON path) a result stored for
        the same device and `signature` is reused without timing. Returns
        (tuning dict, [(tuning dict, ms)] trials)."""
        w, h = self.GetResolution()
        return _autotune(self, lib().yx_renderer_autotune, out, reps, cache, stream,
                         f"renderer:{w}x{h}:{signature}")

    @staticmethod
    def _outputs(out):
        return _lib.Outputs(*[_addr(out.get(k)) for k in ("hit_id", "voxel", "t", "rgba", "shadow", "visits")])

    @property
    def handle(self):
        return self._h


def _device_key():
    import torch
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{p.name}/{getattr(p, 'gcnArchName', '')}/{p.multi_processor_count}"


def _autotune(obj, fn, out, reps, cache, stream, sig):
    key = f"{_device_key()}|{sig}"
    if cache is not None:
        path = pathlib.Path(cache)
        if path.exists():
            stored = json.loads(path.read_text()).get(key)
            if stored is not None:
                obj.set_tuning(**stored)
                return obj.get_tuning(), []
    cap = 256
    trials = (_lib.TuneTrial * cap)()
    n = ctypes.c_int32()
    o = obj._outputs(out) if isinstance(out, dict) else out
    check(fn(obj.handle, ctypes.byref(o), int(reps), trials, cap, ctypes.byref(n),
             ctypes.c_void_p(_stream(stream))))
    best = obj.get_tuning()
    log = [(trials[i].tuning.as_dict(), float(trials[i].ms)) for i in range(min(n.value, cap))]
    if cache is not None:
        path = pathlib.Path(cache)
        data = json.loads(path.read_text()) if path.exists() else {}
        data[key] = best
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(path.suffix + f".{os.getpid()}.tmp")
        tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
        os.replace(tmp, path)
    return best, log


def _addr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _stream(s):
    if s is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(s, int):
        return s
    return s.cuda_stream


__all__ = ["SVOData", "SVORenderer", "YoxelError"]
