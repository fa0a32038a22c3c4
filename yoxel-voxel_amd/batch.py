"""Batched rendering: many scenes x many cameras in one persistent launch.

BatchRenderer replaces a loop of per-scene SetScene / SetView / Render calls
(demo/Demo.cpp:138-174 drives one SVORenderer per scene and view) with one
launch over every (scene, camera) job; outputs are those of rendering each job
alone with SVORenderer.Render, concatenated in job order.
"""
import ctypes
import hashlib
import json

from . import _lib
from ._lib import check, lib
from .renderer import _addr, _autotune, _stream

_OUT_SPEC = {"hit_id": (1, "int32"), "voxel": (3, "int32"), "t": (1, "float32"), "rgba": (1, "int32"),
             "shadow": (1, "uint8"), "visits": (1, "int32")}


def job_struct(scene, width, height, pos, dir, up=(0.0, 0.0, 1.0), fov=60.0,
               light_mode=_lib.YX_LIGHT_HEADLIGHT, light=(0.0, 0.0, 1.0)):
    j = _lib.BatchJob()
    j.scene = int(scene)
    j.width = int(width)
    j.height = int(height)
    j.pos[:] = [float(v) for v in pos]
    j.dir[:] = [float(v) for v in dir]
    j.up[:] = [float(v) for v in up]
    j.fov = float(fov)
    j.light_mode = int(light_mode)
    j.light[:] = [float(v) for v in light]
    return j


class BatchRenderer:
    """yx_batch: scenes are uploaded once into one device record pool; jobs
    are dicts with the keys of job_struct (scene index, width, height, pos,
    dir, up, fov, light_mode, light)."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        check(lib().yx_batch_create(ctypes.byref(self._h)))
        self.jobs = []
        self.total_pixels = 0
        self._scene_sig = ""

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.yx_batch_destroy(h)
            self._h = ctypes.c_void_p()

    def set_scenes(self, scenes):
        arr = (ctypes.c_void_p * max(1, len(scenes)))(*[s.handle.value for s in scenes])
        check(lib().yx_batch_set_scenes(self._h, arr, len(scenes)))
        self.jobs = []
        self.total_pixels = 0
        self._scene_sig = ";".join(json.dumps(s.info(), sort_keys=True) for s in scenes)

    def set_jobs(self, jobs):
        arr = (_lib.BatchJob * max(1, len(jobs)))(*[job_struct(**j) for j in jobs])
        total = ctypes.c_uint64()
        check(lib().yx_batch_set_jobs(self._h, arr, len(jobs), ctypes.byref(total)))
        self.jobs = [dict(j) for j in jobs]
        self.total_pixels = total.value
        return self.total_pixels

    def job_offset(self, j):
        o = ctypes.c_uint64()
        check(lib().yx_batch_job_offset(self._h, int(j), ctypes.byref(o)))
        return o.value

    def job_slices(self):
        """[(first_pixel, width, height)] per job."""
        return [(self.job_offset(i), j["width"], j["height"]) for i, j in enumerate(self.jobs)]

    def get_tuning(self):
        t = _lib.Tuning()
        check(lib().yx_batch_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = _lib.Tuning()
        check(lib().yx_batch_get_tuning(self._h, ctypes.byref(t)))
        for k, v in kw.items():
            if not hasattr(t, k):
                raise KeyError(k)
            setattr(t, k, int(v))
        check(lib().yx_batch_set_tuning(self._h, ctypes.byref(t)))

    def set_graph(self, use_graph):
        check(lib().yx_batch_set_graph(self._h, 1 if use_graph else 0))

    def set_payload(self, fmt):
        """YX_PAYLOAD_RGBA8 (default) or YX_PAYLOAD_C16N16 for every scene."""
        check(lib().yx_batch_set_payload(self._h, int(fmt)))

    @staticmethod
    def _outputs(out):
        return _lib.Outputs(*[_addr(out.get(k)) for k in ("hit_id", "voxel", "t", "rgba", "shadow", "visits")])

    def render(self, stream=None, **out):
        """One frame of every job into caller-owned device buffers of
        total_pixels entries (voxel: 3 per pixel), asynchronously on `stream`."""
        o = self._outputs(out)
        check(lib().yx_render_batch(self._h, ctypes.byref(o), ctypes.c_void_p(_stream(stream))))

    def alloc_outputs(self, outputs=("hit_id", "voxel", "t", "rgba", "shadow")):
        import torch
        dev = torch.device("cuda", torch.cuda.current_device())
        res = {}
        for k in outputs:
            per, dt = _OUT_SPEC[k]
            shape = (self.total_pixels, 3) if per == 3 else (self.total_pixels,)
            res[k] = torch.empty(shape, dtype=getattr(torch, dt), device=dev)
        return res

    def render_tensors(self, outputs=("hit_id", "voxel", "t", "rgba", "shadow"), stream=None):
        out = self.alloc_outputs(outputs)
        self.render(stream=stream, **out)
        return out

    def signature(self):
        """Hash of the scenes and job layout (what the best tuning depends on)."""
        lay = [(j["scene"], j["width"], j["height"], j.get("light_mode", 0)) for j in self.jobs]
        return hashlib.sha1((self._scene_sig + json.dumps(lay)).encode()).hexdigest()[:16]

    def autotune(self, out, reps=2, cache=None, stream=None):
        """Pick the launch shape for this batch with the library's autotuner
        (yx_batch_autotune). With `cache` (a JSON path) a result stored for
        the same device and batch signature is reused without timing.
        Returns (tuning dict, [(tuning dict, ms)] trials)."""
        return _autotune(self, lib().yx_batch_autotune, out, reps, cache, stream, "batch:" + self.signature())

    def scene_bytes(self):
        b = ctypes.c_uint64()
        check(lib().yx_batch_scene_bytes(self._h, ctypes.byref(b)))
        return b.value

    @property
    def handle(self):
        return self._h


__all__ = ["BatchRenderer", "job_struct"]
