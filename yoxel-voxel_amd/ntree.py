"""Brick n-tree volumes: GPU-built 64-ary trees of 8^3 u8 bricks and their
DDA ray cast (yx_ntree_* of include/yoxel.h).

The tree is the n-tree of ntree/ntree/nodes.h:13-100 (4^3-cell grid nodes,
brick leaves) under an R^3 root grid; the cast is castRay / walkBrick of
reaction/trace/trace.cu:42-139 with a 256-entry RGBA8 transfer table
(volvis.py:145-152), empty cells skipped at every tree level and brick voxels
crossed one by one.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .renderer import _addr, _stream


class BrickTree:
    """Device-resident n-tree (yx_ntree)."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.yx_ntree_free(h)
            self._h = ctypes.c_void_p()

    @classmethod
    def build_fbm(cls, size, levels, octaves, frequency, threshold, scale, seed, stream=None):
        d = _lib.NTreeDesc(size, levels, octaves, frequency, threshold, scale, seed)
        h = ctypes.c_void_p()
        check(lib().yx_ntree_build_fbm(ctypes.byref(d), ctypes.byref(h), ctypes.c_void_p(_stream(stream))))
        return cls(h)

    @classmethod
    def from_dense(cls, dense, levels):
        """dense: [z, y, x] u8 array of side size."""
        dense = np.ascontiguousarray(dense, np.uint8)
        h = ctypes.c_void_p()
        check(lib().yx_ntree_from_dense(dense.ctypes.data, dense.shape[0], int(levels), ctypes.byref(h)))
        return cls(h)

    def info(self):
        i = _lib.NTreeInfo()
        check(lib().yx_ntree_get_info(self._h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_}

    def download(self):
        """Host copies (root [R^3] u32, nodes [n, 4] u32, bricks [m, 512] u8)."""
        i = self.info()
        root = np.zeros(i["root_side"] ** 3, np.uint32)
        nodes = np.zeros((i["num_nodes"], 4), np.uint32)
        bricks = np.zeros((i["num_bricks"], 512), np.uint8)
        check(lib().yx_ntree_download(self._h, root.ctypes.data, nodes.ctypes.data if nodes.size else None,
                                      bricks.ctypes.data if bricks.size else None))
        return root, nodes, bricks

    def set_launch(self, persistent=False, refill_min=16):
        """One lane per pixel (default) or persistent waves with ray refill;
        identical outputs."""
        check(lib().yx_ntree_set_launch(self._h, 1 if persistent else 0, int(refill_min)))

    def render(self, view, width, height, palette, stream=None, **out):
        """Cast into caller-owned device buffers (rgba, t, voxel, nodes,
        voxels); `palette` is a device tensor of 256 RGBA8 (uint32/int32)."""
        v = (ctypes.c_float * 12)(*[float(x) for x in np.asarray(view, np.float32)])
        o = _lib.NTreeOutputs(*[_addr(out.get(k)) for k in ("rgba", "t", "voxel", "nodes", "voxels")])
        check(lib().yx_ntree_render(self._h, v, int(width), int(height), ctypes.c_void_p(_addr(palette)),
                                    ctypes.byref(o), ctypes.c_void_p(_stream(stream))))

    def render_tensors(self, view, width, height, palette, outputs=("rgba", "t", "voxel"), stream=None):
        import torch
        dev = torch.device("cuda", torch.cuda.current_device())
        spec = {"rgba": ((height, width), torch.int32), "t": ((height, width), torch.float32),
                "voxel": ((height, width, 3), torch.int32), "nodes": ((height, width), torch.int32),
                "voxels": ((height, width), torch.int32)}
        out = {k: torch.empty(spec[k][0], dtype=spec[k][1], device=dev) for k in outputs}
        self.render(view, width, height, palette, stream=stream, **out)
        return out

    @property
    def handle(self):
        return self._h


__all__ = ["BrickTree"]
