"""Differentiable front-to-back compositing over brick voxels.

Host side of the yx_volume / yx_compositor C ABI (include/yoxel.h) and a
torch.autograd.Function over it. The forward model is castRay / walkBrick of
reaction/trace/trace.cu:42-139 (brick grid with occupancy table, samples on the
lattice t = k * dt with dt one voxel, front-to-back accumulation), with
sigma = softplus(raw) and per-voxel RGB as the learnable parameters.
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import check, f3, lib

# differentiable compositing workload: 512^3 grid as 8^3 bricks, ~10 % of the
# bricks allocated from an fBm threshold, 16 views of 512x512 rays
COMPOSITE_512 = dict(grid=512, brick=8, occupancy=0.1, octaves=5, frequency=4.0, seed=7)


class Volume:
    """Synthetic brick volume: host arrays (numpy) built by the library."""

    def __init__(self, grid, brick, occupancy=0.1, octaves=5, frequency=4.0, seed=0):
        d = _lib.VolumeDesc(grid, brick, occupancy, octaves, frequency, seed)
        h = ctypes.c_void_p()
        check(lib().yx_volume_build_synthetic(ctypes.byref(d), ctypes.byref(h)))
        try:
            info = _lib.VolumeInfo()
            check(lib().yx_volume_get_info(h, ctypes.byref(info)))
            t = ctypes.c_void_p()
            p = ctypes.c_void_p()
            check(lib().yx_volume_arrays(h, ctypes.byref(t), ctypes.byref(p)))
            g = grid // brick
            self.grid, self.brick, self.num_slots = info.grid, info.brick, info.num_slots
            self.table = np.frombuffer((ctypes.c_int32 * (g ** 3)).from_address(t.value), np.int32).copy()
            nvals = self.num_slots * brick ** 3 * 4
            self.params = (np.frombuffer((ctypes.c_float * nvals).from_address(p.value), np.float32).copy()
                           if nvals else np.zeros(0, np.float32)).reshape(-1, 4)
        finally:
            lib().yx_volume_free(h)

    def to(self, device):
        import torch
        return (torch.from_numpy(self.table).to(device), torch.from_numpy(self.params).to(device))


def camera_views(cams, fov, w, h):
    """[(pos, dir)] -> float32 array [n, 12] (pos + InitRayDir basis)."""
    out = np.zeros((len(cams), 12), np.float32)
    for i, (pos, d) in enumerate(cams):
        v = (ctypes.c_float * 12)()
        check(lib().yx_camera_basis(f3(pos), f3(d), f3((0, 0, 1)), float(fov), int(w), int(h), v))
        out[i] = np.array(v[:], np.float32)
    return out


def _check(name, x, dtype, shape_tail=None):
    """The C ABI takes raw device pointers: refuse tensors whose bytes would be
    misread (wrong dtype, host memory, strided views)."""
    import torch
    if not isinstance(x, torch.Tensor) or x.dtype != dtype or not x.is_cuda or not x.is_contiguous():
        got = (f"{x.dtype} on {x.device}, contiguous={x.is_contiguous()}" if isinstance(x, torch.Tensor)
               else type(x).__name__)
        raise TypeError(f"{name} must be a contiguous CUDA tensor of {dtype}, got {got}")
    if shape_tail is not None and tuple(x.shape[1:]) != tuple(shape_tail):
        raise ValueError(f"{name} must have shape [n, {', '.join(map(str, shape_tail))}], got {tuple(x.shape)}")


def _vol_dev(table, params, grid, brick):
    import torch
    _check("params", params, torch.float32, (4,))
    _check("table", table, torch.int32)
    n_slots = params.shape[0] // (brick ** 3)
    return _lib.VolumeDev(table.data_ptr(), params.data_ptr(), grid, brick, n_slots)


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Compositor:
    """Owns the backward's device workspace."""

    def __init__(self):
        self._h = ctypes.c_void_p()
        check(lib().yx_compositor_create(ctypes.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.yx_compositor_destroy(h)
            self._h = ctypes.c_void_p()

    def forward(self, table, params, grid, brick, views, w, h):
        """params: [slots * brick^3, 4] float32 CUDA; views: [n, 12] float32 CUDA.
        Returns rgb [n, h, w, 3], transmittance [n, h, w], samples [n, h, w]."""
        import torch
        _check("views", views, torch.float32, (12,))
        n = views.shape[0]
        dev = params.device
        rgb = torch.empty((n, h, w, 3), dtype=torch.float32, device=dev)
        trans = torch.empty((n, h, w), dtype=torch.float32, device=dev)
        nsamp = torch.empty((n, h, w), dtype=torch.int32, device=dev)
        vol = _vol_dev(table, params, grid, brick)
        check(lib().yx_composite_forward(self._h, ctypes.byref(vol), views.data_ptr(), n, w, h, rgb.data_ptr(),
                                         trans.data_ptr(), nsamp.data_ptr(), _stream()))
        return rgb, trans, nsamp

    def backward(self, table, params, grid, brick, views, w, h, rgb, nsamp, grad_rgb):
        import torch
        _check("views", views, torch.float32, (12,))
        _check("rgb", rgb, torch.float32)
        _check("nsamp", nsamp, torch.int32)
        n = views.shape[0]
        vol = _vol_dev(table, params, grid, brick)
        grad = torch.empty(params.shape, dtype=params.dtype, device=params.device)
        g = grad_rgb.to(torch.float32).contiguous()
        if g.shape != rgb.shape:
            raise ValueError(f"grad_rgb shape {tuple(g.shape)} differs from rgb {tuple(rgb.shape)}")
        check(lib().yx_composite_backward(self._h, ctypes.byref(vol), views.data_ptr(), n, w, h, rgb.data_ptr(),
                                          nsamp.data_ptr(), g.data_ptr(), grad.data_ptr(), _stream()))
        return grad

    def reserve(self, records):
        """Workspace capacity in records, so that no backward waits to size it."""
        check(lib().yx_compositor_reserve(self._h, int(records)))

    def last_stats(self):
        r = ctypes.c_uint64()
        u = ctypes.c_uint64()
        check(lib().yx_compositor_last_stats(self._h, ctypes.byref(r), ctypes.byref(u)))
        return r.value, u.value


def composite_function():
    """torch.autograd.Function: rgb = composite(params) for fixed views/table."""
    import torch

    class VolumeComposite(torch.autograd.Function):
        @staticmethod
        def forward(ctx, params, table, views, grid, brick, w, h, comp):
            rgb, trans, nsamp = comp.forward(table, params, grid, brick, views, w, h)
            ctx.save_for_backward(params, table, views, rgb, nsamp)
            ctx.meta = (grid, brick, w, h, comp)
            return rgb

        @staticmethod
        def backward(ctx, grad_rgb):
            params, table, views, rgb, nsamp = ctx.saved_tensors
            grid, brick, w, h, comp = ctx.meta
            grad = comp.backward(table, params, grid, brick, views, w, h, rgb, nsamp, grad_rgb)
            return grad, None, None, None, None, None, None, None

    return VolumeComposite


def orbit_views(n, fov, w, h, seed=0):
    from .workloads import orbit_cameras
    return camera_views(orbit_cameras(n, seed=seed), fov, w, h)


__all__ = ["Volume", "Compositor", "composite_function", "camera_views", "orbit_views", "COMPOSITE_512",
           "BrickPool"]


class BrickPool:
    """CuBrickPoolManager (reaction/voxtree/brickpool/BrickPool.cpp:21-62) over
    caller-owned torch tensors: `table` int32 [(grid/brick)^3] and `params`
    float32 [slots * brick^3, 4] on the GPU. alloc / release / update change
    them in place (see yx_brick_pool_* in include/yoxel.h)."""

    def __init__(self, table, params, grid, brick):
        self.table, self.params, self.grid, self.brick = table, params, grid, brick
        self._dev = _vol_dev(table, params, grid, brick)
        self._h = ctypes.c_void_p()
        check(lib().yx_brick_pool_create(ctypes.byref(self._dev), _stream(), ctypes.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.yx_brick_pool_destroy(h)
            self._h = ctypes.c_void_p()

    def alloc(self, bricks):
        """Slots for the listed bricks (int32 device tensor of linear brick
        indices); returns (slot of every listed brick, number of new slots)."""
        import torch
        b = bricks.to(torch.int32).contiguous()
        slots = torch.empty_like(b)
        n = ctypes.c_uint32()
        check(lib().yx_brick_pool_alloc(self._h, ctypes.c_void_p(b.data_ptr()), b.numel(),
                                        ctypes.c_void_p(slots.data_ptr()), ctypes.byref(n), _stream()))
        return slots, n.value

    def release(self, bricks):
        import torch
        b = bricks.to(torch.int32).contiguous()
        n = ctypes.c_uint32()
        check(lib().yx_brick_pool_release(self._h, ctypes.c_void_p(b.data_ptr()), b.numel(), ctypes.byref(n),
                                          _stream()))
        return n.value

    def update(self, bricks, packed):
        """UpdateBrickPool: packed float32 [n * brick^3, 4] into the listed bricks' slots."""
        import torch
        b = bricks.to(torch.int32).contiguous()
        pk = packed.to(torch.float32).contiguous()
        if pk.numel() != b.numel() * self.brick ** 3 * 4:
            raise ValueError("packed must hold brick^3 float4 per listed brick")
        check(lib().yx_brick_pool_update(self._h, ctypes.c_void_p(b.data_ptr()), b.numel(),
                                         ctypes.c_void_p(pk.data_ptr()), _stream()))

    def stats(self):
        u = ctypes.c_uint32()
        c = ctypes.c_uint32()
        check(lib().yx_brick_pool_stats(self._h, ctypes.byref(u), ctypes.byref(c)))
        return u.value, c.value
