"""Brick classification of a bit volume on the GPU: the voxtree pipeline of
reaction/voxtree/voxtree.py:80-95 (CalcDensity, MarkBricks, the column-start
cumsum, PackBricks of reaction/voxtree/density.cu) over the yx_voxtree_* C ABI
(include/yoxel.h). Arrays are torch tensors on the GPU; formats as in the
header: bits [B/128][B][B][4] uint32, density [V][V][V/8] uint32 of 4-bit
counts, brick_state [G][G][G], column counts / starts [G][G], packed
x | y << 8 | z << 16 per mixed brick.
"""
import ctypes

from ._lib import check, lib

SOLID = 0x80000000
EMPTY = 0x40000000


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _i32(t):
    import torch
    return t if t.dtype == torch.int32 else t.view(torch.int32)


def pack_bits(vol):
    """bool tensor [x, y, z] of side B (a power of two, 128..1024) -> int32
    words [B/128][B (y)][B (x)][4]: bit b of word l of slice s is voxel
    z = 128 s + 32 l + b (the voxelizer's dump layout, voxtree.py:76-82)."""
    import torch
    B = vol.shape[0]
    v = vol.reshape(B, B, B // 128, 4, 32).permute(2, 1, 0, 3, 4)
    w = torch.zeros(v.shape[:-1], dtype=torch.int64, device=vol.device)
    for b in range(32):
        w |= v[..., b].to(torch.int64) << b
    w = torch.where(w >= 2 ** 31, w - 2 ** 32, w)
    return w.to(torch.int32).reshape(-1).contiguous()


def calc_density(bits, bit_size, out=None):
    """CalcDensity (density.cu:32-60): 2x2x2 counts of the bit volume."""
    import torch
    V = bit_size // 2
    dens = out if out is not None else torch.empty(V * V * V // 8, dtype=torch.int32, device=bits.device)
    check(lib().yx_voxtree_calc_density(_ptr(_i32(bits)), int(bit_size), _ptr(dens), _stream()))
    return dens


def mark_bricks(density, vol_size, out=None):
    """MarkBricks (density.cu:131-176): (brick_state, column_count)."""
    import torch
    G = vol_size // 4
    if out is not None:
        state, cnt = out
    else:
        state = torch.empty(G ** 3, dtype=torch.int32, device=density.device)
        cnt = torch.empty(G * G, dtype=torch.int32, device=density.device)
    check(lib().yx_voxtree_mark_bricks(_ptr(_i32(density)), int(vol_size), _ptr(state), _ptr(cnt), _stream()))
    return state, cnt


def pack_bricks(brick_state, column_start, vol_size, capacity, out=None):
    """PackBricks (density.cu:210-221) into a list of `capacity` entries."""
    import torch
    packed = out if out is not None else torch.zeros(max(1, capacity), dtype=torch.int32, device=brick_state.device)
    check(lib().yx_voxtree_pack_bricks(_ptr(_i32(brick_state)), _ptr(_i32(column_start)), int(vol_size),
                                       _ptr(packed), int(capacity), _stream()))
    return packed[:capacity]


def classify(bits, bit_size, capacity=None):
    """The whole pipeline with the column scan on the GPU. Returns a dict of
    density, brick_state, column_count, column_start (G*G + 1 entries, the
    last the total), packed (the mixed bricks) and total."""
    import torch
    V = bit_size // 2
    G = V // 4
    dev = bits.device
    cap = G ** 3 if capacity is None else int(capacity)
    dens = torch.empty(V * V * V // 8, dtype=torch.int32, device=dev)
    state = torch.empty(G ** 3, dtype=torch.int32, device=dev)
    cnt = torch.empty(G * G, dtype=torch.int32, device=dev)
    start = torch.empty(G * G + 1, dtype=torch.int32, device=dev)
    packed = torch.zeros(max(1, cap), dtype=torch.int32, device=dev)
    total = ctypes.c_uint32()
    check(lib().yx_voxtree_classify(_ptr(_i32(bits)), int(bit_size), _ptr(dens), _ptr(state), _ptr(cnt),
                                    _ptr(start), _ptr(packed), cap, ctypes.byref(total), _stream()))
    return dict(density=dens, brick_state=state, column_count=cnt, column_start=start,
                packed=packed[:total.value], total=total.value)
