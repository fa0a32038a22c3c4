"""The device primitives (csrc/prims.hip) against numpy: the stable LSD radix
sort (u32 and u64 keys, any bit range, carried ids or positions, a count held
on the device), the exclusive scans, the reduction and select-distinct. Every
result is compared bit for bit. These primitives stand in for the library
sort / scan / unique calls of the reference's GPU builders; the stable sort's
order is numpy's stable argsort of the sorted bit field."""
import ctypes

import numpy as np
import pytest

# sizes around the 64-lane wave, the 4096-key sort tile, the 8192-item
# single-block scan and several sort blocks per compute unit
SIZES = [1, 63, 64, 4095, 4096, 4097, 8191, 8192, 8193, 100_003, 3_000_017]


def _dev(torch, a):
    """numpy array -> CUDA tensor holding the same bytes."""
    view = {np.dtype(np.uint32): np.int32, np.dtype(np.uint64): np.int64, np.dtype(np.uint8): np.uint8}
    return torch.from_numpy(np.ascontiguousarray(a).view(view[a.dtype])).cuda()


def _host(t, dtype):
    return t.cpu().numpy().view(dtype)


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _keys(rng, kind, n, dtype):
    bits = 8 * np.dtype(dtype).itemsize
    hi = np.iinfo(dtype).max
    if kind == "uniform":
        return rng.integers(0, hi, size=n, dtype=dtype, endpoint=True)
    if kind == "few":  # many equal keys: only stability orders them
        return (rng.integers(0, 7, size=n, dtype=dtype) * dtype(0x01010101 if bits == 32 else 0x0101010101010101))
    if kind == "runs":  # keys that arrive in runs, as records emitted tile by tile do
        vals = rng.integers(0, hi, size=max(1, n), dtype=dtype, endpoint=True)
        return np.repeat(vals, rng.integers(1, 10, size=vals.size))[:n]  # every repeat >= 1: n keys
    if kind == "sorted":
        return np.sort(rng.integers(0, hi, size=n, dtype=dtype, endpoint=True))
    if kind == "reversed":
        return np.sort(rng.integers(0, hi, size=n, dtype=dtype, endpoint=True))[::-1].copy()
    return np.full(n, dtype(12345), dtype=dtype)  # "equal"


def _expect(keys, b0, b1):
    dt = keys.dtype.type
    width = b1 - b0
    field = keys >> dt(b0)
    if width < 8 * keys.dtype.itemsize:
        field = field & dt((1 << width) - 1)
    order = np.argsort(field, kind="stable")
    return keys[order], order.astype(np.uint32)


def _sort(yx, torch, keys, b0, b1, ids=None, n_dev=None, with_ids=True):
    lib = yx.lib()
    n = keys.size
    fn = lib.yx_prim_sort_u32 if keys.dtype == np.uint32 else lib.yx_prim_sort_u64
    dk = _dev(torch, keys)
    di = _dev(torch, ids) if ids is not None else None
    ko = torch.empty_like(dk)
    io = torch.empty(n, dtype=torch.int32, device="cuda") if with_ids else None
    nd = _dev(torch, np.array([n_dev], np.uint32)) if n_dev is not None else None
    yx._lib.check(fn(dk.data_ptr(), di.data_ptr() if di is not None else None, n,
                nd.data_ptr() if nd is not None else None, b0, b1, ko.data_ptr(),
                io.data_ptr() if io is not None else None, _stream(torch)))
    return _host(ko, keys.dtype), (_host(io, np.uint32) if io is not None else None)


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES)
def test_sort_u32_sizes(yx, cuda, n):
    rng = np.random.default_rng(n)
    keys = _keys(rng, "uniform", n, np.uint32)
    for b0, b1 in [(0, 32), (9, 25)]:
        k, i = _sort(yx, cuda, keys, b0, b1)
        ek, ei = _expect(keys, b0, b1)
        assert np.array_equal(k, ek) and np.array_equal(i, ei), (n, b0, b1)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["few", "runs", "sorted", "reversed", "equal"])
def test_sort_u32_distributions(yx, cuda, kind):
    rng = np.random.default_rng(7)
    n = 250_001
    keys = _keys(rng, kind, n, np.uint32)
    for b0, b1 in [(0, 32), (0, 8), (3, 17), (24, 32), (30, 31)]:
        k, i = _sort(yx, cuda, keys, b0, b1)
        ek, ei = _expect(keys, b0, b1)
        assert np.array_equal(k, ek) and np.array_equal(i, ei), (kind, b0, b1)


@pytest.mark.gpu
def test_sort_carries_given_ids(yx, cuda):
    rng = np.random.default_rng(3)
    n = 123_457
    keys = _keys(rng, "few", n, np.uint32)
    ids = rng.permutation(n).astype(np.uint32) * np.uint32(3) + np.uint32(11)
    k, i = _sort(yx, cuda, keys, 0, 32, ids=ids)
    ek, order = _expect(keys, 0, 32)
    assert np.array_equal(k, ek) and np.array_equal(i, ids[order])
    # keys only
    k, i = _sort(yx, cuda, keys, 4, 20, with_ids=False)
    assert i is None and np.array_equal(k, _expect(keys, 4, 20)[0])


@pytest.mark.gpu
def test_sort_empty_bit_range_copies(yx, cuda):
    rng = np.random.default_rng(4)
    keys = _keys(rng, "uniform", 5000, np.uint32)
    ids = np.arange(5000, dtype=np.uint32)[::-1].copy()
    k, i = _sort(yx, cuda, keys, 12, 12, ids=ids)
    assert np.array_equal(k, keys) and np.array_equal(i, ids)


@pytest.mark.gpu
@pytest.mark.parametrize("n,n_dev", [(100_003, 77_777), (100_003, 0), (100_003, 1), (9000, 9000),
                                     (9000, 20_000), (3_000_017, 1_234_567)])
def test_sort_device_count(yx, cuda, n, n_dev):
    rng = np.random.default_rng(n_dev + 1)
    keys = _keys(rng, "runs", n, np.uint32)
    k, i = _sort(yx, cuda, keys, 9, 25, n_dev=n_dev)
    m = min(n_dev, n)
    ek, ei = _expect(keys[:m], 9, 25)
    assert np.array_equal(k[:m], ek) and np.array_equal(i[:m], ei)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.uint32, np.uint64])
@pytest.mark.parametrize("n", [5, 4099, 100_003])
def test_sort_unaligned_keys(yx, cuda, dtype, n):
    """Keys that start one element past a 16-byte boundary take the upsweep's
    one-key-per-lane path (aligned keys are read 16 bytes at a time)."""
    lib = yx.lib()
    rng = np.random.default_rng(n)
    keys = _keys(rng, "runs", n + 1, dtype)
    assert keys.size == n + 1
    dk = _dev(cuda, keys)
    assert dk.data_ptr() % 16 == 0
    ko = cuda.empty(n, dtype=dk.dtype, device="cuda")
    io = cuda.empty(n, dtype=cuda.int32, device="cuda")
    fn = lib.yx_prim_sort_u32 if dtype == np.uint32 else lib.yx_prim_sort_u64
    bits = (3, 27) if dtype == np.uint32 else (5, 45)
    yx._lib.check(fn(dk.data_ptr() + keys.itemsize, None, n, None, bits[0], bits[1], ko.data_ptr(),
                     io.data_ptr(), _stream(cuda)))
    ek, ei = _expect(keys[1:], *bits)
    assert np.array_equal(_host(ko, dtype), ek) and np.array_equal(_host(io, np.uint32), ei)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 4097, 100_003, 2_000_003])
def test_sort_u64(yx, cuda, n):
    rng = np.random.default_rng(n + 5)
    keys = _keys(rng, "uniform", n, np.uint64)
    for b0, b1 in [(0, 64), (0, 33), (30, 61), (60, 64)]:
        k, i = _sort(yx, cuda, keys, b0, b1)
        ek, ei = _expect(keys, b0, b1)
        assert np.array_equal(k, ek) and np.array_equal(i, ei), (n, b0, b1)
    # Morton-like codes of a depth-11 build: 33 bits, many shared prefixes
    codes = np.sort(rng.integers(0, 1 << 20, size=n, dtype=np.uint64)) << np.uint64(13)
    codes ^= rng.integers(0, 1 << 13, size=n, dtype=np.uint64)
    codes = codes[rng.permutation(n)]
    k, i = _sort(yx, cuda, codes, 0, 33)
    ek, ei = _expect(codes, 0, 33)
    assert np.array_equal(k, ek) and np.array_equal(i, ei)


def _scan(yx, torch, fn, a, out_dtype):
    n = a.size
    d = _dev(torch, a)
    out = torch.empty(n, dtype=torch.int32 if out_dtype == np.uint32 else torch.int64, device="cuda")
    tot = torch.zeros(1, dtype=out.dtype, device="cuda")
    yx._lib.check(fn(d.data_ptr(), out.data_ptr(), n, tot.data_ptr(), _stream(torch)))
    return _host(out, out_dtype), _host(tot, out_dtype)[0]


def _excl(a, dtype):
    c = np.cumsum(a.astype(np.uint64), dtype=np.uint64)
    e = np.concatenate([[0], c[:-1]]).astype(np.uint64) if a.size else c
    tot = c[-1] if a.size else np.uint64(0)
    if dtype == np.uint32:
        return (e & 0xFFFFFFFF).astype(np.uint32), np.uint32(int(tot) & 0xFFFFFFFF)
    return e, np.uint64(tot)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 2, 4095, 4096, 4097, 8191, 8192, 8193, 65_537, 1_000_003, 5_000_011])
def test_scans(yx, cuda, n):
    lib = yx.lib()
    rng = np.random.default_rng(n)
    a32 = rng.integers(0, 1 << 31, size=n, dtype=np.uint32)  # the u32 sums wrap
    out, tot = _scan(yx, cuda, lib.yx_prim_scan_u32, a32, np.uint32)
    e, et = _excl(a32, np.uint32)
    assert np.array_equal(out, e) and tot == et
    out, tot = _scan(yx, cuda, lib.yx_prim_scan_u32_u64, a32, np.uint64)
    e, et = _excl(a32, np.uint64)
    assert np.array_equal(out, e) and tot == et
    a8 = rng.integers(0, 256, size=n, dtype=np.uint8)
    out, tot = _scan(yx, cuda, lib.yx_prim_scan_u8, a8, np.uint32)
    e, et = _excl(a8, np.uint32)
    assert np.array_equal(out, e) and tot == et
    flags = (rng.random(n) < 0.3).astype(np.uint8)
    out, tot = _scan(yx, cuda, lib.yx_prim_scan_u8, flags, np.uint32)
    e, et = _excl(flags, np.uint32)
    assert np.array_equal(out, e) and tot == et
    d = _dev(cuda, a32)
    t = cuda.zeros(1, dtype=cuda.int32, device="cuda")
    yx._lib.check(lib.yx_prim_reduce_u32(d.data_ptr(), n, t.data_ptr(), _stream(cuda)))
    assert _host(t, np.uint32)[0] == _excl(a32, np.uint32)[1]


@pytest.mark.gpu
def test_scan_in_place(yx, cuda):
    lib = yx.lib()
    rng = np.random.default_rng(9)
    a = rng.integers(0, 1000, size=300_001, dtype=np.uint32)
    d = _dev(cuda, a)
    yx._lib.check(lib.yx_prim_scan_u32(d.data_ptr(), d.data_ptr(), a.size, None, _stream(cuda)))
    assert np.array_equal(_host(d, np.uint32), _excl(a, np.uint32)[0])


@pytest.mark.gpu
@pytest.mark.parametrize("n,shift", [(1, 0), (4097, 3), (100_003, 3), (2_000_003, 6), (50_000, 0), (50_000, 63)])
def test_distinct(yx, cuda, n, shift):
    lib = yx.lib()
    rng = np.random.default_rng(n + shift)
    a = np.sort(rng.integers(0, 1 << 40, size=n, dtype=np.uint64) >> np.uint64(rng.integers(0, 30)))
    d = _dev(cuda, a)
    out = cuda.empty(n, dtype=cuda.int64, device="cuda")
    cnt = cuda.zeros(1, dtype=cuda.int64, device="cuda")
    yx._lib.check(lib.yx_prim_distinct_u64(d.data_ptr(), out.data_ptr(), n, shift, cnt.data_ptr(), _stream(cuda)))
    m = int(_host(cnt, np.uint64)[0])
    exp = np.unique(a >> np.uint64(shift))
    assert m == exp.size and np.array_equal(_host(out, np.uint64)[:m], exp)


def test_argument_errors(yx):
    """Checked before any device work (runs without a GPU)."""
    lib = yx.lib()
    assert lib.yx_prim_sort_u32(None, None, 0, None, 0, 32, None, None, None) == 0  # nothing to sort
    assert lib.yx_prim_sort_u32(None, None, 10, None, 0, 32, None, None, None) == -1
    p = ctypes.c_void_p(16)  # never dereferenced: the bit range is rejected first
    assert lib.yx_prim_sort_u32(p, None, 10, None, 0, 33, p, None, None) == -1
    assert lib.yx_prim_sort_u32(p, None, 10, None, 9, 8, p, None, None) == -1
    assert lib.yx_prim_sort_u64(p, None, 10, None, -1, 8, p, None, None) == -1
    assert lib.yx_prim_sort_u64(p, None, 1 << 31, None, 0, 64, p, None, None) == -7  # YX_ERR_RANGE
    assert lib.yx_prim_distinct_u64(p, p, 10, 64, p, None) == -1
    assert lib.yx_prim_reduce_u32(p, 10, None, None) == -1
