"""Brick classification on the GPU (yx_voxtree_*: CalcDensity, MarkBricks,
PackBricks of reaction/voxtree/density.cu:32-60,131-176,210-221 and the column
scan of voxtree.py:88-89) against the numpy restatement, bit for bit, from the
smallest volume to the reference's 1024^3, plus uniform volumes, single
voxels at the corners, a capacity too small and bad sizes."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def u32(t):
    return t.cpu().numpy().view(np.uint32)


def check_all(yx, vo, vol, B):
    words = vo.pack_bits(vol.cpu().numpy())
    bits = yx.voxtree.pack_bits(vol)
    assert np.array_equal(u32(bits), words)
    exp = vo.classify(words, B)
    r = yx.voxtree.classify(bits, B)
    torch.cuda.synchronize()
    for k in ("density", "brick_state", "column_count", "packed"):
        got = u32(r[k])
        if not np.array_equal(got, exp[k]):
            bad = np.nonzero(got != exp[k])[0] if got.shape == exp[k].shape else []
            raise AssertionError(f"B={B} {k}: shapes {got.shape} {exp[k].shape}, {len(bad)} differ, first {bad[:4]}")
    start = u32(r["column_start"])
    assert np.array_equal(start[:-1], exp["column_start"]) and start[-1] == exp["total"] == r["total"]
    # the stage entry points give the same arrays
    V = B // 2
    d = yx.voxtree.calc_density(bits, B)
    st, cnt = yx.voxtree.mark_bricks(d, V)
    pk = yx.voxtree.pack_bricks(st, r["column_start"][:-1].contiguous(), V, r["total"])
    assert torch.equal(d, r["density"]) and torch.equal(st, r["brick_state"]) and torch.equal(cnt, r["column_count"])
    assert torch.equal(pk, r["packed"])
    return exp


@pytest.mark.parametrize("B", [128, 256, 512, 1024])
def test_classify_matches_oracle(yx, cuda, oracle_mod, B):
    vo = oracle_mod.VoxtreeOracle()
    print(f"voxtree B={B}", flush=True)
    exp = check_all(yx, vo, yx.workloads.voxtree_volume(B, B), B)
    G = B // 8
    print(f"  {exp['total']} mixed of {G ** 3} bricks", flush=True)
    assert 0 < exp["total"] < G ** 3


def test_uniform_and_single_voxels(yx, cuda, oracle_mod):
    vo = oracle_mod.VoxtreeOracle()
    B = 256
    G = B // 8
    exp = check_all(yx, vo, torch.ones((B, B, B), dtype=torch.bool, device="cuda"), B)
    assert exp["total"] == G ** 3 - (G - 1) ** 3
    exp = check_all(yx, vo, torch.zeros((B, B, B), dtype=torch.bool, device="cuda"), B)
    assert exp["total"] == 0
    for p in [(0, 0, 0), (B - 1, B - 1, B - 1), (0, B - 1, 64), (8, 16, 24)]:
        vol = torch.zeros((B, B, B), dtype=torch.bool, device="cuda")
        vol[p] = True
        check_all(yx, vo, vol, B)


def test_errors(yx, cuda):
    bits = yx.voxtree.pack_bits(yx.workloads.voxtree_volume(128, 3))
    r = yx.voxtree.classify(bits, 128)
    with pytest.raises(yx.YoxelError):
        yx.voxtree.classify(bits, 128, capacity=r["total"] - 1)
    ok = yx.voxtree.classify(bits, 128, capacity=r["total"])
    assert torch.equal(ok["packed"], r["packed"])
    for bad in (96, 64, 2048):
        with pytest.raises(yx.YoxelError):
            yx.voxtree.calc_density(bits, bad)
    with pytest.raises(yx.YoxelError):
        yx.voxtree.mark_bricks(r["density"], 32)
    with pytest.raises(yx.YoxelError):  # misaligned words
        yx.voxtree.calc_density(bits[1:], 128)
