"""Host-side argument checks of the compositing binding (composite.py): the C
ABI receives raw device pointers, so tensors whose bytes it would misread are
refused before the call."""
import pytest
import torch


def test_tensor_checks(yx):
    chk = yx.composite._check
    ok = torch.zeros((8, 4), dtype=torch.float32)
    # host tensors are refused (the entry points read device memory)
    with pytest.raises(TypeError, match="CUDA"):
        chk("params", ok, torch.float32, (4,))
    # wrong dtype and non-contiguous views are refused before the device check
    with pytest.raises(TypeError, match="float32"):
        chk("params", ok.double(), torch.float32, (4,))
    with pytest.raises(TypeError, match="contiguous"):
        chk("views", torch.zeros((12, 3)).t(), torch.float32, (12,))
    with pytest.raises(TypeError):
        chk("table", [1, 2, 3], torch.int32)


def test_forward_refuses_host_tensors(yx):
    comp = yx.composite.Compositor()
    table = torch.zeros(8, dtype=torch.int32)
    params = torch.zeros((8 * 64, 4), dtype=torch.float32)
    views = torch.zeros((1, 12), dtype=torch.float32)
    with pytest.raises(TypeError):
        comp.forward(table, params, 8, 4, views, 4, 4)
