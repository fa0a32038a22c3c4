"""C ABI: the library loads, exports every entry point include/yoxel.h
declares, and reports errors for the failures the reference can be forced into
(no scene, zero resolution, bad arguments) without touching a GPU."""
import ctypes
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(yx_[a-z0-9_]+)\s*\(", text))
    return names


def test_exports_every_declared_symbol(yx):
    lib = ctypes.CDLL(str(yx.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header(yx):
    from yoxel_voxel_amd import _lib
    assert declared_symbols() <= set(_lib.SIGNATURES)


def test_version(yx):
    assert yx.lib().yx_version().startswith(b"yoxel")


def test_renderer_state_and_errors(yx):
    r = yx.SVORenderer()
    # RendererBase defaults (cell/renderer_base.h:25)
    assert r.GetResolution() == (640, 480)
    r.SetResolution(32, 24)
    assert r.GetResolution() == (32, 24)
    # RenderFrame without a scene returns NULL (cell/ppu_renderer.cpp:78-79)
    assert r.RenderFrame() is None
    with pytest.raises(yx.YoxelError) as e:
        r.SetResolution(0, 10)
    assert e.value.code == -1
    with pytest.raises(yx.YoxelError) as e:
        r.SetFOV(0.0)
    assert e.value.code == -1
    with pytest.raises(yx.YoxelError) as e:
        r.SetLight((0, 0, 1), 7)
    assert e.value.code == -1
    with pytest.raises(yx.YoxelError) as e:
        r.Render(rgba=0, stream=0)
    assert e.value.code == -6  # YX_ERR_NO_SCENE
    lib = yx.lib()
    assert lib.yx_renderer_set_scene(None, None) == -1
    assert b"null" in lib.yx_last_error()


def test_ray_basis_matches_oracle(yx, orc):
    r = yx.SVORenderer()
    r.SetViewPos((0.5, 0.5, 0.3))
    r.SetViewDir((-1, -1, -1.5))  # the reference test view, cell/main.cpp:25-27
    r.SetResolution(1024, 768)
    b = r.ray_basis()
    ob = orc.init_ray_dir((-1, -1, -1.5), (0, 0, 1), 70.0, 1024, 768)
    assert (b == ob).all()


def test_empty_scene_needs_no_device(yx):
    r = yx.SVORenderer()
    r.SetScene(yx.SVOData.build_synthetic(depth=3, num_spheres=0))
    assert r.scene_bytes() == 0
