"""Host-side behaviour of the shading-pipeline settings (no GPU work): the
SVORenderer defaults (demo/SVORenderer.cpp:6-16), argument checks, and the
missing-scene error of yx_render_pipeline."""
import math

import pytest


def test_defaults_and_setters(yx):
    r = yx.SVORenderer()
    assert r.GetDetailCoef() == 1.0  # m_detailCoef(1.0)
    assert r.GetSSNA() is True  # m_ssna(true)
    r.SetDetailCoef(2.5)
    r.SetSSNA(False)
    assert r.GetDetailCoef() == 2.5 and r.GetSSNA() is False
    r.SetShowNormals(True)
    r.SetLigth(0, (0.5, 0.5, 2.0))
    r.SetLigth(3, (0.0, 0.0, 0.0), enabled=False)


def test_argument_checks(yx):
    r = yx.SVORenderer()
    for bad in (-1.0, math.nan, math.inf):
        with pytest.raises(yx.YoxelError):
            r.SetDetailCoef(bad)
    for i in (-1, 4):
        with pytest.raises(yx.YoxelError):
            r.SetLigth(i, (0, 0, 0))
    with pytest.raises(yx.YoxelError):
        r.SetLigth(0, (0, 0, 0), attenuation=(0.0, 1.0, 0.0))
    with pytest.raises(yx.YoxelError):
        r.SetLigth(0, (0, 0, 0), attenuation=(1.0, -1.0, 0.0))
    with pytest.raises(yx.YoxelError):
        r.SetLigth(0, (math.nan, 0, 0))


def test_pipeline_without_scene(yx):
    r = yx.SVORenderer()
    with pytest.raises(yx.YoxelError) as e:
        r.RenderPipeline(stream=0)
    assert e.value.code == -6  # YX_ERR_NO_SCENE
