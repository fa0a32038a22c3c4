"""Pins for the FP64 compositing reference (oracle/composite_oracle.c):
  * a single ray's result equals the front-to-back recurrence evaluated with
    numpy.cumprod in float64 over the voxels the ray samples;
  * an axis-aligned ray samples every voxel of its row once (closed-form count);
  * the analytic gradient equals central finite differences of the fp64 loss;
  * a planted sign mistake in the adjoint (ORACLE_PLANT=6) fails the FD pin;
  * the synthetic volume builder's brick occupancy and layout."""
import numpy as np
import pytest


def small_volume(yx, grid=32, brick=8, occupancy=0.5, seed=3):
    return yx.composite.Volume(grid=grid, brick=brick, occupancy=occupancy, seed=seed)


def softplus(x):
    return np.where(x > 30, x, np.log1p(np.exp(np.minimum(x, 30))))


def test_single_ray_matches_cumprod(yx, oracle_mod):
    co = oracle_mod.CompositeOracle()
    v = small_volume(yx, grid=64, occupancy=0.4)
    views = yx.composite.orbit_views(3, 60.0, 24, 20, seed=1)
    rgb, trans, ns = co.forward(v.table, v.params, 64, 8, views, 24, 20)
    checked = 0
    for ray in range(0, 3 * 24 * 20, 37):
        vox = co.ray_voxels(v.table, 64, 8, v.num_slots, views, 24, 20, ray)
        view, rem = divmod(ray, 24 * 20)
        y, x = divmod(rem, 24)
        assert vox.size == ns[view, y, x]
        q = v.params[vox].astype(np.float64)
        sig = softplus(q[:, 0])
        alpha = -np.expm1(-sig)
        T = np.concatenate([[1.0], np.cumprod(np.exp(-sig))])
        C = ((T[:-1] * alpha)[:, None] * q[:, 1:]).sum(axis=0)
        np.testing.assert_allclose(rgb[view, y, x], C, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(trans[view, y, x], T[-1], rtol=1e-12, atol=1e-300)
        checked += vox.size > 0
    assert checked > 5


def test_axis_ray_sample_count(yx, oracle_mod):
    co = oracle_mod.CompositeOracle()
    v = small_volume(yx, grid=64, occupancy=0.5, seed=4)
    n, b = 64, 8
    g = n // b
    table = v.table.reshape(g, g, g)  # [z, y, x]
    # one pixel looking along +x from x = -0.5 through the centre of voxel row (y, z)
    for (vy, vz) in [(5, 9), (33, 40), (63, 0)]:
        pos = np.array([-0.5, (vy + 0.5) / n, (vz + 0.5) / n], np.float32)
        view = np.zeros(12, np.float32)
        view[:3] = pos
        view[3:6] = (1.0, 0.0, 0.0)  # dir0; du = dv = 0
        rgb, trans, ns = co.forward(v.table, v.params, n, b, view[None], 1, 1)
        # lattice samples t = k/64 for k in [32, 96): x voxel = k - 32, each once
        expected = int((table[vz // b, vy // b, :] >= 0).sum()) * b
        assert ns[0, 0, 0] == expected


def _fd_check(co, v, n, b, views, w, h, eps=1e-3, nsel=24, seed=0):
    rng = np.random.default_rng(seed)
    rgb, _, _ = co.forward(v.table, v.params, n, b, views, w, h)
    target = (rgb + rng.normal(0, 0.05, size=rgb.shape)).astype(np.float32)
    # voxels that some ray samples
    nrays = views.shape[0] * w * h
    seen = set()
    for ray in range(nrays):
        seen.update(co.ray_voxels(v.table, n, b, v.num_slots, views, w, h, ray).tolist())
    sel = rng.choice(sorted(seen), size=min(nsel, len(seen)), replace=False)
    loss, grad, rays = co.grad(v.table, v.params, n, b, views, w, h, target, sel)
    fd = np.zeros_like(grad)
    for i, idx in enumerate(sel):
        for comp in range(4):
            lp = co.loss_rays(v.table, v.params, n, b, views, w, h, target, rays[i], idx, comp, eps)
            lm = co.loss_rays(v.table, v.params, n, b, views, w, h, target, rays[i], idx, comp, -eps)
            fd[i, comp] = (lp - lm) / (2 * eps)
    return grad, fd


def test_gradient_matches_central_differences(yx, oracle_mod):
    co = oracle_mod.CompositeOracle()
    v = small_volume(yx, grid=32, occupancy=0.5, seed=5)
    views = yx.composite.orbit_views(2, 60.0, 16, 16, seed=2)
    grad, fd = _fd_check(co, v, 32, 8, views, 16, 16)
    np.testing.assert_allclose(grad, fd, rtol=1e-5, atol=1e-9 * np.abs(fd).max())
    assert np.abs(grad).max() > 0


def test_planted_adjoint_sign_is_caught(yx, oracle_mod, tmp_path):
    bad = oracle_mod.CompositeOracle(oracle_mod.build(plant=6, out=tmp_path / "liboracle_plant6.so"))
    assert bad.plant() == 6
    v = small_volume(yx, grid=32, occupancy=0.5, seed=5)
    views = yx.composite.orbit_views(2, 60.0, 16, 16, seed=2)
    grad, fd = _fd_check(bad, v, 32, 8, views, 16, 16)
    with pytest.raises(AssertionError):
        np.testing.assert_allclose(grad, fd, rtol=1e-5, atol=1e-9 * np.abs(fd).max())


def test_volume_builder(yx):
    v = yx.composite.Volume(grid=64, brick=8, occupancy=0.25, seed=11)
    g = 8
    assert v.table.shape == (g ** 3,)
    alloc = v.table >= 0
    assert alloc.sum() == v.num_slots
    assert abs(v.num_slots - 0.25 * g ** 3) <= 3  # quantile threshold, ties aside
    # slots numbered in table order, params per slot
    assert np.array_equal(v.table[alloc], np.arange(v.num_slots))
    assert v.params.shape == (v.num_slots * 512, 4)
    assert np.all((v.params[:, 1:] >= 0) & (v.params[:, 1:] <= 1))
    # deterministic
    w = yx.composite.Volume(grid=64, brick=8, occupancy=0.25, seed=11)
    assert np.array_equal(v.params, w.params) and np.array_equal(v.table, w.table)
