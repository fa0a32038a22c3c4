"""GPU parity of the differentiable compositing (forward and adjoint) against
the FP64 reference, through the C ABI. The discrete part (which samples a ray
takes) must match exactly (per-ray sample counts); colours and transmittance
within 1e-5; gradients within 1e-3 relative of the FP64 reference gradient and
of FP64 central finite differences (eps 1e-3) on 256 random voxels; two
backward passes give bitwise identical gradients."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def setup_case(yx, torch, grid, brick, occupancy, seed, n_views, w, h, cam_seed=0):
    v = yx.composite.Volume(grid=grid, brick=brick, occupancy=occupancy, seed=seed)
    views = yx.composite.orbit_views(n_views, 60.0, w, h, seed=cam_seed)
    dev = torch.device("cuda", 0)
    table, params = v.to(dev)
    vd = torch.from_numpy(views).to(dev)
    return v, views, table, params, vd


def perturbed_target(yx, torch, comp, v, table, params, vd, grid, brick, w, h, seed=1):
    rng = np.random.default_rng(seed)
    p2 = v.params.copy()
    p2[:, 0] += rng.normal(0.0, 0.1, size=p2.shape[0]).astype(np.float32)
    p2d = torch.from_numpy(p2).to(params.device)
    rgb_t, _, _ = comp.forward(table, p2d, grid, brick, vd, w, h)
    torch.cuda.synchronize()
    return rgb_t.cpu().numpy().astype(np.float32)


def check_grad(gpu, ref, what):
    err = np.linalg.norm(gpu - ref) / max(np.linalg.norm(ref), 1e-30)
    assert err < 1e-3, f"{what}: norm-wise relative error {err:.2e}"
    bad = np.abs(gpu - ref) > 1e-3 * np.abs(ref) + 1e-4 * np.abs(ref).max()
    assert not bad.any(), f"{what}: {int(bad.sum())} elements off, e.g. gpu {gpu[bad][:4]} ref {ref[bad][:4]}"


def run_parity(yx, orc_c, torch, grid, brick, occupancy, seed, n_views, w, h, nsel=256, threads=16,
               fd=True):
    v, views, table, params, vd = setup_case(yx, torch, grid, brick, occupancy, seed, n_views, w, h)
    comp = yx.composite.Compositor()
    rgb, trans, nsamp = comp.forward(table, params, grid, brick, vd, w, h)
    torch.cuda.synchronize()
    o_rgb, o_trans, o_ns = orc_c.forward(v.table, v.params, grid, brick, views, w, h, threads=threads)
    g_ns = nsamp.cpu().numpy().view(np.uint32)
    assert np.array_equal(g_ns, o_ns), f"sample counts differ at {int((g_ns != o_ns).sum())} rays"
    np.testing.assert_allclose(rgb.cpu().numpy(), o_rgb, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(trans.cpu().numpy(), o_trans, rtol=1e-5, atol=1e-6)

    target = perturbed_target(yx, torch, comp, v, table, params, vd, grid, brick, w, h)
    t_dev = torch.from_numpy(target).to(params.device)
    grad_rgb = 2.0 * (rgb - t_dev)
    g1 = comp.backward(table, params, grid, brick, vd, w, h, rgb, nsamp, grad_rgb)
    g2 = comp.backward(table, params, grid, brick, vd, w, h, rgb, nsamp, grad_rgb)
    torch.cuda.synchronize()
    g1n, g2n = g1.cpu().numpy(), g2.cpu().numpy()
    assert np.array_equal(g1n.view(np.uint32), g2n.view(np.uint32)), "backward is not deterministic"
    records, unique = comp.last_stats()
    assert records == int(o_ns.sum()) and 0 < unique <= records

    # voxels some ray sampled, 256 of them at random
    touched = np.nonzero(np.abs(g1n).sum(axis=1) > 0)[0]
    assert touched.size <= unique, f"{touched.size} voxels with a gradient, {unique} distinct keys counted"
    rng = np.random.default_rng(seed)
    sel = rng.choice(touched, size=min(nsel, touched.size), replace=False)
    loss, ref, rays = orc_c.grad(v.table, v.params, grid, brick, views, w, h, target, sel, max_rays=4096,
                                 threads=threads)
    check_grad(g1n[sel], ref, "gradient vs FP64")
    if fd:
        eps = 1e-3
        fdv = np.zeros_like(ref)
        for i, idx in enumerate(sel):
            for c in range(4):
                lp = orc_c.loss_rays(v.table, v.params, grid, brick, views, w, h, target, rays[i], idx, c, eps)
                lm = orc_c.loss_rays(v.table, v.params, grid, brick, views, w, h, target, rays[i], idx, c, -eps)
                fdv[i, c] = (lp - lm) / (2 * eps)
        check_grad(g1n[sel], fdv, "gradient vs central differences")
    return records


def test_small_grid_parity(yx, oracle_mod, cuda):
    run_parity(yx, oracle_mod.CompositeOracle(), cuda, 64, 8, 0.3, 3, 4, 48, 40, nsel=128)


def test_uneven_bricks_and_dense_volume(yx, oracle_mod, cuda):
    # brick 4, most bricks allocated, odd resolution
    run_parity(yx, oracle_mod.CompositeOracle(), cuda, 32, 4, 0.9, 8, 3, 33, 17, nsel=64)


@pytest.mark.parametrize("grid,brick,occupancy", [(16, 1, 0.5), (32, 2, 0.6), (64, 16, 0.5)])
def test_brick_sizes(yx, oracle_mod, cuda, grid, brick, occupancy):
    # bricks of 1 and 2 voxels go through the per-slot reduction with 1 and 8
    # voxel sums per slot; 16^3 bricks through the whole-key sort and run sums
    run_parity(yx, oracle_mod.CompositeOracle(), cuda, grid, brick, occupancy, 5, 2, 24, 20, nsel=64)


def test_composite_512_full(yx, oracle_mod, cuda):
    """The full workload: 512^3 as 8^3 bricks at ~10 % occupancy, 16 views of
    512x512 rays; forward checked on every ray, gradient on 256 voxels."""
    cfg = yx.composite.COMPOSITE_512
    print("composite_512 parity", flush=True)
    records = run_parity(yx, oracle_mod.CompositeOracle(), cuda, cfg["grid"], cfg["brick"], cfg["occupancy"],
                         cfg["seed"], 16, 512, 512, nsel=256, threads=16)
    assert records > 10_000_000


def test_autograd_function(yx, cuda):
    torch = cuda
    grid, brick, w, h = 64, 8, 32, 32
    v, views, table, params, vd = setup_case(yx, torch, grid, brick, 0.3, 3, 2, w, h)
    comp = yx.composite.Compositor()
    fn = yx.composite.composite_function()
    target = perturbed_target(yx, torch, comp, v, table, params, vd, grid, brick, w, h)
    t_dev = torch.from_numpy(target).to(params.device)
    p = params.clone().requires_grad_(True)
    rgb = fn.apply(p, table, vd, grid, brick, w, h, comp)
    loss = ((rgb - t_dev) ** 2).sum()
    loss.backward()
    # same as the library backward called directly
    rgb2, _, ns = comp.forward(table, params, grid, brick, vd, w, h)
    g = comp.backward(table, params, grid, brick, vd, w, h, rgb2, ns, 2.0 * (rgb2 - t_dev))
    assert torch.equal(p.grad, g)
    # plain gradient descent with a step scaled to the first gradient lowers the loss
    lr = 0.02 / float(p.grad.abs().max())
    losses = []
    for _ in range(15):
        p.grad = None
        l = ((fn.apply(p, table, vd, grid, brick, w, h, comp) - t_dev) ** 2).sum()
        l.backward()
        with torch.no_grad():
            p -= lr * p.grad
        losses.append(float(l.detach()))
    assert all(b <= a for a, b in zip(losses, losses[1:])) and losses[-1] < 0.9 * losses[0]


@pytest.mark.parametrize("brick", [8, 16])
def test_backward_capacity(yx, cuda, brick):
    """The record total stays on the device. A reserved capacity gives the same
    bits as a backward that sized its own workspace; a capacity below the total
    gives a NaN gradient and a reported error, and the next call grows the
    workspace and is exact again. brick 8 reduces per slot, 16 by whole keys."""
    torch = cuda
    grid, w, h = 64, 40, 32
    v, views, table, params, vd = setup_case(yx, torch, grid, brick, 0.4, 11, 3, w, h)
    ref = yx.composite.Compositor()
    rgb, _, ns = ref.forward(table, params, grid, brick, vd, w, h)
    g_rgb = torch.sin(rgb * 7.0).contiguous()
    g_ref = ref.backward(table, params, grid, brick, vd, w, h, rgb, ns, g_rgb)
    torch.cuda.synchronize()
    total = int(ns.to(torch.int64).sum())
    assert ref.last_stats()[0] == total and total > 1000

    exact = yx.composite.Compositor()
    exact.reserve(total)
    g = exact.backward(table, params, grid, brick, vd, w, h, rgb, ns, g_rgb)
    torch.cuda.synchronize()
    assert torch.equal(g, g_ref)

    small = yx.composite.Compositor()
    small.reserve(total // 2)
    g = small.backward(table, params, grid, brick, vd, w, h, rgb, ns, g_rgb)
    torch.cuda.synchronize()
    assert torch.isnan(g).all(), "an overflowed backward must not return a gradient"
    with pytest.raises(yx.YoxelError) as e:
        small.last_stats()
    assert e.value.code == yx._lib.YX_ERR_RANGE
    g = small.backward(table, params, grid, brick, vd, w, h, rgb, ns, g_rgb)
    torch.cuda.synchronize()
    assert torch.equal(g, g_ref)
    assert small.last_stats()[0] == total


def test_table_edited_between_calls(yx, oracle_mod, cuda):
    """Both walks skip whole empty super-bricks (4^3 bricks) from a summary of
    the brick table that every call rebuilds: after the table is edited in
    place (bricks released, others allocated into their slots), a compositor
    that already ran on the old table gives the oracle's sample counts and
    the same bits as a fresh compositor, forward and backward."""
    torch = cuda
    grid, brick, w, h = 128, 8, 48, 40
    v, views, table, params, vd = setup_case(yx, torch, grid, brick, 0.15, 21, 3, w, h)
    comp = yx.composite.Compositor()
    rgb, _, ns = comp.forward(table, params, grid, brick, vd, w, h)
    # room for the edited table's records too (its total may be larger)
    comp.reserve(4 * int(ns.to(torch.int64).sum()) + 4096)
    comp.backward(table, params, grid, brick, vd, w, h, rgb, ns, torch.ones_like(rgb))
    torch.cuda.synchronize()
    # release every allocated brick of one half of the volume and give their
    # slots to empty bricks of the other half, in place
    g = grid // brick
    tab = v.table.copy().reshape(g, g, g)
    alloc = np.argwhere(tab >= 0)
    free = np.argwhere(tab < 0)
    gone = alloc[alloc[:, 2] < g // 2]
    new = free[free[:, 2] >= g // 2][: len(gone)]
    slots = tab[gone[:, 0], gone[:, 1], gone[:, 2]].copy()
    tab[gone[:, 0], gone[:, 1], gone[:, 2]] = -1
    tab[new[:, 0], new[:, 1], new[:, 2]] = slots[: len(new)]
    assert len(gone) > 10 and len(new) == len(gone)
    table.copy_(torch.from_numpy(tab.reshape(-1)))
    rgb1, tr1, ns1 = comp.forward(table, params, grid, brick, vd, w, h)
    fresh = yx.composite.Compositor()
    rgb2, tr2, ns2 = fresh.forward(table, params, grid, brick, vd, w, h)
    torch.cuda.synchronize()
    _, _, o_ns = oracle_mod.CompositeOracle().forward(tab.reshape(-1), v.params, grid, brick, views, w, h,
                                                      threads=8)
    assert np.array_equal(ns1.cpu().numpy().view(np.uint32), o_ns)
    assert not np.array_equal(ns1.cpu().numpy(), ns.cpu().numpy()), "the edit must change the samples"
    assert torch.equal(rgb1, rgb2) and torch.equal(tr1, tr2) and torch.equal(ns1, ns2)
    gr = torch.sin(rgb1 * 5.0).contiguous()
    g1 = comp.backward(table, params, grid, brick, vd, w, h, rgb1, ns1, gr)
    g2 = fresh.backward(table, params, grid, brick, vd, w, h, rgb2, ns2, gr)
    torch.cuda.synchronize()
    assert torch.equal(g1, g2)
