"""Pins of the march oracle (FP32 C restatement + FP64 compositing) against
closed forms, background-only cases, the literal restatement of the
reference's pure-PyTorch renderer, and FP64 finite differences."""
import math

import numpy as np
import pytest
import torch

import oracle.oracle as O
from tests.conftest import host_grid
from svox2_amd import scenes


def _uniform_grid(N, sigma, dc, basis_dim=9):
    links = np.arange(N ** 3, dtype=np.int32).reshape(N, N, N)
    dens = np.full(N ** 3, sigma, np.float64)
    sh = np.zeros((N ** 3, 3 * basis_dim))
    for c in range(3):
        sh[:, c * basis_dim] = dc[c]
    Nf = np.array([N, N, N], np.float64)
    return O.Grid(links, dens, sh, basis_dim, offset=0.5 * Nf - 0.5, scaling=0.5 * Nf)


def _rays_through_box(n, seed):
    g = torch.Generator().manual_seed(seed)
    return scenes.sphere_rays(n, g, radius=1.7, target=0.5)


def test_uniform_field_closed_form(march):
    """sigma, colour uniform: C = c (1 - e^{-sigma ws M}) + bg e^{-sigma ws M},
    M = number of samples in [tmin, tmax]; M from FP64 geometry."""
    N, sigma = 16, 3.0
    dc = [0.8, -0.2, 0.3]
    g = _uniform_grid(N, sigma, dc)
    o, d = _rays_through_box(300, 0)
    o, d = o.numpy(), d.numpy()
    opt = O.Opts(background_brightness=0.7, stop_thresh=0.0, sigma_thresh=0.0)
    # FP64 geometry
    scale = N / 2.0
    og = o.astype(np.float64) * scale + (scale - 0.5)
    dg = d.astype(np.float64) * scale
    dn = np.linalg.norm(dg, axis=1)
    dgu = dg / dn[:, None]
    with np.errstate(divide="ignore"):
        t1 = (-0.5 - og) / dgu
        t2 = (N - 0.5 - og) / dgu
    tmin = np.maximum(np.minimum(t1, t2).max(1), 0.0)
    tmax = np.maximum(t1, t2).min(1)
    span = (tmax - tmin) / opt.step_size
    frac = span - np.floor(span)
    keep = (tmax > tmin) & (frac > 0.01) & (frac < 0.99)
    M = np.floor(span[keep]).astype(np.int64) + 1
    ws = opt.step_size / dn[keep]
    col = np.maximum(np.array(dc) * O.SH_C0 + 0.5, 0.0)
    T = np.exp(-sigma * ws * M)
    expect = col[None, :] * (1 - T)[:, None] + opt.background_brightness * T[:, None]

    smp = march.samples(g, o[keep], d[keep], opt)
    counts = np.bincount(smp["ray"], minlength=keep.sum())
    assert np.array_equal(counts, M)  # stepping pinned
    rgb64 = O.render_fp64(g, o[keep], d[keep], opt, torch.from_numpy(g.density),
                          torch.from_numpy(g.sh), samples=smp).numpy()
    assert np.abs(rgb64 - expect).max() < 1e-9
    rgb32, _, _ = march.render(g, o[keep], d[keep], opt)
    assert np.abs(rgb32 - expect).max() < 2e-5


def test_background_only_cases(march):
    N = 8
    opt = O.Opts(background_brightness=0.35)
    # zero density: exactly the background
    g = _uniform_grid(N, 0.0, [1.0, 1.0, 1.0])
    o, d = _rays_through_box(50, 1)
    rgb, lt, st = march.render(g, o.numpy(), d.numpy(), opt)
    assert np.all(rgb == np.float32(0.35)) and np.all(lt == 0)
    # rays missing the box
    g = _uniform_grid(N, 5.0, [1.0, 1.0, 1.0])
    o = np.array([[3.0, 3.0, 3.0], [0.0, 0.0, 5.0]], np.float32)
    d = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0]], np.float32)
    rgb, lt, st = march.render(g, o, d, opt)
    assert np.all(rgb == np.float32(0.35)) and np.all(st[:, 0] == 0)
    # all-empty links
    g.links[:] = -1
    o2, d2 = _rays_through_box(20, 2)
    rgb, _, _ = march.render(g, o2.numpy(), d2.numpy(), opt)
    assert np.all(rgb == np.float32(0.35))


def test_axis_aligned_rays(march):
    """Direction components exactly 0 are ignored by the slab test."""
    N, sigma = 8, 2.0
    g = _uniform_grid(N, sigma, [0.5, 0.5, 0.5])
    o = np.array([[-2.0, 0.1, 0.2], [0.1, -2.0, 0.3], [0.05, 0.1, 2.0]], np.float32)
    d = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, -1.0]], np.float32)
    opt = O.Opts(stop_thresh=0.0, sigma_thresh=0.0)
    setup = march.ray_setup(g, o, d, opt)
    assert np.all(setup[:, 7] == 4.0) and np.all(setup[:, 8] == 12.0)  # whole box crossed
    smp = march.samples(g, o, d, opt)
    counts = np.bincount(smp["ray"], minlength=3)
    assert np.all(counts == N * 2 + 1)
    rgb64 = O.render_fp64(g, o, d, opt, torch.from_numpy(g.density), torch.from_numpy(g.sh),
                          samples=smp).numpy()
    ws = 0.5 / (N / 2)
    T = math.exp(-sigma * ws * (2 * N + 1))
    col = 0.5 * O.SH_C0 + 0.5
    assert np.allclose(rgb64, col * (1 - T) + T, atol=1e-12)


def test_literal_reference_agrees(dense64, march):
    g, o, d, gt, sc = dense64
    opt = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)
    ref = O.volume_render_gradcheck_lerp(torch.from_numpy(g.links).long(),
                                         torch.from_numpy(g.density).view(-1, 1),
                                         torch.from_numpy(g.sh), 9, 1.0, 0.0,
                                         torch.from_numpy(o), torch.from_numpy(d),
                                         dtype=torch.float64).numpy()
    nothr = O.render_fp64(g, o, d, opt, torch.from_numpy(g.density), torch.from_numpy(g.sh),
                          thresholds=False).numpy()
    thr = O.render_fp64(g, o, d, opt, torch.from_numpy(g.density),
                        torch.from_numpy(g.sh)).numpy()
    rgb32, _, _ = march.render(g, o, d, opt)
    # positions differ only by FP32 vs FP64 stepping at 64^3
    assert np.abs(ref - nothr).max() < 2e-6
    # thresholds change the result by at most the skipped transmittance
    assert np.abs(thr - nothr).max() < 1e-6
    assert np.abs(rgb32 - thr).max() < 1e-5


def test_skip_matches_no_skip(march):
    """Empty-space skipping never changes the colour (skipped samples have only
    empty corners)."""
    gen = torch.Generator().manual_seed(5)
    sc = scenes.make_blob_grid(48, scenes.BlobSpec(6, (0.08, 0.12), (30.0, 60.0)), gen)
    links = O.accel_dist_prop(sc.links.numpy())
    g = host_grid(sc, links=links)
    o, d = scenes.sphere_rays(400, gen, radius=1.6, target=0.4)
    opt = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)
    rgb_s, _, st_s = march.render(g, o.numpy(), d.numpy(), opt, skip=True)
    rgb_n, _, st_n = march.render(g, o.numpy(), d.numpy(), opt, skip=False)
    assert st_s[:, 0].sum() < st_n[:, 0].sum()  # skipping happened
    assert np.abs(rgb_s - rgb_n).max() < 1e-5


def test_mse_gradient_finite_differences(march):
    """FP64 central differences of the oracle loss against its autograd
    gradient, including the beta and sparsity terms."""
    gen = torch.Generator().manual_seed(7)
    sc = scenes.make_blob_grid(12, scenes.BlobSpec(3, (0.3, 0.4), (3.0, 6.0)), gen, dense=True)
    g = host_grid(sc, torch.float64)
    o, d = scenes.sphere_rays(64, gen, radius=1.6, target=0.3)
    gt = torch.rand(64, 3, generator=gen, dtype=torch.float64).numpy()
    opt = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)
    smp = march.samples(g, o.numpy(), d.numpy(), opt)
    kw = dict(beta_loss=0.3, sparsity_loss=0.05)

    def parts(dens, sh):
        gg = O.Grid(g.links, dens, sh, 9, g.offset, g.scaling)
        extra = {}
        rgb = O._composite(gg, d.numpy(), opt, torch.from_numpy(dens), torch.from_numpy(sh),
                           dict(smp), 64, True, extra)
        return rgb, extra

    # the beta weight b_r is a constant of the backward (it is not differentiated)
    _, ex0 = parts(g.density, g.sh)
    t0 = ex0["t_end"]
    b = kw["beta_loss"] / 64 * (1.0 - t0 / (1.0 - t0 + 1e-3))

    def loss(dens, sh):
        rgb, extra = parts(dens, sh)
        val = ((rgb - torch.from_numpy(gt)) ** 2).mean() - (b * extra["depth"]).sum()
        s_ = extra["sigma"][extra["counted"]]
        return float(val + kw["sparsity_loss"] * torch.log1p(2 * s_ * s_).sum())

    _, gd, gs = O.mse_grads_fp64(g, o.numpy(), d.numpy(), opt, gt, samples=smp, **kw)
    rng = np.random.default_rng(0)
    h = 1e-4
    rows = np.nonzero(np.abs(gd) > 1e-6)[0]
    for r in rng.choice(rows, 6, replace=False):
        dp, dm = g.density.copy(), g.density.copy()
        dp[r] += h
        dm[r] -= h
        fd = (loss(dp, g.sh) - loss(dm, g.sh)) / (2 * h)
        assert fd == pytest.approx(gd[r], rel=1e-4, abs=1e-9)
    srows = np.argwhere(np.abs(gs) > 1e-6)
    for r, c in srows[rng.choice(len(srows), 6, replace=False)]:
        sp, sm = g.sh.copy(), g.sh.copy()
        sp[r, c] += h
        sm[r, c] -= h
        fd = (loss(g.density, sp) - loss(g.density, sm)) / (2 * h)
        assert fd == pytest.approx(gs[r, c], rel=1e-4, abs=1e-9)


def test_touched_rows_match_contributions(dense64, march):
    g, o, d, gt, sc = dense64
    opt = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)
    smp = march.samples(g, o[:256], d[:256], opt)
    m = O.touched_rows(g, smp)
    _, gd, gs = O.mse_grads_fp64(g, o[:256], d[:256], opt, gt[:256], samples=smp)
    # every row with a nonzero gradient is touched
    assert not np.any((np.abs(gd) > 0) & ~m)
    assert m.sum() > 0


@pytest.mark.parametrize("cut", [None, 5])
def test_touched_rows_axis_ray_closed_form(march, cut):
    """The touched set in closed form, both ways (nothing missing, nothing
    extra).  A ray along +x at grid (y, z) = (3.25, 4.5) through an N = 8
    grid, world = grid coordinates, uniform density 0.05 (no stop):
    samples sit at x = -0.5 + k/2 clamped to [0, 7], in cells
    l = (min(trunc(x), 6), 3, 4), and each marks its 8 corners
    (render_util.cuh:137-140), so the touched rows are the cells
    x in [0, 7], y in {3, 4}, z in {4, 5}.  With the density zeroed at
    x >= cut, samples in cells l_x >= cut have density 0 (all corners zero)
    and do not contribute, so x stops at cut."""
    N = 8
    links = np.arange(N ** 3, dtype=np.int32).reshape(N, N, N)
    dens = np.full(N ** 3, 0.05)
    if cut is not None:
        dens.reshape(N, N, N)[cut:] = 0.0
    sh = np.random.default_rng(0).standard_normal((N ** 3, 27))
    g = O.Grid(links, dens, sh, 9, offset=np.zeros(3), scaling=np.ones(3))
    opt = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)
    o = np.array([[-3.0, 3.25, 4.5]], np.float32)
    d = np.array([[1.0, 0.0, 0.0]], np.float32)
    smp = march.samples(g, o, d, opt)
    got = O.touched_rows(g, smp).reshape(N, N, N)
    want = np.zeros((N, N, N), bool)
    want[: N if cut is None else cut + 1, 3:5, 4:6] = True
    assert np.array_equal(got, want)
    # a hole in the links: its row is never touched, the rest stays
    links2 = links.copy()
    links2[2, 3, 4] = -1
    g2 = O.Grid(links2, dens, sh, 9, offset=np.zeros(3), scaling=np.ones(3))
    got2 = O.touched_rows(g2, march.samples(g2, o, d, opt)).reshape(N, N, N)
    want2 = want.copy()
    want2[2, 3, 4] = False
    assert np.array_equal(got2, want2)


def test_c_fused_port_matches_fp64(march):
    """The scalar C port of the fused forward+backward (the CPU baseline)
    against the FP64 autograd oracle, with per-ray clips, beta and sparsity."""
    gen = torch.Generator().manual_seed(21)
    sc = scenes.make_blob_grid(32, scenes.BlobSpec(4, (0.15, 0.25), (10.0, 40.0)), gen)
    links = O.accel_dist_prop(sc.links.numpy())
    g = host_grid(sc, links=links)
    o, d = scenes.sphere_rays(512, gen, radius=1.6, target=0.4)
    gt = torch.rand(512, 3, generator=gen, dtype=torch.float64).numpy()
    clip = scenes.config_clip(512, gen).numpy() * 0.7
    opt = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)
    kw = dict(beta_loss=0.2, sparsity_loss=0.01)
    rgb, gd, gs, m = march.fused(g, o.numpy(), d.numpy(), opt, gt, clip=clip, **kw)
    smp = march.samples(g, o.numpy(), d.numpy(), opt, clip=clip)
    rgb64, gd64, gs64 = O.mse_grads_fp64(g, o.numpy(), d.numpy(), opt, gt, samples=smp, **kw)
    assert np.abs(rgb - rgb64).max() < 1e-5
    assert np.abs(gd - gd64).max() < 1e-4 * np.abs(gd64).max()
    assert np.abs(gs - gs64).max() < 1e-4 * np.abs(gs64).max()
    assert np.array_equal(m, O.touched_rows(g, smp))


def test_compacted_blocks_match_whole_grid(march):
    """The FP64 oracle's fast path (rays marched in blocks of 2048 on worker
    threads; per block, samples with no row at any corner dropped and autograd
    over the used rows only) against autograd over the whole grid with every
    sample, on a sparse grid with per-ray clips and both extra loss terms:
    colours bit-identical (dropped samples add exact zeros), gradients equal
    to FP64 rounding.  The same holds when the sample list is passed in and
    cut into small chunks."""
    gen = torch.Generator().manual_seed(5)
    sc = scenes.make_blob_grid(32, scenes.BlobSpec(4, (0.15, 0.25), (10.0, 40.0)), gen)
    g = host_grid(sc, torch.float64)
    n = 4500
    o, d = scenes.sphere_rays(n, gen, radius=1.6, target=0.4)
    gt = torch.rand(n, 3, generator=gen, dtype=torch.float64).numpy()
    clip = scenes.config_clip(n, gen).numpy() * 0.7
    opt = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)
    kw = dict(beta_loss=0.2, sparsity_loss=0.01)
    smp = march.samples(g, o.numpy(), d.numpy(), opt, clip=clip)
    # Threaded blocks give the sample list of separate marches, in ray order.
    parts = [march.samples(g, o[a:a + 1500].numpy(), d[a:a + 1500].numpy(), opt,
                           clip=clip[a:a + 1500]) for a in (0, 1500, 3000)]
    for k in smp:
        cat = np.concatenate([q[k] + (a if k == "ray" else 0)
                              for q, a in zip(parts, (0, 1500, 3000))])
        assert np.array_equal(smp[k], cat), k
    dens = torch.from_numpy(np.asarray(g.density, np.float64).copy()).requires_grad_(True)
    sh = torch.from_numpy(np.asarray(g.sh, np.float64).copy()).requires_grad_(True)
    extra = {}
    rgb_w = O._composite(g, d.numpy(), opt, dens, sh, dict(smp), n, True, extra)
    t_end = extra["t_end"].detach()
    b = kw["beta_loss"] / n * (1.0 - t_end / (1.0 - t_end + 1e-3))
    s_ = extra["sigma"][extra["counted"]]
    loss = (((rgb_w - torch.from_numpy(gt)) ** 2).sum() / (3.0 * n)
            - (b * extra["depth"]).sum() + kw["sparsity_loss"] * torch.log1p(2 * s_ * s_).sum())
    loss.backward()
    for extra_kw in (dict(), dict(samples=smp, max_samples=20000)):
        rgb, gd, gs = O.mse_grads_fp64(g, o.numpy(), d.numpy(), opt, gt, clip=clip, **kw,
                                       **extra_kw)
        assert np.array_equal(rgb, rgb_w.detach().numpy())
        assert np.abs(gd - dens.grad.numpy()).max() <= 1e-12 * np.abs(dens.grad.numpy()).max()
        assert np.abs(gs - sh.grad.numpy()).max() <= 1e-12 * np.abs(sh.grad.numpy()).max()


def _axis_grid(N, density):
    links = np.arange(N ** 3, dtype=np.int32).reshape(N, N, N)
    sh = np.zeros((N ** 3, 27))
    return O.Grid(links, np.asarray(density, np.float64).reshape(-1), sh, 9,
                  offset=np.zeros(3), scaling=np.ones(3))


@pytest.mark.parametrize("sigma,stop", [(0.2, 1e-7), (5.0, 1e-7), (5.0, 0.0)])
def test_depth_expected_closed_form(march, sigma, stop):
    """Expected termination on a uniform field, world = grid coordinates, a
    ray along +x entering at x = -0.5: samples at t_k = 2.5 + k/2 up to
    t = 18.5 (x = 15.5), world step 1/2, weights a^k (1 - a) with
    a = exp(-sigma/2), the ray stopping after the first sample whose
    transmittance a^(k+1) falls below stop_thresh; depth = sum w_k t_k."""
    N = 16
    g = _axis_grid(N, np.full(N ** 3, sigma))
    o = np.array([[-3.0, 7.25, 8.5], [-3.0, 7.25, 8.5]], np.float32)
    d = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]], np.float32)  # the second points away
    opt = O.Opts(sigma_thresh=1e-8, stop_thresh=stop)
    k = np.arange(33)
    t = 2.5 + 0.5 * k
    a = np.exp(-sigma * 0.5)
    w = a ** k * (1 - a)
    if stop > 0:
        stopped = a ** (k + 1) < stop
        if stopped.any():
            last = int(np.argmax(stopped))
            w = w[:last + 1]
            t = t[:last + 1]
    got = O.depth_fp64(g, o, d, opt)
    assert got[0] == pytest.approx(float((w * t).sum()), rel=1e-12)
    assert got[1] == 0.0


def test_depth_sigma_thresh_closed_form(march):
    """First crossing on a linear ramp: density = x index, so the trilinear
    density at x is x itself (exact in float32); with threshold 6.3 the
    first sample above it is x = 6.5, at t = 9.5 along the ray from x = -3,
    depth (t / step) ws = 9.5.  A threshold above every density gives 0."""
    N = 16
    x = np.arange(N, dtype=np.float64)[:, None, None] * np.ones((1, N, N))
    g = _axis_grid(N, x)
    o = np.array([[-3.0, 7.25, 8.5]], np.float32)
    d = np.array([[1.0, 0.0, 0.0]], np.float32)
    opt = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)
    assert O.depth_fp64(g, o, d, opt, sigma_thresh=6.3)[0] == 9.5
    assert O.depth_fp64(g, o, d, opt, sigma_thresh=0.25)[0] == 3.5  # x = 0.5
    assert O.depth_fp64(g, o, d, opt, sigma_thresh=20.0)[0] == 0.0
