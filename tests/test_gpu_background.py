"""The MSI background (render_lerp_kernel_cuvol.cu:386-607, :775-860) and its
total variation (loss_kernel.cu:248-333,606-641) on the GPU against the FP64
oracle (oracle.render_bg_fp64, oracle.msi_tv_grad_sparse), on the 64^3 blob
grid with a 16-pixel, 8-layer background with empty columns, negative
densities and colours on both sides of the clamp.

Tolerances: colours within 2e-5 of the oracle (the background's sample
positions go through float32 asin / atan2 on the GPU and float64 in the
oracle; measured errors are printed); gradients to the shared gradient
tolerances; masks as stated per test."""
import numpy as np
import pytest
import torch
from torch import nn

import oracle.oracle as O
import svox2_amd
from svox2_amd import _C, scenes
from tests.conftest import assert_grad_close, host_grid

pytestmark = pytest.mark.gpu

COLOUR_TOL = 2e-5


def _opts_of(grid):
    o = grid.opt
    return O.Opts(o.background_brightness, o.step_size, o.sigma_thresh, o.stop_thresh,
                  o.near_clip, o.use_spheric_clip, o.last_sample_opaque)


def _attach_background(grid, reso, nl, seed, sigma_hi=3.0):
    """Random MSI on grid: raster links with some empty columns, colours
    normal(0, 1.5), sigma uniform(-0.5, sigma_hi); returns its host copy."""
    gen = torch.Generator().manual_seed(seed)
    links = torch.arange(2 * reso * reso, dtype=torch.int32).view(2 * reso, reso).clone()
    empty = torch.randperm(links.numel(), generator=gen)[: links.numel() // 10]
    links.view(-1)[empty] = -1
    data = torch.empty(2 * reso * reso, nl, 4)
    data[..., :3] = torch.randn(2 * reso * reso, nl, 3, generator=gen) * 1.5
    data[..., 3] = torch.rand(2 * reso * reso, nl, generator=gen) * (sigma_hi + 0.5) - 0.5
    dev = grid.links.device
    grid.background_nlayers, grid.background_reso = nl, reso
    grid.background_links = links.to(dev)
    grid.background_data = nn.Parameter(data.to(dev))
    return O.Background(links.numpy(), data.double().numpy())


@pytest.fixture(scope="module")
def blobs(gpu):
    gen = torch.Generator().manual_seed(7)
    sc = scenes.make_blob_grid(64, scenes.SPARSE256, gen)
    o, d = scenes.sphere_rays(600, gen, radius=1.8, target=0.5)
    return sc, o, d, gen


def _grid(sc, dev, dtype=torch.float32):
    g = scenes.to_sparse_grid(sc, device=dev, dtype=dtype)
    g.opt.sigma_thresh = 1e-8
    return g


def _oracle_rgb(hg, bgm, o, d, opts, requires_grad=False, sparsity=False):
    dens = torch.tensor(hg.density, requires_grad=requires_grad)
    sh = torch.tensor(hg.sh, requires_grad=requires_grad)
    bd = torch.tensor(bgm.data, requires_grad=requires_grad)
    out = O.render_bg_fp64(hg, bgm, o, d, opts, dens, sh, bd, sparsity_energy=sparsity)
    return out, (dens, sh, bd)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_background_forward_vs_fp64(gpu, blobs, dtype):
    sc, o, d, _ = blobs
    for stop, sigma_hi in ((1e-7, 3.0), (1e-2, 30.0)):
        grid = _grid(sc, gpu, dtype)
        grid.opt.stop_thresh = stop
        grid.opt.background_brightness = 0.3
        bgm = _attach_background(grid, 16, 8, seed=3, sigma_hi=sigma_hi)
        hg = host_grid(sc, dtype, links=grid.links.cpu().numpy())
        with torch.no_grad():
            rgb = grid.volume_render(svox2_amd.Rays(o.to(gpu), d.to(gpu))).cpu().numpy()
        ref, _ = _oracle_rgb(hg, bgm, o.numpy(), d.numpy(), _opts_of(grid))
        ref = ref.numpy()
        fg_only = O.render_fp64(hg, o.numpy(), d.numpy(), _opts_of(grid),
                                torch.from_numpy(hg.density), torch.from_numpy(hg.sh)).numpy()
        err = np.abs(rgb - ref).max()
        print(f"background {dtype} stop {stop}: colour max|err| {err:.2e}, "
              f"background share {np.abs(ref - fg_only).max():.2f}")
        assert err < COLOUR_TOL
        assert np.abs(ref - fg_only).max() > 0.1  # the background shows


def test_background_image_matches_rays(gpu, blobs):
    """The in-kernel image path composites the background behind each pixel
    as the ray path does for the camera's rays."""
    sc, _, _, _ = blobs
    grid = _grid(sc, gpu)
    _attach_background(grid, 16, 8, seed=4)
    cams = scenes.CameraSet(scenes.hemisphere_cameras(2, torch.Generator().manual_seed(1)),
                            width=40, height=30, focal=30.0)
    cam = cams.camera(0)
    cam.c2w = cam.c2w.to(gpu)
    with torch.no_grad():
        img = grid.volume_render_image(cam)
        ref = grid.volume_render(cam.gen_rays())
    err = (img.reshape(-1, 3) - ref).abs().max().item()
    print(f"background image vs rays: max|err| {err:.2e}")
    assert err < 1e-5


def test_background_autograd_vs_fp64(gpu, blobs):
    """volume_render + backward: density, SH and background gradients against
    the oracle's autograd for a random dL/drgb."""
    sc, o, d, _ = blobs
    grid = _grid(sc, gpu)
    bgm = _attach_background(grid, 16, 8, seed=5)
    hg = host_grid(sc, links=grid.links.cpu().numpy())
    gout = torch.randn(len(o), 3, generator=torch.Generator().manual_seed(9))
    rgb = grid.volume_render(svox2_amd.Rays(o.to(gpu), d.to(gpu)))
    (rgb * gout.to(gpu)).sum().backward()
    ref, (dens, sh, bd) = _oracle_rgb(hg, bgm, o.numpy(), d.numpy(), _opts_of(grid), True)
    (ref * gout.double()).sum().backward()
    assert_grad_close(grid.density_data.grad.cpu().numpy().reshape(-1), dens.grad.numpy(),
                      "grad density")
    assert_grad_close(grid.sh_data.grad.cpu().numpy(), sh.grad.numpy(), "grad sh")
    assert_grad_close(grid.background_data.grad.cpu().numpy(), bd.grad.numpy(),
                      "grad background")
    assert np.abs(bd.grad.numpy()[..., 3]).max() > 0 and np.abs(bd.grad.numpy()[..., :3]).max() > 0


@pytest.mark.parametrize("sparsity", [0.0, 0.2])
def test_background_fused_vs_fp64(gpu, blobs, sparsity):
    """The fused MSE step: colours equal the forward's; gradients against the
    oracle's autograd of mean((rgb - gt)^2) (with the Cauchy sparsity term of
    the background's contributing samples; the grid's own sparsity term is
    checked elsewhere, so its density gradient is compared without sparsity
    only); the background mask covers every layer with a density gradient and
    only columns the rays reached."""
    sc, o, d, gen = blobs
    grid = _grid(sc, gpu)
    bgm = _attach_background(grid, 16, 8, seed=6)
    hg = host_grid(sc, links=grid.links.cpu().numpy())
    gt = torch.rand(len(o), 3, generator=torch.Generator().manual_seed(12))
    rays = svox2_amd.Rays(o.to(gpu), d.to(gpu))
    with torch.no_grad():
        rgb_fwd = grid.volume_render(rays)
    rgb = grid.volume_render_fused(rays, gt.to(gpu), sparsity_loss=sparsity)
    assert torch.equal(rgb, rgb_fwd)
    (ref, bg_sp), (dens, sh, bd) = _oracle_rgb(hg, bgm, o.numpy(), d.numpy(), _opts_of(grid),
                                               True, sparsity=True)
    loss = ((ref - gt.double()) ** 2).sum() / (3 * len(o)) + sparsity * bg_sp
    loss.backward()
    if sparsity == 0.0:
        assert_grad_close(grid.density_data.grad.cpu().numpy().reshape(-1), dens.grad.numpy(),
                          "fused grad density")
    assert_grad_close(grid.sh_data.grad.cpu().numpy(), sh.grad.numpy(), "fused grad sh")
    assert_grad_close(grid.background_data.grad.cpu().numpy(), bd.grad.numpy(),
                      "fused grad background")
    mask = grid.sparse_background_indexer.cpu().numpy()
    g_sigma = np.abs(bd.grad.numpy()[..., 3])
    assert mask[g_sigma > 1e-12 * g_sigma.max()].all()
    touched_rows = np.abs(bd.grad.numpy()).reshape(len(bd.grad), -1).max(1) > 0
    assert not mask[~touched_rows].any()


def test_background_fused_matches_autograd_large(gpu):
    """65536 rays (ordering workspace and the background's buffers after it,
    in both the fused call and the separate backward): the fused call's
    colours equal the autograd forward's and its gradients the autograd
    backward's for the MSE loss."""
    gen = torch.Generator().manual_seed(21)
    sc = scenes.make_blob_grid(64, scenes.SPARSE256, gen)
    o, d = scenes.sphere_rays(65536, gen, radius=1.8, target=0.5)
    grid = _grid(sc, gpu)
    _attach_background(grid, 32, 16, seed=8)
    rays = svox2_amd.Rays(o.to(gpu), d.to(gpu))
    gt = torch.rand(len(o), 3, generator=gen).to(gpu)
    rgb_f = grid.volume_render_fused(rays, gt).clone()
    g_f = [p.grad.clone() for p in (grid.density_data, grid.sh_data, grid.background_data)]
    for p in (grid.density_data, grid.sh_data, grid.background_data):
        p.grad = None
    rgb = grid.volume_render(rays)
    torch.nn.functional.mse_loss(rgb, gt).backward()
    assert torch.equal(rgb.detach(), rgb_f)
    for a, p in zip(g_f, (grid.density_data, grid.sh_data, grid.background_data)):
        b = p.grad
        rel = ((a - b).abs().max() / b.abs().max().clamp_min(1e-30)).item()
        assert rel < 1e-5, rel


def test_msi_tv_grad_sparse(gpu):
    """msi_tv_grad_sparse against the oracle on a background with empty
    columns, listed cells including the image edges and top layer, different
    colour and sigma scales; masks bit-exact.  The facade's draw
    (inplace_tv_background_grad) lists the cells the oracle is given when
    np.random is seeded alike."""
    rng = np.random.default_rng(3)
    reso, nl = 16, 8
    links = np.arange(2 * reso * reso, dtype=np.int32).reshape(2 * reso, reso)
    links.reshape(-1)[rng.choice(links.size, 50, replace=False)] = -1
    msi = rng.normal(size=(2 * reso * reso, nl, 4)).astype(np.float32)
    total = 2 * reso * reso * nl
    cells = np.concatenate([rng.choice(total, 3000, replace=False), [total - 1, nl - 1]])
    cells = cells.astype(np.int32)
    grad = torch.zeros(msi.shape, device=gpu)
    mask = torch.zeros(msi.shape[:2], dtype=torch.bool, device=gpu)
    _C.msi_tv_grad_sparse(torch.from_numpy(links).to(gpu), torch.from_numpy(msi).to(gpu),
                          torch.from_numpy(cells).to(gpu), mask, 0.7, 0.2, grad)
    ref, rmask = O.msi_tv_grad_sparse(links, msi.astype(np.float64), cells, 0.7, 0.2)
    g = grad.cpu().numpy()
    assert np.abs(g - ref).max() <= 1e-5 * np.abs(ref).max()
    assert np.array_equal(mask.cpu().numpy(), rmask)

    grid = _grid(scenes.make_blob_grid(16, scenes.SPARSE256, torch.Generator().manual_seed(1)),
                 gpu)
    bgm = _attach_background(grid, 8, 4, seed=2)
    gfac = torch.zeros_like(grid.background_data)
    np.random.seed(5)
    grid.inplace_tv_background_grad(gfac, scaling=0.5, scaling_density=0.1, sparse_frac=0.05,
                                    contiguous=True)
    np.random.seed(5)
    cells = grid._get_rand_cells_background(0.05, contiguous=True).cpu().numpy()
    ref, _ = O.msi_tv_grad_sparse(bgm.links, bgm.data, cells, 0.5, 0.1)
    assert np.abs(gfac.cpu().numpy() - ref).max() <= 1e-5 * np.abs(ref).max()


def test_background_optim_sparsify_checkpoint(gpu, blobs, tmp_path):
    """optim_background_step: RMSprop over the touched layers only, the colour
    channels at lr_color and sigma at lr_sigma (closed form of one step from
    zero state, where rms becomes g^2, optim_kernel.cu:21); sparsify_background: columns without a layer of sigma >=
    the threshold (after dilation) are dropped and the others renumbered in
    raster order, the render unchanged when the dropped columns hold zeros;
    save / load keeps the background and the render."""
    sc, o, d, gt = blobs
    grid = _grid(sc, gpu)
    _attach_background(grid, 16, 8, seed=7)
    rays = svox2_amd.Rays(o.to(gpu), d.to(gpu))
    grid.volume_render_fused(rays, torch.rand(len(o), 3, generator=gt).to(gpu))
    mask = grid.sparse_background_indexer.clone()
    data0 = grid.background_data.data.clone()
    g = grid.background_data.grad.clone()
    grid.optim_background_step(lr_sigma=0.3, lr_color=0.05, beta=0.9, epsilon=1e-8)
    rms = g * g  # a zero state takes the first squared gradient whole
    lr = torch.tensor([0.05, 0.05, 0.05, 0.3], device=gpu)
    expect = data0 - lr * g / (torch.sqrt(rms) + 1e-8)
    m = mask[..., None].expand_as(data0)
    got = grid.background_data.data
    assert torch.equal(got[~m], data0[~m])
    assert torch.allclose(got[m], expect[m], rtol=1e-5, atol=1e-6)
    assert m.any()

    # sparsify: keep columns whose sigma reaches 1.0 somewhere; zero the rest
    reso, nl = 16, 8
    bd = grid.background_data.data
    col_keep = (bd[..., 3] >= 1.0).any(-1)
    links = grid.background_links.clone()
    rows_of = links.view(-1)
    dropped = (rows_of >= 0) & ~col_keep[rows_of.clamp_min(0).long()]
    bd[rows_of[dropped].long()] = 0.0
    with torch.no_grad():
        before = grid.volume_render(rays).clone()
    grid.sparsify_background(sigma_thresh=1.0, dilate=0)
    kept = (rows_of >= 0) & ~dropped
    new_links = grid.background_links.view(-1)
    assert torch.equal(new_links[~kept], torch.full_like(new_links[~kept], -1))
    assert torch.equal(new_links[kept], torch.arange(int(kept.sum()), dtype=torch.int32,
                                                     device=gpu))
    assert grid.background_data.shape == (int(kept.sum()), nl, 4)
    with torch.no_grad():
        after = grid.volume_render(rays)
    assert torch.equal(after, before)

    path = str(tmp_path / "bg.npz")
    grid.save(path)
    g2 = svox2_amd.SparseGrid.load(path, device=gpu)
    g2.opt.sigma_thresh = grid.opt.sigma_thresh
    assert g2.use_background and torch.equal(g2.background_links, grid.background_links)
    with torch.no_grad():
        bg_only = g2.volume_render(rays) - grid.volume_render(rays)
    # SH data is stored as float16 in the checkpoint; the background is exact
    assert bg_only.abs().max().item() < 2e-2
    assert torch.equal(g2.background_data.data, grid.background_data.data)
