"""The svox1 backend (render_svox1_kernel.cu) on the GPU against the FP64
oracle over the FP32 walk (oracle.svox1_fp64 for colours,
oracle.svox1_grads_fp64 for the reference's backward), on the 64^3 dense grid
and the accelerated 64^3 blob grid (empty blocks skipped), in float32 and
bfloat16 storage, with the stop's renormalisation and the opaque last sample.

Tolerances: colours within 2e-5 (absolute) of the oracle on the stored
values; gradients to the shared gradient tolerances; masks bit-exact."""
import numpy as np
import pytest
import torch

import oracle.oracle as O
import svox2_amd
from svox2_amd import _C, scenes
from tests.conftest import assert_grad_close, host_grid

pytestmark = pytest.mark.gpu


def _grid(sc, dev, dtype=torch.float32, stop=1e-7, opaque=False):
    g = scenes.to_sparse_grid(sc, device=dev, dtype=dtype)
    g.opt.backend = "svox1"
    g.opt.sigma_thresh = 1e-8
    g.opt.stop_thresh = stop
    g.opt.last_sample_opaque = opaque
    g.opt.background_brightness = 0.6
    return g


def _opts_of(grid):
    o = grid.opt
    return O.Opts(o.background_brightness, o.step_size, o.sigma_thresh, o.stop_thresh,
                  o.near_clip, o.use_spheric_clip, o.last_sample_opaque)


@pytest.fixture(scope="module")
def blobs(gpu):
    gen = torch.Generator().manual_seed(7)
    sc = scenes.make_blob_grid(64, scenes.SPARSE256, gen)
    o, d = scenes.sphere_rays(600, gen, radius=1.8, target=0.5)
    return sc, o.numpy(), d.numpy(), gen


def _cases(dense64, blobs):
    _, o0, d0, _, sc0 = dense64
    sc1, o1, d1, _ = blobs
    return (("dense64", sc0, np.ascontiguousarray(o0[:400]), np.ascontiguousarray(d0[:400])),
            ("blobs64", sc1, o1, d1))


def _rays(o, d, dev):
    return svox2_amd.Rays(torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev))


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_svox1_forward_vs_fp64(gpu, dense64, blobs, dtype):
    for label, sc, o, d in _cases(dense64, blobs):
        for stop, opaque in ((1e-7, False), (1e-2, False), (1e-7, True)):
            grid = _grid(sc, gpu, dtype, stop, opaque)
            hg = host_grid(sc, dtype, links=grid.links.cpu().numpy())
            rgb = _C.volume_render_svox1(grid._to_cpp(), _rays(o, d, gpu)._to_cpp(),
                                         grid.opt._to_cpp()).cpu().numpy()
            smp = O.march().svox1_samples(hg, o, d, _opts_of(grid))
            ref = O.svox1_fp64(hg, o, d, _opts_of(grid), torch.from_numpy(hg.density),
                               torch.from_numpy(hg.sh), samples=smp).numpy()
            err = np.abs(rgb - ref).max()
            print(f"svox1 {label} {dtype} stop {stop} opaque {opaque}: colour max|err| {err:.2e}, "
                  f"stopped rays {np.unique(smp['ray'][(smp['flag'] & 2) != 0]).size}")
            assert err < 2e-5, (label, stop, opaque)
            assert (np.abs(ref - 0.6) > 1e-3).any()
            if stop == 1e-2 and label == "dense64":
                assert ((smp["flag"] & 2) != 0).sum() > 50  # renormalised stops occur


def test_svox1_backward_vs_fp64(gpu, dense64, blobs):
    rng = np.random.default_rng(8)
    for label, sc, o, d in _cases(dense64, blobs):
        for stop in (1e-7, 1e-2):
            grid = _grid(sc, gpu, stop=stop)
            hg = host_grid(sc, links=grid.links.cpu().numpy())
            rays = _rays(o, d, gpu)
            rgb = _C.volume_render_svox1(grid._to_cpp(), rays._to_cpp(), grid.opt._to_cpp())
            gout = rng.normal(size=(len(o), 3)).astype(np.float32)
            gd, gs, _, _ = grid._get_data_grads()
            gd.zero_()
            gs.zero_()
            mask = torch.zeros(gd.size(0), dtype=torch.bool, device=gpu)
            holder = _C.GridOutputGrads()
            holder.grad_density_out, holder.grad_sh_out, holder.mask_out = gd, gs, mask
            _C.volume_render_svox1_backward(grid._to_cpp(), rays._to_cpp(), grid.opt._to_cpp(),
                                            torch.from_numpy(gout).to(gpu), rgb, holder)
            rgd, rgs, rmask = O.svox1_grads_fp64(hg, o, d, _opts_of(grid),
                                                 gout.astype(np.float64),
                                                 rgb.cpu().numpy().astype(np.float64))
            assert_grad_close(gd.cpu().numpy().reshape(-1), rgd, f"{label} {stop} grad density")
            assert_grad_close(gs.cpu().numpy(), rgs, f"{label} {stop} grad sh")
            assert np.array_equal(mask.cpu().numpy(), rmask), (label, stop)
            assert rmask.sum() > 50


def test_svox1_fused_vs_forward_and_fp64(gpu, blobs):
    """The fused call's colours equal the forward's bit for bit; its gradients
    are the backward's for dL/drgb = (rgb - gt) * 2 / (3 n); beta and
    sparsity are ignored."""
    sc, o, d, gen = blobs
    grid = _grid(sc, gpu)
    hg = host_grid(sc, links=grid.links.cpu().numpy())
    gt = torch.rand(len(o), 3, generator=gen)
    rays = _rays(o, d, gpu)
    rgb_f = _C.volume_render_svox1(grid._to_cpp(), rays._to_cpp(), grid.opt._to_cpp())
    gd, gs, _, _ = grid._get_data_grads()
    gd.zero_()
    gs.zero_()
    mask = torch.zeros(gd.size(0), dtype=torch.bool, device=gpu)
    holder = _C.GridOutputGrads()
    holder.grad_density_out, holder.grad_sh_out, holder.mask_out = gd, gs, mask
    rgb = torch.empty(len(o), 3, device=gpu)
    _C.volume_render_svox1_fused(grid._to_cpp(), rays._to_cpp(), grid.opt._to_cpp(),
                                 gt.to(gpu), 0.5, 0.5, rgb, holder)
    assert torch.equal(rgb, rgb_f)
    rgb64 = rgb.double().cpu().numpy()
    gout = (rgb64 - gt.double().numpy()) * (2.0 / (3 * len(o)))
    rgd, rgs, rmask = O.svox1_grads_fp64(hg, o, d, _opts_of(grid), gout, rgb64)
    assert_grad_close(gd.cpu().numpy().reshape(-1), rgd, "fused grad density")
    assert_grad_close(gs.cpu().numpy(), rgs, "fused grad sh")
    assert np.array_equal(mask.cpu().numpy(), rmask)


def test_svox1_facade(gpu, blobs):
    """SparseGrid with backend 'svox1': volume_render (autograd) with an MSE
    loss gives the fused call's gradients; an image renders through the ray
    path with the ray kernel's colours; ray clips are refused."""
    sc, o, d, gen = blobs
    grid = _grid(sc, gpu)
    rays = _rays(o, d, gpu)
    gt = torch.rand(len(o), 3, generator=gen).to(gpu)
    rgb_fused = grid.volume_render_fused(rays, gt).detach().clone()
    gd_f = grid.density_data.grad.cpu().numpy().copy()
    gs_f = grid.sh_data.grad.cpu().numpy().copy()
    grid.density_data.grad = None
    grid.sh_data.grad = None
    rgb = grid.volume_render(rays)
    torch.nn.functional.mse_loss(rgb, gt).backward()
    assert torch.equal(rgb.detach(), rgb_fused)
    rel = lambda a, b: np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
    assert rel(grid.density_data.grad.cpu().numpy(), gd_f) < 1e-5
    assert rel(grid.sh_data.grad.cpu().numpy(), gs_f) < 1e-5
    cams = scenes.CameraSet(scenes.hemisphere_cameras(2, torch.Generator().manual_seed(1)),
                            width=40, height=30, focal=50.0)
    cam = cams.camera(0)
    cam.c2w = cam.c2w.to(gpu)
    with torch.no_grad():
        img = grid.volume_render_image(cam, batch_size=512)
        ref = _C.volume_render_svox1(grid._to_cpp(), cam.gen_rays()._to_cpp(), grid.opt._to_cpp())
    assert torch.equal(img.reshape(-1, 3), ref)
    clipped = svox2_amd.Rays(rays.origins, rays.dirs,
                             torch.zeros(len(o), 2, device=gpu) + torch.tensor([0.0, 9.0],
                                                                               device=gpu))
    with pytest.raises(RuntimeError, match="clip"):
        _C.volume_render_svox1(grid._to_cpp(), clipped._to_cpp(), grid.opt._to_cpp())
