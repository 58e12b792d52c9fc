"""GPU depth rendering (volume_render_expected_term / volume_render_sigma_thresh,
render_lerp_kernel_cuvol.cu:122-225,1154-1197) against the FP64 depth oracle
over the FP32 march's samples.

Tolerances: expected termination |d - d64| <= 2e-5 max(1, |d64|) (float32
weights and sums over up to a few hundred samples); first-crossing depth
within two float32 roundings of the FP64 product (t / step) ws of the same
float32 t and ws."""
import numpy as np
import pytest
import torch

import oracle.oracle as O
import svox2_amd
from svox2_amd import _C, scenes
from tests.conftest import host_grid

pytestmark = pytest.mark.gpu

OPT = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)


def _grid(sc, dev, dtype=torch.float32):
    g = scenes.to_sparse_grid(sc, device=dev, dtype=dtype)
    g.opt.sigma_thresh, g.opt.stop_thresh = OPT.sigma_thresh, OPT.stop_thresh
    return g


def _check(grid, g, o, d, dev, clip=None):
    rays = svox2_amd.Rays(o.to(dev), d.to(dev), None if clip is None else clip.to(dev))
    cl = None if clip is None else clip.numpy()
    e = grid.volume_render_depth(rays).cpu().numpy()
    e64 = O.depth_fp64(g, o.numpy(), d.numpy(), OPT, clip=cl)
    err = np.abs(e - e64) / np.maximum(1.0, np.abs(e64))
    print(f"expected depth: max err {err.max():.3e} (max depth {e64.max():.3f})")
    assert err.max() <= 2e-5
    assert (e64 > 0).any() and (e64 == 0).any() == (e == 0).any()
    for thr in (0.5, 5.0, 40.0):
        s = grid.volume_render_depth(rays, sigma_thresh=thr).cpu().numpy()
        s64 = O.depth_fp64(g, o.numpy(), d.numpy(), OPT, sigma_thresh=thr, clip=cl)
        assert np.array_equal(s == 0, s64 == 0), thr
        assert np.all(np.abs(s - s64) <= 2.0 ** -22 * np.abs(s64)), thr
    return e


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_depth_dense64(gpu, dense64, dtype):
    _, o, d, _, sc = dense64
    g = host_grid(sc, dtype)
    grid = _grid(sc, gpu, dtype)
    o, d = torch.from_numpy(o), torch.from_numpy(d)
    _check(grid, g, o, d, gpu)
    # rays that miss the grid get 0
    miss = grid.volume_render_depth(svox2_amd.Rays(torch.tensor([[3.0, 3.0, 3.0]], device=gpu),
                                                   torch.tensor([[1.0, 0.0, 0.0]], device=gpu)))
    assert float(miss[0]) == 0.0


def test_depth_sparse_with_clips_and_options(gpu, sparse256):
    """The 256^3 sparse grid (empty-space skips, cell map) with per-ray clips;
    the cell map on and off agree bit for bit; last_sample_opaque does not
    change depth (the reference's depth march has none)."""
    sc, grid, links_acc, o, d = sparse256
    g = host_grid(sc, links=links_acc)
    clip = scenes.config_clip(o.shape[0], torch.Generator().manual_seed(8))
    e = _check(grid, g, o, d, gpu, clip)
    rays = svox2_amd.Rays(o.to(gpu), d.to(gpu), clip.to(gpu))
    spec = grid._to_cpp()
    spec.cell_map = False
    e_nomap = _C.volume_render_expected_term(spec, rays._to_cpp(), grid.opt._to_cpp())
    assert np.array_equal(e, e_nomap.cpu().numpy())
    grid.opt.last_sample_opaque = True
    try:
        assert np.array_equal(e, grid.volume_render_depth(rays).cpu().numpy())
    finally:
        grid.opt.last_sample_opaque = False


def test_depth_image(gpu, dense64):
    """volume_render_depth_image equals the ray depth of Camera.gen_rays."""
    _, _, _, _, sc = dense64
    grid = _grid(sc, gpu)
    cams = scenes.CameraSet(scenes.hemisphere_cameras(2, torch.Generator().manual_seed(1)),
                            width=40, height=30, focal=50.0)
    cam = cams.camera(0)
    cam.c2w = cam.c2w.to(gpu)
    img = grid.volume_render_depth_image(cam, batch_size=512)
    assert img.shape == (30, 40)
    rays = cam.gen_rays()
    assert torch.equal(img.reshape(-1), grid.volume_render_depth(rays))
