"""Checkpoint interchange on the GPU (SparseGrid.save / load,
svox2.py:1526-1628): files in the reference's layouts, written here directly
with numpy key by key (current keys with float16 sh_data; the legacy "data"
array [density | sh] without radius or center), load onto the GPU, which runs
accelerate(); the links then equal the oracle's accel_dist_prop of the file's
links bit for bit, and a render and fused gradient of the loaded grid match
the FP64 oracle on the file's values (colours 2e-5, gradients as in
tests/conftest.py)."""
import numpy as np
import pytest
import torch

import oracle.oracle as O
import svox2_amd
from svox2_amd import scenes
from tests.conftest import assert_grad_close

pytestmark = pytest.mark.gpu

OPT = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)


def _scene():
    gen = torch.Generator().manual_seed(31)
    sc = scenes.make_blob_grid(64, scenes.BlobSpec(6, (0.08, 0.16), (20.0, 60.0)), gen)
    o, d = scenes.sphere_rays(4096, gen, radius=1.6, target=0.4)
    gt = torch.rand(4096, 3, generator=gen)
    return sc, o, d, gt


def _write(path, sc, layout):
    links = sc.links.numpy().astype(np.int32)
    dens = sc.density.float().numpy().reshape(-1, 1)
    sh = sc.sh.float().numpy()
    if layout == "current":
        np.savez(path, radius=np.float32([1.0, 1.0, 1.0]), center=np.float32([0.0, 0.0, 0.0]),
                 links=links, density_data=dens, sh_data=sh.astype(np.float16), basis_type=1)
        return dens, sh.astype(np.float16).astype(np.float32)
    np.savez_compressed(path, links=links, data=np.concatenate([dens, sh], 1))
    return dens, sh


@pytest.mark.parametrize("layout", ["current", "legacy_data"])
def test_load_reference_layout_render(gpu, tmp_path, layout):
    sc, o, d, gt = _scene()
    path = str(tmp_path / f"{layout}.npz")
    dens, sh = _write(path, sc, layout)
    grid = svox2_amd.SparseGrid.load(path, device=gpu)
    grid.opt.sigma_thresh, grid.opt.stop_thresh = OPT.sigma_thresh, OPT.stop_thresh
    links_acc = O.accel_dist_prop(sc.links.numpy())
    assert np.array_equal(grid.links.cpu().numpy(), links_acc)
    assert np.array_equal(grid.density_data.data.cpu().numpy(), dens)
    assert np.array_equal(grid.sh_data.data.cpu().numpy(), sh)
    assert torch.equal(grid.radius, torch.ones(3)) and torch.equal(grid.center, torch.zeros(3))
    N = np.array(links_acc.shape, np.float64)
    g = O.Grid(links_acc, dens.astype(np.float64).reshape(-1), sh.astype(np.float64), 9,
               offset=0.5 * N - 0.5, scaling=0.5 * N)
    rgb = grid.volume_render_fused(svox2_amd.Rays(o.to(gpu), d.to(gpu)), gt.to(gpu))
    rgb64, gd64, gs64 = O.mse_grads_fp64(g, o.numpy(), d.numpy(), OPT, gt.numpy())
    assert np.abs(rgb.cpu().numpy() - rgb64).max() < 2e-5
    assert_grad_close(grid.density_data.grad.cpu().numpy().reshape(-1), gd64, "grad_density")
    assert_grad_close(grid.sh_data.grad.cpu().numpy(), gs64, "grad_sh")


def test_save_load_roundtrip_gpu(gpu, tmp_path):
    """A GPU grid saved and loaded again (links already accelerated in the
    file: accelerate() is idempotent) renders the same colours bit for bit
    when its SH values are float16-exact."""
    sc, o, d, _ = _scene()
    sc.sh = sc.sh.half().double()
    grid = scenes.to_sparse_grid(sc, device=gpu)
    path = str(tmp_path / "rt.npz")
    grid.save(path)
    z = np.load(path)
    assert sorted(z.files) == ["basis_type", "center", "density_data", "links", "radius",
                               "sh_data"]
    assert z["sh_data"].dtype == np.float16 and z["links"].dtype == np.int32
    back = svox2_amd.SparseGrid.load(path, device=gpu)
    assert torch.equal(back.links, grid.links)
    rays = svox2_amd.Rays(o.to(gpu), d.to(gpu))
    with torch.no_grad():
        assert torch.equal(back.volume_render(rays), grid.volume_render(rays))
