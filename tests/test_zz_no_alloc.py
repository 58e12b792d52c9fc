"""svox2_volume_render_cuvol_fused_noalloc through the C ABI: the fused call
whose workspace is the library's fixed device array.  Two calls in a row (the
second smaller, so the array still holds the first call's records) against
the fused call with a caller workspace from torch: colours and masks equal,
gradients within float32 atomic-order noise; a batch too large for the array
is refused with a message and changes nothing."""
import ctypes

import numpy as np
import pytest
import torch

import svox2_amd
from svox2_amd import _C, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("needs a GPU")
    _C.load_library()
    return torch.device("cuda:0")


def _grid(gpu):
    sc, _, _, _ = scenes.workload_dense64(seed=0, n_rays=16)
    grid = scenes.to_sparse_grid(sc, device=gpu)
    grid.opt.sigma_thresh = 1e-8
    grid.opt.stop_thresh = 1e-7
    return grid


def _call(lib, grid, rays, gt, noalloc):
    """One fused call through the C ABI; returns (status, colours, grad
    density, grad SH, mask) with gradients and mask starting from zero."""
    gd, gs, _, _ = grid._get_data_grads()
    gd.zero_()
    gs.zero_()
    n = rays.origins.size(0)
    mask = torch.zeros(grid.density_data.size(0), dtype=torch.bool, device=gd.device)
    holder = _C.GridOutputGrads()
    holder.grad_density_out, holder.grad_sh_out, holder.mask_out = gd, gs, mask
    g = _C._grid_struct(grid._to_cpp())
    oo, dd, clip = _C._rays(rays._to_cpp())
    gstruct = _C._grads_struct(holder, g.rows, g.sh_data_dim)
    out = torch.full((n, 3), -7.0, device=gd.device)
    opt = _C._opt_struct(grid.opt._to_cpp())
    if noalloc:
        st = lib.svox2_volume_render_cuvol_fused_noalloc(
            ctypes.byref(g), _C._ptr(oo), _C._ptr(dd), _C._ptr(clip), n, ctypes.byref(opt),
            _C._ptr(gt), 0.0, 0.0, _C._ptr(out), None, ctypes.byref(gstruct), _C._stream(oo))
    else:
        nbytes = lib.svox2_volume_render_cuvol_fused_workspace_bytes(n)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=gd.device)
        st = lib.svox2_volume_render_cuvol_fused(
            ctypes.byref(g), _C._ptr(oo), _C._ptr(dd), _C._ptr(clip), n, ctypes.byref(opt),
            _C._ptr(gt), 0.0, 0.0, _C._ptr(out), None, ctypes.byref(gstruct), _C._ptr(ws),
            nbytes, _C._stream(oo))
    torch.cuda.synchronize()
    return (st, out.cpu().numpy(), gd.cpu().numpy().copy(), gs.cpu().numpy().copy(),
            mask.cpu().numpy())


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def test_noalloc_matches_fused_with_workspace(gpu):
    lib = _C.load_library()
    grid = _grid(gpu)
    gen = torch.Generator().manual_seed(21)
    for n in (5000, 1200):  # the second call is smaller: stale records behind it
        o, d = scenes.sphere_rays(n, gen)
        rays = svox2_amd.Rays(o.to(gpu), d.to(gpu))
        gt = torch.rand(n, 3, generator=gen).to(gpu)
        st_a, rgb_a, gd_a, gs_a, m_a = _call(lib, grid, rays, gt, noalloc=True)
        assert st_a == 0, lib.svox2_last_error()
        st_b, rgb_b, gd_b, gs_b, m_b = _call(lib, grid, rays, gt, noalloc=False)
        assert st_b == 0, lib.svox2_last_error()
        print(f"noalloc n={n}: grad density rel {_rel(gd_a, gd_b):.2e}, "
              f"grad sh rel {_rel(gs_a, gs_b):.2e}", flush=True)
        assert np.array_equal(rgb_a, rgb_b), n
        assert np.array_equal(m_a, m_b), n
        assert m_a.any()
        # the same kernels on the same records; only the float32 atomic order differs
        assert _rel(gd_a, gd_b) < 1e-5 and _rel(gs_a, gs_b) < 1e-5, n


def test_noalloc_refuses_oversize_batch(gpu):
    lib = _C.load_library()
    grid = _grid(gpu)
    n = 40000  # about 180 MB of records, beyond the 40 MiB array
    assert lib.svox2_volume_render_cuvol_fused_workspace_bytes(n) > 40 << 20
    gen = torch.Generator().manual_seed(22)
    o, d = scenes.sphere_rays(n, gen)
    rays = svox2_amd.Rays(o.to(gpu), d.to(gpu))
    gt = torch.rand(n, 3, generator=gen).to(gpu)
    st, rgb, gd, gs, m = _call(lib, grid, rays, gt, noalloc=True)
    assert st == 1  # SVOX2_ERR_INVALID_ARGUMENT
    msg = lib.svox2_last_error().decode()
    assert "no-allocation" in msg and "8192" in msg, msg
    assert np.all(rgb == -7.0) and not gd.any() and not gs.any() and not m.any()
