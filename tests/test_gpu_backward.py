"""GPU gradients (fused MSE kernel and the autograd backward) against the FP64
oracle's autograd gradient and FP64 finite differences.

Tolerances: colours 2e-5 absolute; gradients against the FP64 oracle
max|g - g64| <= 2e-5 * max|g64| per tensor and |g - g64| <= 1e-2 * |g64| per
element above 1e-3 of the max (tests/conftest.py: float32 accumulation through
atomics in arbitrary order, so the results are not bit-reproducible); the
touched-row mask is bit-exact against the rows that are corners of a
contributing sample of the FP32 march.
"""
import numpy as np
import pytest
import torch

import oracle.oracle as O
import svox2_amd
from svox2_amd import scenes
from tests.conftest import assert_grad_close, grad_errors, host_grid

pytestmark = pytest.mark.gpu

OPT = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)


def _grid(sc, dev, dtype=torch.float32):
    g = scenes.to_sparse_grid(sc, device=dev, dtype=dtype)
    g.opt.sigma_thresh = OPT.sigma_thresh
    g.opt.stop_thresh = OPT.stop_thresh
    return g


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _fused(grid, o, d, gt, dev, **kw):
    rays = svox2_amd.Rays(o.to(dev), d.to(dev))
    rgb = grid.volume_render_fused(rays, gt.to(dev), **kw)
    return (rgb.cpu().numpy(), grid.density_data.grad.cpu().numpy().reshape(-1),
            grid.sh_data.grad.cpu().numpy(), grid.sparse_grad_indexer.cpu().numpy())


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_fused_dense64_vs_fp64(gpu, dense64, march, dtype):
    _, o, d, gt, sc = dense64
    grid = _grid(sc, gpu, dtype)
    g = host_grid(sc, dtype)
    rgb, gd, gs, mask = _fused(grid, torch.from_numpy(o), torch.from_numpy(d),
                               torch.from_numpy(gt), gpu)
    smp = march.samples(g, o, d, OPT)
    rgb64, gd64, gs64 = O.mse_grads_fp64(g, o, d, OPT, gt, samples=smp)
    print(f"dense64 fused {dtype}: colour {np.abs(rgb - rgb64).max():.3e}")
    assert np.abs(rgb - rgb64).max() < 2e-5
    assert_grad_close(gd, gd64, "grad_density")
    assert_grad_close(gs, gs64, "grad_sh")
    assert np.array_equal(mask, O.touched_rows(g, smp))


def test_fused_beta_and_sparsity(gpu, dense64, march):
    _, o, d, gt, sc = dense64
    grid = _grid(sc, gpu)
    g = host_grid(sc)
    kw = dict(beta_loss=0.3, sparsity_loss=0.05)
    rgb, gd, gs, _ = _fused(grid, torch.from_numpy(o[:1024]), torch.from_numpy(d[:1024]),
                            torch.from_numpy(gt[:1024]), gpu, **kw)
    rgb64, gd64, gs64 = O.mse_grads_fp64(g, o[:1024], d[:1024], OPT, gt[:1024], **kw)
    assert np.abs(rgb - rgb64).max() < 2e-5
    assert_grad_close(gd, gd64, "grad_density")
    assert_grad_close(gs, gs64, "grad_sh")


def test_autograd_backward_matches_oracle(gpu, dense64, march):
    _, o, d, gt, sc = dense64
    grid = _grid(sc, gpu)
    g = host_grid(sc)
    n = 2048
    rays = svox2_amd.Rays(torch.from_numpy(o[:n]).to(gpu), torch.from_numpy(d[:n]).to(gpu))
    rgb = grid.volume_render(rays)
    loss = ((rgb - torch.from_numpy(gt[:n]).to(gpu)) ** 2).mean()
    loss.backward()
    gd = grid.density_data.grad.cpu().numpy().reshape(-1)
    gs = grid.sh_data.grad.cpu().numpy()
    rgb64, gd64, gs64 = O.mse_grads_fp64(g, o[:n], d[:n], OPT, gt[:n])
    assert np.abs(rgb.detach().cpu().numpy() - rgb64).max() < 2e-5
    assert_grad_close(gd, gd64, "grad_density")
    assert_grad_close(gs, gs64, "grad_sh")


def test_finite_differences_with_ray_clip(gpu, march):
    """Per-ray clipped rays on a 64^3 instance: the GPU gradient (autograd
    Function, transmittance recomputed in the backward) against FP64 central
    differences (h = 1e-4) of the oracle loss on a random 1% of rows, along
    random directions supported on that subset and per element."""
    gen = torch.Generator().manual_seed(11)
    sc = scenes.make_blob_grid(64, scenes.DENSE64, gen, dense=True)
    o, d = scenes.sphere_rays(2048, gen)
    gt = torch.rand(2048, 3, generator=gen, dtype=torch.float64).float()
    clip = scenes.config_clip(2048, gen) * 0.6
    grid = _grid(sc, gpu)
    rays = svox2_amd.Rays(o.to(gpu), d.to(gpu), clip.to(gpu))
    loss = ((grid.volume_render(rays) - gt.to(gpu)) ** 2).mean()
    loss.backward()
    gd = grid.density_data.grad.cpu().double().numpy().reshape(-1)
    gs = grid.sh_data.grad.cpu().double().numpy()
    g = host_grid(sc)
    smp = march.samples(g, o.numpy(), d.numpy(), OPT, clip=clip.numpy())
    R = g.density.shape[0]
    rng = np.random.default_rng(0)
    subset = rng.choice(R, R // 100, replace=False)
    # Directions move only rows whose density stays far above sigma_thresh under
    # the perturbation: near-zero densities sit on the sigma_thresh kink, where
    # the loss has no derivative.
    smooth = subset[g.density[subset] > 1e-2]

    def L(dens, sh):
        return float(((O.render_fp64(g, o.numpy(), d.numpy(), OPT, torch.from_numpy(dens),
                                     torch.from_numpy(sh), samples=smp)
                       - gt.double()) ** 2).mean())

    h = 1e-4
    for k in range(4):
        vd = np.zeros(R)
        vd[smooth] = rng.standard_normal(smooth.size)
        vs = np.zeros_like(g.sh)
        vs[smooth] = rng.standard_normal((smooth.size, g.sh.shape[1]))
        fd = (L(g.density + h * vd, g.sh + h * vs) - L(g.density - h * vd, g.sh - h * vs)) / (2 * h)
        an = float(gd @ vd + (gs * vs).sum())
        assert an == pytest.approx(fd, rel=2e-3, abs=1e-9), k
    top = smooth[np.argsort(-np.abs(gd[smooth]))[:4]]
    for r in top:
        e = np.zeros(R)
        e[r] = 1
        fd = (L(g.density + h * e, g.sh) - L(g.density - h * e, g.sh)) / (2 * h)
        assert gd[r] == pytest.approx(fd, rel=2e-3, abs=1e-9)


def test_training_step_facade(gpu):
    """One training step of the reference loop (opt/opt.py:492-603): fused
    render, TV on density and SH, RMSprop on both; gradients are zeroed and only
    touched rows move."""
    gen = torch.Generator().manual_seed(2)
    sc = scenes.make_blob_grid(64, scenes.SPARSE256, gen, device=gpu)
    cams = scenes.CameraSet(scenes.hemisphere_cameras(10, gen))
    o, d = cams.random_batch(5000, gen)
    gt = torch.rand(5000, 3, generator=gen)
    grid = _grid(sc, gpu)
    before_d = grid.density_data.data.clone()
    before_s = grid.sh_data.data.clone()
    rays = svox2_amd.Rays(o.to(gpu), d.to(gpu))
    grid.volume_render_fused(rays, gt.to(gpu))
    np.random.seed(0)
    grid.inplace_tv_grad(grid.density_data.grad, scaling=1e-5, sparse_frac=0.01)
    grid.inplace_tv_color_grad(grid.sh_data.grad, scaling=1e-3, sparse_frac=0.01)
    touched = grid.sparse_grad_indexer.clone()
    grid.optim_density_step(30.0, beta=0.95)
    grid.optim_sh_step(1e-2, beta=0.95)
    assert torch.count_nonzero(grid.density_data.grad) == 0
    assert torch.count_nonzero(grid.sh_data.grad) == 0
    moved = (grid.density_data.data != before_d).flatten()
    assert moved.any() and not (moved & ~touched).any()
    moved_s = (grid.sh_data.data != before_s).any(1)
    assert moved_s.any() and not (moved_s & ~grid.sparse_sh_grad_indexer).any()


@pytest.mark.parametrize("size,n_rays", [(64, 2000), (256, 5000)])
def test_training_step_values(gpu, march, size, n_rays):
    """The training step's values against the oracle's composition of the
    reference loop (opt/opt.py:492-603): on a 64^3 sparse grid, and at the
    size of the 256^3 training configuration (about 8% occupied, 5000 rays
    from 100 cameras, 1% of the cells per TV term).

    - The fused gradients plus both TV terms (the same contiguous cell runs,
      drawn from np.random as the reference does) are compared with the FP64
      fused gradients plus the oracle's TV, to the gradient tolerances.
    - The touched-row masks are compared with the fused mask or-ed with the
      TV masks.
    - Then one RMSprop step per quantity (lr 30 and 1e-2, beta 0.95, fresh
      state) is compared with the oracle's closed form applied to the GPU's
      own gradients on those rows."""
    if size == 64:
        gen = torch.Generator().manual_seed(4)
        sc = scenes.make_blob_grid(64, scenes.SPARSE256, gen, device=gpu)
        cams = scenes.CameraSet(scenes.hemisphere_cameras(10, gen))
    else:
        sc, cams, gen = scenes.workload_blobs(256, scenes.SPARSE256, seed=0, device=gpu)
    o, d = cams.random_batch(n_rays, gen)
    gt = torch.rand(n_rays, 3, generator=gen)
    grid = _grid(sc, gpu)
    links = grid.links.cpu().numpy()
    g = host_grid(sc, links=links)
    dens0 = grid.density_data.data.cpu().double().numpy().copy()
    sh0 = grid.sh_data.data.cpu().double().numpy().copy()
    grid.volume_render_fused(svox2_amd.Rays(o.to(gpu), d.to(gpu)), gt.to(gpu))
    np.random.seed(3)
    grid.inplace_tv_grad(grid.density_data.grad, scaling=1e-5, sparse_frac=0.01)
    grid.inplace_tv_color_grad(grid.sh_data.grad, scaling=1e-3, sparse_frac=0.01)
    gd = grid.density_data.grad.cpu().double().numpy()
    gs = grid.sh_data.grad.cpu().double().numpy()
    mask_d = grid.sparse_grad_indexer.cpu().numpy().copy()
    mask_s = grid.sparse_sh_grad_indexer.cpu().numpy().copy()

    cells_total = links.size
    num = max(int(0.01 * cells_total), 1)
    np.random.seed(3)
    start_d = np.random.randint(0, cells_total)
    start_s = np.random.randint(0, cells_total)
    run = lambda st: (np.arange(st, st + num) % cells_total).astype(np.int64)
    smp = march.samples(g, o.numpy(), d.numpy(), OPT)
    _, gd64, gs64 = O.mse_grads_fp64(g, o.numpy(), d.numpy(), OPT, gt.numpy(), samples=smp)
    tv_d, m_d = O.tv_grad_sparse(links, np.asarray(g.density).reshape(-1, 1), run(start_d), 0, 1,
                                 1e-5, False, False)
    tv_s, m_s = O.tv_grad_sparse(links, g.sh, run(start_s), 0, g.sh.shape[1], 1e-3, True, False)
    assert_grad_close(gd, gd64.reshape(-1, 1) + tv_d, "grad_density + TV")
    assert_grad_close(gs, gs64 + tv_s, "grad_sh + TV")
    fused_mask = O.touched_rows(g, smp)
    assert np.array_equal(mask_d, fused_mask | m_d)
    assert np.array_equal(mask_s, fused_mask | m_s)

    grid.optim_density_step(30.0, beta=0.95)
    grid.optim_sh_step(1e-2, beta=0.95)
    for data0, grad, mask, lr, new in (
            (dens0, gd, mask_d, 30.0, grid.density_data.data),
            (sh0, gs, mask_s, 1e-2, grid.sh_data.data)):
        want, _, _ = O.rmsprop_step(data0, np.zeros_like(data0), grad, np.nonzero(mask)[0],
                                    0.95, lr, 1e-8, -1e9, lr)
        got = new.cpu().double().numpy()
        assert np.abs(got - want).max() <= 1e-5 * max(1.0, np.abs(want).max())


def test_fused_ray_order(gpu, dense64):
    """The fused call's ray orders on a 65536-ray batch (above the ordering
    threshold of an MI355X, 40960 rays): with a caller-owned workspace the
    backward's order is a permutation of the rays sorted by descending chunk
    count (its histogram is the workspace's second), and colours and
    gradients match the input-order call."""
    import ctypes
    from svox2_amd import _C
    _, _, _, _, sc = dense64
    grid = _grid(sc, gpu)
    gen = torch.Generator().manual_seed(7)
    o, d = scenes.sphere_rays(65536, gen)
    rays = svox2_amd.Rays(o.to(gpu), d.to(gpu))
    n = rays.origins.size(0)
    gt_t = torch.rand(n, 3, generator=gen).to(gpu)
    lib = _C.load_library()
    nbytes = lib.svox2_ray_order_workspace_bytes(n)
    # [forward-order histogram | backward-order histogram | cursors | cost | order]
    assert nbytes >= 4 * (3 * 1024 + 2 * n)
    ws = torch.zeros(nbytes // 4, dtype=torch.int32, device=gpu)

    def run(workspace):
        gd, gs, _, _ = grid._get_data_grads()
        gd.zero_()
        gs.zero_()
        holder = _C.GridOutputGrads()
        holder.grad_density_out, holder.grad_sh_out = gd, gs
        spec = grid._to_cpp()
        g = _C._grid_struct(spec)
        oo, dd, clip = _C._rays(rays._to_cpp())
        gstruct = _C._grads_struct(holder, g.rows, g.sh_data_dim)
        out = torch.empty(n, 3, device=gpu)
        st = lib.svox2_volume_render_cuvol_fused(
            ctypes.byref(g), _C._ptr(oo), _C._ptr(dd), _C._ptr(clip), n,
            ctypes.byref(_C._opt_struct(grid.opt._to_cpp())), _C._ptr(gt_t), 0.0, 0.0,
            _C._ptr(out), None, ctypes.byref(gstruct),
            _C._ptr(workspace) if workspace is not None else None,
            nbytes if workspace is not None else 0, _C._stream(oo))
        assert st == 0, lib.svox2_last_error()
        torch.cuda.synchronize()
        return out.cpu().numpy(), gd.cpu().numpy().copy(), gs.cpu().numpy().copy()

    rgb_a, gd_a, gs_a = run(ws)
    rgb_b, gd_b, gs_b = run(None)
    # The workspace the call asks for at this size adds the chunk log: the
    # backward evaluates the forward's logged chunk placements.
    nbytes_log = lib.svox2_volume_render_cuvol_fused_workspace_bytes(n)
    assert nbytes_log > nbytes
    ws_log = torch.zeros(nbytes_log // 4, dtype=torch.int32, device=gpu)
    nbytes, nbytes_ord = nbytes_log, nbytes
    rgb_c, gd_c, gs_c = run(ws_log)
    nbytes = nbytes_ord
    assert np.array_equal(rgb_c, rgb_b)
    assert _rel(gd_c, gd_b) < 1e-5 and _rel(gs_c, gs_b) < 1e-5
    w = ws.cpu().numpy().view(np.uint32)
    hist, cost, order = w[1024:2048], w[3072:3072 + n], w[3072 + n:3072 + 2 * n]
    assert int(hist.sum()) == n
    assert cost.max() > cost.min()  # rays of different lengths
    assert np.array_equal(np.sort(order), np.arange(n, dtype=np.uint32))
    assert np.all(np.diff(cost[order].astype(np.int64)) <= 0)
    assert np.array_equal(np.bincount(cost, minlength=1024), hist)
    assert np.array_equal(rgb_a, rgb_b)
    assert _rel(gd_a, gd_b) < 1e-5 and _rel(gs_a, gs_b) < 1e-5


@pytest.mark.parametrize("which", ["dense64", "sparse_long"])
def test_fused_record_replay_matches_remarch(gpu, dense64, which):
    """The fused call with the record stream (the backward replays the
    forward's samples) against the re-marching backward: colours bit-exact,
    gradients within 1e-5 of each other (atomic order differs), masks equal.
    On the dense 64^3 grid most rays hold far more contributing samples than
    the per-ray share of the stream, so late rays overflow it and are
    re-marched: both paths run in one launch."""
    from svox2_amd import _C
    _, o, d, gt, sc = dense64
    grid = _grid(sc, gpu)
    if which == "dense64":
        rays = svox2_amd.Rays(torch.from_numpy(o).to(gpu), torch.from_numpy(d).to(gpu))
        gt_t = torch.from_numpy(gt).to(gpu)
    else:
        # Below the ordering threshold (40960 rays on 256 CUs): records on.
        gen = torch.Generator().manual_seed(3)
        oo, dd = scenes.sphere_rays(32768, gen)
        rays = svox2_amd.Rays(oo.to(gpu), dd.to(gpu))
        gt_t = torch.rand(32768, 3, generator=gen).to(gpu)
    out = {}
    for record in (True, False):
        gd, gs, _, _ = grid._get_data_grads()
        gd.zero_()
        gs.zero_()
        holder = _C.GridOutputGrads()
        holder.grad_density_out, holder.grad_sh_out = gd, gs
        holder.mask_out = torch.zeros(grid.density_data.size(0), dtype=torch.bool, device=gpu)
        rgb = torch.empty_like(gt_t)
        _C.volume_render_cuvol_fused(grid._to_cpp(), rays._to_cpp(), grid.opt._to_cpp(), gt_t,
                                     0.0, 0.0, rgb, holder, ordered=True, record=record)
        torch.cuda.synchronize()
        out[record] = (rgb.cpu().numpy(), gd.cpu().numpy().copy(), gs.cpu().numpy().copy(),
                       holder.mask_out.cpu().numpy())
    a, b = out[True], out[False]
    assert np.array_equal(a[0], b[0])
    assert _rel(a[1], b[1]) < 1e-5 and _rel(a[2], b[2]) < 1e-5
    assert np.array_equal(a[3], b[3])
