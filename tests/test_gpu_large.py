"""The 512^3 (about 3% occupied, SH-2) grid of the benchmark: bit-exact sample
indexing, and colours and gradients against the FP64 oracle, on the full
65,536-ray batch, for float32 and bfloat16 storage.  Tolerances as in the
smaller tests (colours 2e-5 vs the oracle on the stored values, 2e-3
including BF16 storage rounding; gradients 2e-5 of their max and 1e-2 per
element above 1e-3 of the max, tests/conftest.py)."""
import time

import numpy as np
import pytest
import torch

import oracle.oracle as O
import svox2_amd
from svox2_amd import _C, scenes
from tests.conftest import assert_grad_close, host_grid

pytestmark = pytest.mark.gpu

OPT = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)


@pytest.fixture(scope="module")
def big(gpu):
    t = time.time()
    sc, cams, gen = scenes.workload_blobs(512, scenes.SPARSE512, seed=0, device=gpu)
    grid = scenes.to_sparse_grid(sc, device=gpu)
    grid.opt.sigma_thresh, grid.opt.stop_thresh = OPT.sigma_thresh, OPT.stop_thresh
    links = grid.links.cpu().numpy()
    o, d = cams.random_batch(65536, gen)
    gt = torch.rand(65536, 3, generator=gen)
    print(f"512^3 grid: {sc.density.shape[0]} rows ({sc.occupancy:.2%}), built in "
          f"{time.time() - t:.1f}s")
    return sc, grid, links, o, d, gt


def test_512_stats_bit_exact(gpu, big, march):
    sc, grid, links, o, d, gt = big
    rays = svox2_amd.Rays(o.to(gpu), d.to(gpu))
    st = _C.volume_render_cuvol_stats(grid._to_cpp(), rays._to_cpp(),
                                      grid.opt._to_cpp()).cpu().numpy()
    g = host_grid(sc, links=links)
    t = time.time()
    _, _, st_ref = march.render(g, o.numpy(), d.numpy(), OPT)
    print(f"C restatement: {time.time() - t:.1f}s; I={st[:, 0].sum()} E={st[:, 1].sum()} "
          f"H={st[:, 2].sum()}")
    assert np.array_equal(st, st_ref)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_512_colours_and_grads(gpu, big, march, dtype):
    sc, grid32, links, o, d, gt = big
    grid = grid32 if dtype == torch.float32 else scenes.to_sparse_grid(sc, device=gpu,
                                                                       dtype=dtype)
    grid.opt.sigma_thresh, grid.opt.stop_thresh = OPT.sigma_thresh, OPT.stop_thresh
    n = 65536
    rays = svox2_amd.Rays(o[:n].to(gpu), d[:n].to(gpu))
    grid.density_data.grad = None
    grid.sh_data.grad = None
    rgb = grid.volume_render_fused(rays, gt[:n].to(gpu)).cpu().numpy()
    gd = grid.density_data.grad.cpu().numpy().reshape(-1)
    gs = grid.sh_data.grad.cpu().numpy()
    g = host_grid(sc, dtype, links=links)
    rgb64, gd64, gs64 = O.mse_grads_fp64(g, o[:n].numpy(), d[:n].numpy(), OPT, gt[:n].numpy())
    ec = np.abs(rgb - rgb64).max()
    print(f"512^3 {dtype}: colour {ec:.3e}")
    assert ec < 2e-5
    assert_grad_close(gd, gd64, f"512^3 {dtype} grad_density")
    assert_grad_close(gs, gs64, f"512^3 {dtype} grad_sh")
    if dtype == torch.bfloat16:
        g32 = host_grid(sc, torch.float64, links=links)
        rgb_s = O.render_fp64(g32, o[:n].numpy(), d[:n].numpy(), OPT,
                              torch.from_numpy(g32.density), torch.from_numpy(g32.sh)).numpy()
        es_ = np.abs(rgb - rgb_s).max()
        print(f"512^3 bf16 colour error including storage rounding: {es_:.3e}")
        assert es_ < 2e-3


def test_512_chunk_log(gpu, big):
    """The fused call on a 131,072-ray batch (the 65,536 rays twice, above the
    batch sizes that record their samples) with the workspace it asks for
    (ray order plus the chunk log: the backward evaluates the forward's logged
    chunk placements, the first 16 chunks of each ray, and places the rest
    itself) against the call without a workspace (input order, every chunk
    placed by the backward): colours bit-exact, gradients within 1e-5 of
    each other (atomic order differs), masks equal.  Each ray's log head
    holds min(chunks, 16) and whether the ray went on, its chunk count taken
    from the plain forward.  Rays with more than 16 chunks occur on this
    batch, so both kinds of chunk run in one backward."""
    import ctypes
    sc, grid, links, o, d, gt = big
    lib = _C.load_library()
    n = 2 * 65536
    rays = svox2_amd.Rays(torch.cat([o, o]).to(gpu), torch.cat([d, d]).to(gpu))
    gt_t = torch.cat([gt, gt]).to(gpu)
    nbytes = lib.svox2_volume_render_cuvol_fused_workspace_bytes(n)
    order_bytes = lib.svox2_ray_order_workspace_bytes(n)
    rec_bytes = (n * 16 * 36 * 4 + 255) // 256 * 256
    assert nbytes == order_bytes + rec_bytes + (16 * n + 255) // 256 * 256
    ws = torch.zeros(nbytes // 4, dtype=torch.int32, device=gpu)

    def run(workspace):
        gd, gs, _, _ = grid._get_data_grads()
        gd.zero_()
        gs.zero_()
        mask = torch.zeros(gd.size(0), dtype=torch.bool, device=gpu)
        holder = _C.GridOutputGrads()
        holder.grad_density_out, holder.grad_sh_out, holder.mask_out = gd, gs, mask
        g = _C._grid_struct(grid._to_cpp())
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
        return (out.cpu().numpy(), gd.cpu().numpy().copy(), gs.cpu().numpy().copy(),
                mask.cpu().numpy().copy())

    rgb_a, gd_a, gs_a, m_a = run(ws)
    rgb_b, gd_b, gs_b, m_b = run(None)
    heads = ws.cpu().numpy()[(order_bytes + rec_bytes) // 4:].reshape(-1, 4)[:n]
    chunks = torch.empty(n, dtype=torch.int32, device=gpu)
    _C.volume_render_cuvol(grid._to_cpp(), rays._to_cpp(), grid.opt._to_cpp(),
                           ray_cost_out=chunks)
    chunks = chunks.cpu().numpy()
    assert chunks.max() < 1023  # below the cost clamp: the counts are exact
    assert np.array_equal(heads[:, 0], np.minimum(chunks, 16))
    assert np.array_equal(heads[:, 1], (chunks > 16).astype(np.int32))
    assert np.any(chunks > 16)  # rays continued past their logged chunks
    assert np.array_equal(rgb_a, rgb_b)
    rel = lambda a, b: np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
    print(f"chunk log: {int((heads[:, 1] == 1).sum())} rays past 16 chunks (max "
          f"{chunks.max()}); grad rel "
          f"{rel(gd_a, gd_b):.2e} / {rel(gs_a, gs_b):.2e}")
    assert rel(gd_a, gd_b) < 1e-5 and rel(gs_a, gs_b) < 1e-5
    assert np.array_equal(m_a, m_b)


def test_512_ordered_forward(gpu, big):
    """The forward with an ordering workspace (svox2_volume_render_cuvol_ordered,
    used by the autograd forward) on a full 65,536-ray batch, with per-ray
    clips: rays run in descending order of segment length, and every ray's
    colour, final log transmittance and chunk count equal those of the
    input-order forward, bit for bit."""
    sc, grid, links, o, d, gt = big
    n = 65536
    gen = torch.Generator().manual_seed(11)
    clip = torch.stack([torch.rand(n, generator=gen) * 0.5, 2 + 2 * torch.rand(n, generator=gen)], 1)
    rays = svox2_amd.Rays(o.to(gpu), d.to(gpu), clip=clip.to(gpu))
    out = {}
    for ordered in (False, True):
        lt = torch.empty(n, device=gpu)
        cost = torch.empty(n, dtype=torch.int32, device=gpu)
        rgb = _C.volume_render_cuvol(grid._to_cpp(), rays._to_cpp(), grid.opt._to_cpp(),
                                     log_transmit_out=lt, ray_cost_out=cost, ordered=ordered)
        torch.cuda.synchronize()
        out[ordered] = (rgb.cpu().numpy(), lt.cpu().numpy(), cost.cpu().numpy())
    for a, b in zip(out[False], out[True]):
        assert np.array_equal(a, b)
    assert out[True][2].max() > out[True][2].min()


def test_512_fused_matches_autograd(gpu, big):
    """The fused call on a full 65,536-ray batch (it records its samples,
    forms their colours chunk-parallel and replays the records, with the
    counters split over 64 partitions) against the autograd path on the same
    rays (volume_render, MSE loss, backward; it re-marches every ray):
    colours bit-exact and gradients within 1e-5 of each other (atomic order
    differs)."""
    sc, grid, links, o, d, gt = big
    n = 65536
    rays = svox2_amd.Rays(o.to(gpu), d.to(gpu))
    gt_t = gt.to(gpu)
    grid.density_data.grad = None
    grid.sh_data.grad = None
    rgb_f = grid.volume_render_fused(rays, gt_t).detach().cpu().numpy()
    gd_f = grid.density_data.grad.cpu().numpy().copy()
    gs_f = grid.sh_data.grad.cpu().numpy().copy()
    grid.density_data.grad = None
    grid.sh_data.grad = None
    rgb = grid.volume_render(rays)
    torch.nn.functional.mse_loss(rgb, gt_t).backward()
    rgb_a = rgb.detach().cpu().numpy()
    gd_a = grid.density_data.grad.cpu().numpy()
    gs_a = grid.sh_data.grad.cpu().numpy()
    grid.density_data.grad = None
    grid.sh_data.grad = None
    rel = lambda a, b: np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
    print(f"fused vs autograd: grad rel {rel(gd_f, gd_a):.2e} / {rel(gs_f, gs_a):.2e}")
    assert np.array_equal(rgb_f, rgb_a)
    assert rel(gd_f, gd_a) < 1e-5 and rel(gs_f, gs_a) < 1e-5


def test_512_autograd_clips_vs_fp64(gpu, big):
    """The million-ray path (SparseGrid.volume_render with per-ray clips
    t_near ~ U[0, 0.5], t_far ~ U[2, 4]; its backward re-marches every ray
    and recomputes the transmittance) on a full 65,536-ray batch against the
    FP64 oracle with the same clips: colours within 2e-5, gradients within
    the gradient tolerances."""
    sc, grid, links, o, d, gt = big
    n = 65536
    gen = torch.Generator().manual_seed(5)
    clip = scenes.config_clip(n, gen)
    grid.density_data.grad = None
    grid.sh_data.grad = None
    rgb = grid.volume_render(svox2_amd.Rays(o.to(gpu), d.to(gpu), clip.to(gpu)))
    torch.nn.functional.mse_loss(rgb, gt.to(gpu)).backward()
    rgb = rgb.detach().cpu().numpy()
    gd = grid.density_data.grad.cpu().numpy().reshape(-1)
    gs = grid.sh_data.grad.cpu().numpy()
    grid.density_data.grad = None
    grid.sh_data.grad = None
    g = host_grid(sc, links=links)
    rgb64, gd64, gs64 = O.mse_grads_fp64(g, o.numpy(), d.numpy(), OPT, gt.numpy(),
                                         clip=clip.numpy())
    ec = np.abs(rgb - rgb64).max()
    print(f"512^3 clipped autograd: colour {ec:.3e}")
    assert ec < 2e-5
    assert_grad_close(gd, gd64, "512^3 clipped grad_density")
    assert_grad_close(gs, gs64, "512^3 clipped grad_sh")


def test_1m_autograd_slices(gpu, big):
    """A 1,048,576-ray clipped batch (rays of the 65,536-ray batch drawn with
    replacement, origins and directions jittered by up to 5e-4) through the
    autograd path with a summed squared error, which is linear in the rays:
    its colours equal those of its 16 slices of 65,536 rays bit for bit, and
    its gradient equals the sum of the slices' gradients within 1e-5 of its
    max (atomic order differs)."""
    sc, grid, links, o, d, gt = big
    n, k = 1 << 20, 16
    gen = torch.Generator().manual_seed(77)
    gi = torch.randint(0, o.shape[0], (n,), generator=gen)
    jitter = (torch.rand(n, 3, generator=gen) - 0.5) * 1e-3
    oo = (o[gi] + jitter).to(gpu)
    dd = torch.nn.functional.normalize(d[gi] + jitter, dim=1).to(gpu)
    clip = scenes.config_clip(n, gen).to(gpu)
    tgt = torch.rand(n, 3, generator=gen).to(gpu)

    def grads(sl):
        grid.density_data.grad = None
        grid.sh_data.grad = None
        rgb = grid.volume_render(svox2_amd.Rays(oo[sl], dd[sl], clip[sl]))
        ((rgb - tgt[sl]) ** 2).sum().backward()
        return rgb.detach(), grid.density_data.grad.clone(), grid.sh_data.grad.clone()

    rgb_all, gd_all, gs_all = grads(slice(0, n))
    gd_sum = torch.zeros_like(gd_all)
    gs_sum = torch.zeros_like(gs_all)
    for i in range(k):
        sl = slice(i * n // k, (i + 1) * n // k)
        rgb, gd, gs = grads(sl)
        assert torch.equal(rgb, rgb_all[sl])
        gd_sum += gd
        gs_sum += gs
    grid.density_data.grad = None
    grid.sh_data.grad = None
    rel = lambda a, b: float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))
    print(f"1M rays: grad rel vs sum of slices {rel(gd_all, gd_sum):.2e} / "
          f"{rel(gs_all, gs_sum):.2e}")
    assert rel(gd_all, gd_sum) < 1e-5 and rel(gs_all, gs_sum) < 1e-5


def test_262144_fused_matches_autograd(gpu, big):
    """A 262,144-ray batch (rays of the 65,536-ray batch drawn with
    replacement and jittered by up to 5e-4): above 4 device fills the fused
    call's backward is the 4-wave build, without chunk prefetch.  Against the
    autograd path on the same rays: colours bit-exact, gradients within 1e-5
    of each other (atomic order differs)."""
    sc, grid, links, o, d, gt = big
    n = 1 << 18
    gen = torch.Generator().manual_seed(21)
    gi = torch.randint(0, o.shape[0], (n,), generator=gen)
    jitter = (torch.rand(n, 3, generator=gen) - 0.5) * 1e-3
    rays = svox2_amd.Rays((o[gi] + jitter).to(gpu),
                          torch.nn.functional.normalize(d[gi] + jitter, dim=1).to(gpu))
    tgt = torch.rand(n, 3, generator=gen).to(gpu)
    grid.density_data.grad = None
    grid.sh_data.grad = None
    rgb_f = grid.volume_render_fused(rays, tgt).detach()
    gd_f = grid.density_data.grad.clone()
    gs_f = grid.sh_data.grad.clone()
    grid.density_data.grad = None
    grid.sh_data.grad = None
    rgb = grid.volume_render(rays)
    torch.nn.functional.mse_loss(rgb, tgt).backward()
    rel = lambda a, b: float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))
    print(f"262144 rays fused vs autograd: grad rel {rel(gd_f, grid.density_data.grad):.2e} / "
          f"{rel(gs_f, grid.sh_data.grad):.2e}")
    assert torch.equal(rgb_f, rgb.detach())
    assert rel(gd_f, grid.density_data.grad) < 1e-5 and rel(gs_f, grid.sh_data.grad) < 1e-5
    grid.density_data.grad = None
    grid.sh_data.grad = None


def test_512_image_800(gpu, big, march):
    """The 800x800 image of the benchmark's 512^3 grid (float32 and bf16
    storage): bit-exact against the ray kernel on the oracle's float32 camera
    rays of the same camera, and the colours of 4096 random pixels within
    2e-5 of the FP64 oracle on the stored values."""
    sc, grid32, links, o, d, gt = big
    cams = scenes.CameraSet(scenes.hemisphere_cameras(3, torch.Generator().manual_seed(9)))
    cam = cams.camera(1)
    ro, rd = march.camera_rays(cam.c2w.numpy(), cam.fx_val, cam.fy_val, cam.cx_val, cam.cy_val,
                               800, 800)
    sel = np.random.default_rng(0).choice(800 * 800, 4096, replace=False)
    for dtype in (torch.float32, torch.bfloat16):
        grid = grid32 if dtype == torch.float32 else scenes.to_sparse_grid(sc, device=gpu,
                                                                           dtype=dtype)
        grid.opt.sigma_thresh, grid.opt.stop_thresh = OPT.sigma_thresh, OPT.stop_thresh
        with torch.no_grad():
            img = grid.volume_render_image(cam).reshape(-1, 3)
            rgb = grid.volume_render(svox2_amd.Rays(torch.from_numpy(ro).to(gpu),
                                                    torch.from_numpy(rd).to(gpu)))
        assert torch.equal(img, rgb), dtype
        g = host_grid(sc, dtype, links=links)
        rgb64 = O.render_fp64(g, ro[sel], rd[sel], OPT, torch.from_numpy(g.density),
                              torch.from_numpy(g.sh)).numpy()
        err = np.abs(img.cpu().numpy()[sel] - rgb64).max()
        print(f"800x800 {dtype}: colour error on 4096 pixels {err:.3e}")
        assert err < 2e-5


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_512_fused_5000(gpu, big, dtype):
    """The reference's training batch size on the 512^3 grid: 5000 rays
    (below the ordering threshold, so the forward records every contributing
    sample and the backward replays the records) against the FP64 oracle."""
    sc, grid32, links, o, d, gt = big
    grid = grid32 if dtype == torch.float32 else scenes.to_sparse_grid(sc, device=gpu,
                                                                       dtype=dtype)
    grid.opt.sigma_thresh, grid.opt.stop_thresh = OPT.sigma_thresh, OPT.stop_thresh
    n = 5000
    grid.density_data.grad = None
    grid.sh_data.grad = None
    rgb = grid.volume_render_fused(svox2_amd.Rays(o[:n].to(gpu), d[:n].to(gpu)),
                                   gt[:n].to(gpu)).cpu().numpy()
    g = host_grid(sc, dtype, links=links)
    rgb64, gd64, gs64 = O.mse_grads_fp64(g, o[:n].numpy(), d[:n].numpy(), OPT, gt[:n].numpy())
    assert np.abs(rgb - rgb64).max() < 2e-5
    assert_grad_close(grid.density_data.grad.cpu().numpy().reshape(-1), gd64,
                      f"5000 rays {dtype} grad_density")
    assert_grad_close(grid.sh_data.grad.cpu().numpy(), gs64, f"5000 rays {dtype} grad_sh")
    grid.density_data.grad = None
    grid.sh_data.grad = None


def test_serving_mix_at_size(gpu, big):
    """The serving configuration at its real size: four resident grids (two
    256^3 about 8% occupied and two 512^3 about 3%, one of each in bf16),
    image sides 128, 256, 512, 800 and 1024, the benchmark's server settings
    (4096-pixel tiles, 128 tiles per launch, graph replay): every request's
    image equals its lone volume_render_image bit for bit."""
    from svox2_amd.serve import RenderRequest, RenderServer
    sc512, grid512, _, _, _, _ = big
    sc256, cams256, _ = scenes.workload_blobs(256, scenes.SPARSE256, seed=1, device=gpu)
    grids = []
    for sc, dt in ((sc256, torch.float32), (sc256, torch.bfloat16), (sc512, torch.float32),
                   (sc512, torch.bfloat16)):
        g = grid512 if (sc is sc512 and dt == torch.float32) else \
            scenes.to_sparse_grid(sc, device=gpu, dtype=dt)
        g.opt.sigma_thresh, g.opt.stop_thresh = OPT.sigma_thresh, OPT.stop_thresh
        grids.append(g)
    cam_sets = [cams256, cams256] + [scenes.CameraSet(
        scenes.hemisphere_cameras(8, torch.Generator().manual_seed(13)))] * 2
    rng = np.random.default_rng(2)
    reqs = []
    for side in (128, 256, 512, 800, 1024, 128, 512, 1024, 256, 800):
        gi = int(rng.integers(4))
        cams = cam_sets[gi]
        cam = cams.camera(int(rng.integers(cams.c2w.shape[0])), width=side, height=side)
        reqs.append(RenderRequest(grid=gi, camera=cam, arrival=0.0))
    server = RenderServer(grids, hbm_budget=1 << 30, tile_pixels=4096, max_tiles=128,
                          max_active=64, policy="fifo", depth=3)
    stats = server.run(reqs, realtime=False)
    assert stats["requests"] == len(reqs)
    for r in reqs:
        with torch.no_grad():
            alone = grids[r.grid].volume_render_image(r.camera)
        assert torch.equal(r.image, alone), (r.grid, r.camera.width)
