"""GPU forward march against the oracle.

Tolerances: colours within 2e-5 (absolute, colours are O(1)) of the FP64
oracle evaluated on the same stored values; sample indexing (iterations,
evaluated samples, contributing samples, and a hash of every visited cell, t
and contribution flag plus the final log transmittance) bit-exact against the
FP32 C restatement.  BF16 storage is also compared with the FP64 oracle of the
unrounded float32 grid, within 2e-3.
"""
import numpy as np
import pytest
import torch

import oracle.oracle as O
import svox2_amd
from svox2_amd import _C, scenes
from tests.conftest import host_grid

pytestmark = pytest.mark.gpu

OPT = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)


def _grid(sc, dev, dtype=torch.float32, accelerate=True):
    g = scenes.to_sparse_grid(sc, device=dev, dtype=dtype, accelerate=accelerate)
    g.opt.sigma_thresh = OPT.sigma_thresh
    g.opt.stop_thresh = OPT.stop_thresh
    return g


def _render(grid, o, d, dev, clip=None):
    rays = svox2_amd.Rays(o.to(dev), d.to(dev), None if clip is None else clip.to(dev))
    with torch.no_grad():
        return grid.volume_render(rays).cpu().numpy()


def _stats(grid, o, d, dev, clip=None):
    rays = svox2_amd.Rays(o.to(dev), d.to(dev), None if clip is None else clip.to(dev))
    return _C.volume_render_cuvol_stats(grid._to_cpp(), rays._to_cpp(),
                                        grid.opt._to_cpp()).cpu().numpy()


def _opts_of(grid):
    o = grid.opt
    return O.Opts(o.background_brightness, o.step_size, o.sigma_thresh, o.stop_thresh,
                  o.near_clip, o.use_spheric_clip, o.last_sample_opaque)


def test_dense64_forward_vs_fp64(gpu, dense64, march):
    g, o, d, gt, sc = dense64
    grid = _grid(sc, gpu)
    rgb = _render(grid, torch.from_numpy(o), torch.from_numpy(d), gpu)
    rgb64 = O.render_fp64(g, o, d, OPT, torch.from_numpy(g.density),
                          torch.from_numpy(g.sh)).numpy()
    err = np.abs(rgb - rgb64).max()
    print(f"dense64 fp32 colour max|err| vs FP64 = {err:.3e}")
    assert err < 2e-5


def test_dense64_stats_bit_exact(gpu, dense64, march):
    g, o, d, gt, sc = dense64
    grid = _grid(sc, gpu)
    st = _stats(grid, torch.from_numpy(o), torch.from_numpy(d), gpu)
    _, lt, st_ref = march.render(g, o, d, OPT)
    assert np.array_equal(st, st_ref)


def test_bf16_storage_forward(gpu, dense64, march):
    _, o, d, gt, sc = dense64
    grid = _grid(sc, gpu, torch.bfloat16)
    rgb = _render(grid, torch.from_numpy(o), torch.from_numpy(d), gpu)
    gb = host_grid(sc, torch.bfloat16)
    rgb64 = O.render_fp64(gb, o, d, OPT, torch.from_numpy(gb.density),
                          torch.from_numpy(gb.sh)).numpy()
    err_arith = np.abs(rgb - rgb64).max()
    g32 = host_grid(sc, torch.float64)
    rgb64s = O.render_fp64(g32, o, d, OPT, torch.from_numpy(g32.density),
                           torch.from_numpy(g32.sh)).numpy()
    err_storage = np.abs(rgb - rgb64s).max()
    print(f"dense64 bf16 colour max|err|: arithmetic {err_arith:.3e}, incl. storage "
          f"{err_storage:.3e}")
    assert err_arith < 2e-5
    assert err_storage < 2e-3
    st = _stats(grid, torch.from_numpy(o), torch.from_numpy(d), gpu)
    _, _, st_ref = march.render(gb, o, d, OPT)
    assert np.array_equal(st, st_ref)


def test_sparse256_accel_and_stats(gpu, sparse256, march):
    sc, grid, links_acc, o, d = sparse256
    assert np.array_equal(links_acc, O.accel_dist_prop(sc.links.cpu().numpy()))
    g = host_grid(sc, links=links_acc)
    st = _stats(grid, o, d, gpu)
    _, _, st_ref = march.render(g, o.numpy(), d.numpy(), OPT)
    assert np.array_equal(st, st_ref)
    assert st[:, 2].sum() > 0 and st[:, 0].sum() > st[:, 1].sum()  # hits and skips happen


def test_sparse256_colours_vs_fp64(gpu, sparse256):
    sc, grid, links_acc, o, d = sparse256
    g = host_grid(sc, links=links_acc)
    rgb = _render(grid, o[:1024], d[:1024], gpu)
    rgb64 = O.render_fp64(g, o[:1024].numpy(), d[:1024].numpy(), OPT,
                          torch.from_numpy(g.density), torch.from_numpy(g.sh)).numpy()
    err = np.abs(rgb - rgb64).max()
    print(f"sparse256 colour max|err| vs FP64 = {err:.3e}")
    assert err < 2e-5


def test_uniform_field_closed_form(gpu):
    N, sigma = 16, 3.0
    links = torch.arange(N ** 3, dtype=torch.int32).view(N, N, N)
    dens = torch.full((N ** 3, 1), sigma, dtype=torch.float64)
    sh = torch.zeros(N ** 3, 27, dtype=torch.float64)
    dc = [0.8, -0.2, 0.3]
    for c in range(3):
        sh[:, 9 * c] = dc[c]
    sc = scenes.SceneData(links, dens, sh, 9, 1.0)
    grid = _grid(sc, gpu, accelerate=False)
    grid.opt.stop_thresh = 0.0
    grid.opt.background_brightness = 0.7
    o = torch.tensor([[-2.0, 0.1, 0.2], [0.05, 0.1, 2.0]])
    d = torch.tensor([[1.0, 0.0, 0.0], [0.0, 0.0, -1.0]])
    rgb = _render(grid, o, d, gpu)
    ws = 0.5 / (N / 2)
    T = np.exp(-sigma * ws * (2 * N + 1))
    col = np.maximum(np.array(dc) * O.SH_C0 + 0.5, 0)
    assert np.abs(rgb - (col * (1 - T) + 0.7 * T)[None]).max() < 2e-6


def test_edge_cases(gpu, dense64, march):
    g, o, d, gt, sc = dense64
    grid = _grid(sc, gpu)
    # empty batch
    out = _render(grid, torch.zeros(0, 3), torch.zeros(0, 3), gpu)
    assert out.shape == (0, 3)
    # rays that miss the grid return exactly the background
    grid.opt.background_brightness = 0.25
    om = torch.tensor([[3.0, 3.0, 3.0], [0.0, 0.0, 5.0], [5.0, 0.0, 0.0]])
    dm = torch.tensor([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    assert np.all(_render(grid, om, dm, gpu) == np.float32(0.25))
    # option variants: bit-exact indexing and FP64 colours
    oo, dd = torch.from_numpy(o[:512]), torch.from_numpy(d[:512])
    variants = [dict(last_sample_opaque=True), dict(near_clip=0.4), dict(use_spheric_clip=True),
                dict(background_brightness=0.0), dict(step_size=0.3), dict(sigma_thresh=5.0)]
    for v in variants:
        grid.opt = svox2_amd.RenderOptions(sigma_thresh=1e-8, stop_thresh=1e-7)
        for k, val in v.items():
            setattr(grid.opt, k, val)
        opt = _opts_of(grid)
        st = _stats(grid, oo, dd, gpu)
        _, _, st_ref = march.render(g, oo.numpy(), dd.numpy(), opt)
        assert np.array_equal(st, st_ref), v
        rgb = _render(grid, oo, dd, gpu)
        rgb64 = O.render_fp64(g, oo.numpy(), dd.numpy(), opt, torch.from_numpy(g.density),
                              torch.from_numpy(g.sh)).numpy()
        assert np.abs(rgb - rgb64).max() < 2e-5, v
    # per-ray (t_near, t_far) clip
    grid.opt = svox2_amd.RenderOptions(sigma_thresh=1e-8, stop_thresh=1e-7)
    clip = scenes.config_clip(512, torch.Generator().manual_seed(3)) * 0.5
    st = _stats(grid, oo, dd, gpu, clip)
    _, _, st_ref = march.render(g, oo.numpy(), dd.numpy(), OPT, clip=clip.numpy())
    assert np.array_equal(st, st_ref)
    rgb = _render(grid, oo, dd, gpu, clip)
    rgb64 = O.render_fp64(g, oo.numpy(), dd.numpy(), OPT, torch.from_numpy(g.density),
                          torch.from_numpy(g.sh), clip=clip.numpy()).numpy()
    assert np.abs(rgb - rgb64).max() < 2e-5


def test_all_empty_grid(gpu):
    N = 8
    sc = scenes.SceneData(torch.full((N, N, N), -1, dtype=torch.int32),
                          torch.zeros(1, 1, dtype=torch.float64),
                          torch.zeros(1, 27, dtype=torch.float64), 9, 0.0)
    grid = _grid(sc, gpu)
    o, d = scenes.sphere_rays(64, torch.Generator().manual_seed(0))
    assert np.all(_render(grid, o, d, gpu) == np.float32(1.0))


def test_image_kernel_matches_ray_kernel(gpu, dense64, march):
    """The image kernel against (1) the ray kernel on the oracle's float32
    camera rays (the same arithmetic as cam2world_ray, so bit-exact), (2)
    the FP64 colours of those rays, and (3) the ray kernel on the rays of
    Camera.gen_rays (the reference's FP64 Python generator, svox2.py:157-183,
    within an ulp of (1)), to the colour tolerance."""
    g, o, d, gt, sc = dense64
    grid = _grid(sc, gpu)
    cams = scenes.CameraSet(scenes.hemisphere_cameras(2, torch.Generator().manual_seed(1)),
                            width=64, height=48, focal=80.0)
    cam = cams.camera(0)
    with torch.no_grad():
        img = grid.volume_render_image(cam).cpu().numpy()
    c2w = cam.c2w.numpy()
    ro, rd = march.camera_rays(c2w, cam.fx_val, cam.fy_val, cam.cx_val, cam.cy_val, 64, 48)
    rgb = _render(grid, torch.from_numpy(ro), torch.from_numpy(rd), gpu)
    assert np.array_equal(img.reshape(-1, 3), rgb)
    rgb64 = O.render_fp64(g, ro, rd, OPT, torch.from_numpy(g.density),
                          torch.from_numpy(g.sh)).numpy()
    assert np.abs(rgb - rgb64).max() < 2e-5
    rays = cam.gen_rays()
    rgb_gen = _render(grid, rays.origins, rays.dirs, gpu)
    print(f"image kernel vs Camera.gen_rays rays: {np.abs(img.reshape(-1, 3) - rgb_gen).max():.3e}")
    assert np.abs(img.reshape(-1, 3) - rgb_gen).max() < 2e-5


def test_contiguous_and_interleaved_layouts_agree(gpu, dense64):
    """Separate contiguous density (R,1) / sh (R,27) tensors, as an external
    caller passes them, render and differentiate exactly as the package's
    interleaved row layout."""
    _, o, d, gt, sc = dense64
    grid = _grid(sc, gpu)
    assert grid.sh_data.stride(0) == 32 and grid.density_data.stride(0) == 32
    spec = grid._to_cpp()
    spec.density_data = grid.density_data.data.contiguous()
    spec.sh_data = grid.sh_data.data.contiguous()
    rays = svox2_amd.Rays(torch.from_numpy(o[:1024]).to(gpu),
                          torch.from_numpy(d[:1024]).to(gpu))._to_cpp()
    opt = grid.opt._to_cpp()
    a = _C.volume_render_cuvol(grid._to_cpp(), rays, opt)
    b = _C.volume_render_cuvol(spec, rays, opt)
    assert torch.equal(a, b)
    gt_t = torch.from_numpy(gt[:1024]).to(gpu)
    outs = []
    for sp, interleaved in ((grid._to_cpp(), True), (spec, False)):
        h = _C.GridOutputGrads()
        if interleaved:
            gdv, gsv = svox2_amd.svox2.interleaved_rows(sp.sh_data.size(0), 27, torch.float32, gpu)
        else:
            gdv = torch.zeros(sp.sh_data.size(0), 1, device=gpu)
            gsv = torch.zeros(sp.sh_data.size(0), 27, device=gpu)
        h.grad_density_out, h.grad_sh_out = gdv, gsv
        out = torch.empty_like(a)
        _C.volume_render_cuvol_fused(sp, rays, opt, gt_t, 0.0, 0.0, out, h)
        outs.append((out, gdv.contiguous(), gsv.contiguous()))
    assert torch.equal(outs[0][0], outs[1][0])
    for x, y in zip(outs[0][1:], outs[1][1:]):
        assert (x - y).abs().max() <= 1e-6 * y.abs().max()


def _cell_map_np(links: np.ndarray, rows: int) -> np.ndarray:
    """The cell map as the header defines it, by plain loops over blocks."""
    s = links.shape
    m = [(n - 2) // 2 + 1 for n in s]
    out = np.zeros(m, np.uint16)
    pad = np.full((s[0] + 2, s[1] + 2, s[2] + 2), -1, np.int64)
    pad[:s[0], :s[1], :s[2]] = links
    valid = (pad >= 0) & (pad < rows)
    for bx in range(m[0]):
        for by in range(m[1]):
            for bz in range(m[2]):
                levels, occ = set(), 0
                for i in range(2):
                    for j in range(2):
                        for k in range(2):
                            x, y, z = 2 * bx + i, 2 * by + j, 2 * bz + k
                            if x > s[0] - 2 or y > s[1] - 2 or z > s[2] - 2:
                                continue
                            l0 = pad[x, y, z]
                            levels.add(int(-l0 - 1) if l0 < -1 else 0)
                            if valid[x:x + 2, y:y + 2, z:z + 2].any():
                                occ |= 1 << (4 * i + 2 * j + k)
                lv = levels.pop() if len(levels) == 1 else 255
                out[bx, by, bz] = min(lv, 255) | (occ << 8)
    return out.reshape(-1)


@pytest.mark.parametrize("shape", [(16, 16, 16), (17, 12, 9)])
def test_cell_map_definition(gpu, shape):
    """svox2_build_cell_map against the header's definition on random links
    mixing rows, -1, skip levels and out-of-range values (odd sizes too)."""
    g = np.random.default_rng(3)
    links = g.integers(-4, 40, size=shape).astype(np.int32)
    links[links > 30] = -1
    cmap = _C.build_cell_map(torch.from_numpy(links).to(gpu), 25)
    torch.cuda.synchronize()
    got = cmap.cpu().numpy().view(np.uint16)
    assert np.array_equal(got, _cell_map_np(links, 25))


def test_cell_map_march_bit_exact(gpu, sparse256, march):
    """The march with the cell map equals the oracle and the march without
    it, bit for bit, on accelerated links and on links whose empty entries are
    arbitrary negative values (blocks of mixed levels take the per-cell path)."""
    sc, grid, links_acc, o, d = sparse256
    g = np.random.default_rng(11)
    mixed = links_acc.copy()
    empty = mixed < 0
    mixed[empty] = -g.integers(1, 6, size=int(empty.sum())).astype(np.int32)
    for lk in (links_acc, mixed):
        grid.links = torch.from_numpy(lk).to(gpu)
        st = {}
        for use_map in (None, False):
            spec = grid._to_cpp()
            spec.cell_map = use_map
            rays = svox2_amd.Rays(o.to(gpu), d.to(gpu))
            st[use_map] = _C.volume_render_cuvol_stats(spec, rays._to_cpp(),
                                                       grid.opt._to_cpp()).cpu().numpy()
        _, _, ref = march.render(host_grid(sc, links=lk), o.numpy(), d.numpy(), OPT)
        assert np.array_equal(st[None], ref) and np.array_equal(st[False], ref)
    grid.links = torch.from_numpy(links_acc).to(gpu)
