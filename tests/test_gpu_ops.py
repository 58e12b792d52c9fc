"""GPU sampling, total variation, RMSprop and accel_dist_prop against the
oracle.  Tolerances: trilinear samples 1e-5 absolute (float32 lerps of O(10)
values); gradients 1e-5 relative to their max; RMSprop data 1e-6 and rms 1e-5
relative (float32 arithmetic of the closed form; BF16 data within its rounding);
accel_dist_prop and masks bit-exact."""
import numpy as np
import pytest
import torch

import oracle.oracle as O
from svox2_amd import _C, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_sparse(gpu):
    gen = torch.Generator().manual_seed(4)
    sc = scenes.make_blob_grid(32, scenes.BlobSpec(5, (0.1, 0.2), (5.0, 20.0)), gen)
    return sc, gen


def test_accel_dist_prop_bit_exact(gpu, small_sparse):
    """Bit-exact against the oracle (itself checked against brute-force
    pyramids) on the blob grid and on random links of odd and degenerate
    shapes: a unit axis (no pyramid level), a pyramid of one 2x2x2 level,
    long thin grids and sizes that are not powers of two."""
    sc, _ = small_sparse
    rng = np.random.default_rng(0)
    shapes = [(13, 9, 20), (1, 5, 7), (2, 2, 2), (3, 300, 2), (70, 65, 129), (33, 17, 2)]
    randoms = [torch.from_numpy(np.where(rng.random(sh) < p, np.arange(np.prod(sh)).reshape(sh),
                                         -1).astype(np.int32))
               for sh, p in zip(shapes, (0.03, 0.2, 0.1, 0.01, 0.002, 0.05))]
    for links in [sc.links] + randoms:
        ref = O.accel_dist_prop(links.numpy())
        t = links.clone().to(gpu)
        _C.accel_dist_prop(t)
        assert np.array_equal(t.cpu().numpy(), ref)
        _C.accel_dist_prop(t)  # idempotent
        assert np.array_equal(t.cpu().numpy(), ref)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_sample_forward_backward(gpu, small_sparse, dtype):
    sc, gen = small_sparse
    grid = scenes.to_sparse_grid(sc, device=gpu, dtype=dtype, accelerate=False)
    pts = torch.rand(3000, 3, generator=gen) * 2.4 - 1.2
    d, s = grid.sample(pts.to(gpu))
    dens64 = grid.density_data.data.double().cpu().reshape(-1).requires_grad_(True)
    sh64 = grid.sh_data.data.double().cpu().requires_grad_(True)
    N = 32
    gp = (pts.double() * (N / 2) + (N / 2 - 0.5)).numpy()
    rd, rs = O.sample_trilerp(sc.links.numpy(), dens64, sh64, gp)
    assert np.abs(d.detach().cpu().numpy() - rd.detach().numpy()).max() < 1e-5
    assert np.abs(s.detach().cpu().numpy() - rs.detach().numpy()).max() < 1e-5
    wd = torch.randn(3000, 1, generator=gen)
    ws = torch.randn(3000, 27, generator=gen)
    ((d * wd.to(gpu)).sum() + (s * ws.to(gpu)).sum()).backward()
    ((rd * wd.double()).sum() + (rs * ws.double()).sum()).backward()
    gd = grid.density_data.grad.cpu().numpy().reshape(-1)
    gs = grid.sh_data.grad.cpu().numpy()
    assert np.abs(gd - dens64.grad.numpy()).max() <= 1e-5 * np.abs(dens64.grad.numpy()).max()
    assert np.abs(gs - sh64.grad.numpy()).max() <= 1e-5 * np.abs(sh64.grad.numpy()).max()
    # density only
    d2, s2 = grid.sample(pts.to(gpu), want_colors=False)
    assert s2.shape[0] == 0 and torch.equal(d2, d.detach())


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_tv_grad_sparse(gpu, small_sparse, dtype):
    sc, gen = small_sparse
    links = sc.links.to(gpu)
    R = sc.density.shape[0]
    cells = torch.arange(32 ** 3 - 700, 32 ** 3 + 300, dtype=torch.int32) % (32 ** 3)
    for data, lo, hi, scale, ie, lz in ((sc.density, 0, 1, 1e-2, False, True),
                                        (sc.sh, 0, 27, 1e-3, True, False),
                                        (sc.sh, 4, 13, 2.0, False, False)):
        dt = data.to(dtype)
        grad = torch.zeros(dt.shape, dtype=torch.float32, device=gpu)
        mask = torch.zeros(R, dtype=torch.bool, device=gpu)
        _C.tv_grad_sparse(links, dt.to(gpu).contiguous(), cells.to(gpu), mask, lo, hi, scale,
                          False, 2.0, ie, lz, -1.0, -1.0, grad)
        rg, rm = O.tv_grad_sparse(sc.links.numpy(), dt.double().numpy(), cells.numpy(), lo, hi,
                                  scale, ie, lz)
        g = grad.cpu().numpy()
        assert np.abs(g - rg).max() <= 1e-5 * max(np.abs(rg).max(), 1e-30)
        assert np.array_equal(mask.cpu().numpy(), rm)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_tv_grad_sparse_run_matches_cell_list(gpu, small_sparse, dtype):
    """tv_grad_sparse_run (cells generated in the kernel) against the oracle
    on the same contiguous runs the cell-list entry point is given, including
    a run that wraps past the last cell and a run of every cell; masks
    bit-exact, gradients within 1e-5 of their max (atomic order)."""
    sc, _ = small_sparse
    links = sc.links.to(gpu)
    R = sc.density.shape[0]
    n = 32 ** 3
    for first, count, data, lo, hi, scale, ie, lz in (
            (n - 700, 1000, sc.density, 0, 1, 1e-2, False, True),
            (5, 4000, sc.sh, 0, 27, 1e-3, True, False),
            (0, n, sc.sh, 4, 13, 2.0, False, False),
            (n - 1, 1, sc.density, 0, 1, 1.0, False, False)):
        dt = data.to(dtype)
        grad = torch.zeros(dt.shape, dtype=torch.float32, device=gpu)
        mask = torch.zeros(R, dtype=torch.bool, device=gpu)
        _C.tv_grad_sparse_run(links, dt.to(gpu).contiguous(), first, count, mask, lo, hi,
                              scale, ie, lz, grad)
        cells = (torch.arange(first, first + count, dtype=torch.int64) % n).to(torch.int32)
        rg, rm = O.tv_grad_sparse(sc.links.numpy(), dt.double().numpy(), cells.numpy(), lo, hi,
                                  scale, ie, lz)
        g = grad.cpu().numpy()
        # Row 0 collects every past-the-edge neighbour (link 0, the reference's
        # edge rule): with every cell listed that is some 3000 float32 atomic
        # terms in arbitrary order, so it gets a bound of its own.
        top = max(np.abs(rg).max(), 1e-30)
        assert np.abs(g[1:] - rg[1:]).max() <= 1e-5 * top, (first, count)
        assert np.abs(g[0] - rg[0]).max() <= 1e-4 * top, (first, count)
        assert np.array_equal(mask.cpu().numpy(), rm), (first, count)
    # a run outside the grid is refused
    with pytest.raises(RuntimeError):
        _C.tv_grad_sparse_run(links, sc.density.to(gpu), n, 1, None, 0, 1, 1.0, False, False,
                              torch.zeros(R, 1, device=gpu))


def _spec(links, sh, basis_dim):
    spec = _C.SparseGridSpec()
    spec.links = links
    spec.sh_data = sh
    spec.density_data = torch.zeros(sh.size(0), 1, dtype=sh.dtype, device=sh.device)
    spec._offset = torch.zeros(3)
    spec._scaling = torch.ones(3)
    spec.basis_dim = basis_dim
    return spec


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_lumisphere_tv_grad_sparse(gpu, small_sparse, dtype):
    """lumisphere_tv_grad_sparse against the oracle: the blob grid's SH rows
    (basis 9) with every flat index listed (those past the size - 1 extents
    included) and a random list with repeats; an odd random grid with basis
    1, 4, 16 (48 coefficients, the third channel past the 32nd) and 25 (the
    second channel cut at the 32nd), in interleaved rows of a wider buffer;
    dir_factor 0 and scale 0 (rows still marked where a term is nonzero).
    Masks bit-exact, gradients within 1e-5 of their max (float32 colour sums
    and atomic order)."""
    sc, _ = small_sparse
    rng = np.random.default_rng(31)
    shape = (19, 13, 10)
    R = 700
    odd = rng.integers(-3, R, size=shape).astype(np.int32)
    odd[rng.random(shape) < 0.4] = -1
    cases = []
    for cells in (np.arange(32 ** 3, dtype=np.int32),
                  rng.integers(0, 32 ** 3, size=5000).astype(np.int32)):
        cases.append((sc.links.numpy(), sc.sh.numpy(), 9, cells, 0.7, 0.5))
    odd_cells = rng.integers(0, int(np.prod(shape)), size=3000).astype(np.int32)
    for B, f, scale in ((1, 0.3, 1.0), (4, 0.0, 2.0), (16, 0.8, 0.5), (25, 1.2, 0.5),
                        (9, 0.4, 0.0)):
        cases.append((odd, rng.standard_normal((R, 3 * B)), B, odd_cells, f, scale))
    for links, data, B, cells, f, scale in cases:
        D = data.shape[1]
        host = torch.from_numpy(np.ascontiguousarray(data)).float().to(dtype)
        buf = torch.zeros(host.size(0), D + 5, dtype=dtype, device=gpu)
        buf[:, 2:2 + D] = host.to(gpu)
        gbuf = torch.zeros(host.size(0), D + 3, device=gpu)
        mask = torch.zeros(host.size(0), dtype=torch.bool, device=gpu)
        sv = torch.from_numpy(rng.standard_normal(B)).float()
        svu = torch.from_numpy(rng.standard_normal(B)).float()
        grads = _C.GridOutputGrads()
        grads.grad_sh_out = gbuf[:, 1:1 + D]
        grads.mask_out = mask
        lk = torch.from_numpy(links).to(gpu)
        _C.lumisphere_tv_grad_sparse(_spec(lk, buf[:, 2:2 + D], B), torch.from_numpy(cells).to(gpu),
                                     sv.to(gpu), svu.to(gpu), scale, -1.0, -1.0, f, grads)
        rg, rm = O.lumisphere_tv_grad_sparse(links, host.double().numpy(), cells, B,
                                             sv.double().numpy(), svu.double().numpy(), scale, f)
        g = gbuf.cpu().numpy()
        tag = (B, D, len(cells), f, scale)
        assert np.abs(g[:, 1:1 + D] - rg).max() <= 1e-5 * max(np.abs(rg).max(), 1e-30), tag
        assert not g[:, 0].any() and not g[:, 1 + D:].any(), tag
        assert np.array_equal(mask.cpu().numpy(), rm) and rm.any(), tag
        if scale == 0.0:
            assert not g.any(), tag


def test_inplace_tv_lumisphere_grad_draws(gpu, small_sparse):
    """The facade draws the reference's random numbers in its order (run start
    from np.random, the direction from torch.randn on the device, one rotation
    axis per listed cell from torch.randn on the CPU; svox2.py:1847-1884) and
    adds the oracle's gradient for the first axis's basis values, which are the
    ones the reference's kernel reads; the mask goes to the SH indexer."""
    from scipy.spatial.transform import Rotation

    sc, _ = small_sparse
    grid = scenes.to_sparse_grid(sc, device=gpu, dtype=torch.float32, accelerate=False)
    grad = grid._get_data_grads()[1]
    grid.sparse_sh_grad_indexer = torch.zeros(grid.sh_data.size(0), dtype=torch.bool,
                                              device=gpu)
    for f in (0.5, 0.0):
        grad.zero_()
        grid.sparse_sh_grad_indexer.zero_()
        np.random.seed(11)  # a run that crosses the blobs
        torch.manual_seed(6)
        grid.inplace_tv_lumisphere_grad(grad, scaling=0.8, sparse_frac=0.05, dir_factor=f,
                                        dir_perturb_radians=0.3)
        np.random.seed(11)
        torch.manual_seed(6)
        n = 32 ** 3
        count = int(0.05 * n)
        cells = (np.random.randint(0, n) + np.arange(count)) % n
        dirs = torch.randn(3, device=gpu)
        dirs /= torch.norm(dirs)
        if f > 0:
            axis = torch.randn((count, 3))
            axis /= torch.norm(axis, dim=-1, keepdim=True)
            axis *= 0.3
            rot = torch.from_numpy(Rotation.from_rotvec(axis.numpy()).as_matrix()).float()
            dirs_u = (rot.to(gpu) * dirs.unsqueeze(-2)).sum(-1)  # every cell's, as the reference
        else:
            dirs_u = dirs[None]
        sv = O.eval_sh_bases(9, dirs.double().cpu()[None])[0].numpy()
        svu = O.eval_sh_bases(9, dirs_u.double().cpu())[0].numpy()
        rg, rm = O.lumisphere_tv_grad_sparse(sc.links.numpy(),
                                             grid.sh_data.data.double().cpu().numpy(), cells, 9,
                                             sv, svu, 0.8, f)
        g = grad.cpu().numpy()
        assert np.abs(g - rg).max() <= 1e-5 * np.abs(rg).max(), f
        assert np.array_equal(grid.sparse_sh_grad_indexer.cpu().numpy(), rm) and rm.any(), f


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_rmsprop_all_mask_index(gpu, dtype):
    gen = torch.Generator().manual_seed(9)
    R, D = 1000, 27
    data0 = torch.randn(R, D, generator=gen).to(dtype)
    rms0 = torch.rand(R, D, generator=gen) * (torch.rand(R, D, generator=gen) > 0.3)
    grad0 = torch.randn(R, D, generator=gen)
    maskt = torch.rand(R, generator=gen) < 0.2
    for kind, indexer, rows in (("all", torch.empty(()), np.arange(R)),
                                ("mask", maskt, np.nonzero(maskt.numpy())[0]),
                                ("index", torch.nonzero(maskt).flatten(),
                                 np.nonzero(maskt.numpy())[0])):
        data, rms, grad = data0.clone().to(gpu), rms0.clone().to(gpu), grad0.clone().to(gpu)
        _C.rmsprop_step(data, rms, grad, indexer.to(gpu), 0.95, 0.1, 1e-8, -1.5, 0.05)
        rd, rr, rgr = O.rmsprop_step(data0.double().numpy(), rms0.double().numpy(),
                                     grad0.double().numpy(), rows, 0.95, 0.1, 1e-8, -1.5, 0.05)
        tol = 1e-6 if dtype == torch.float32 else 8e-3
        assert np.allclose(data.double().cpu().numpy(), rd, rtol=tol, atol=tol), kind
        # rms = fma(beta, rms - g^2, g^2) cancels when g^2 >> rms: float32 keeps
        # about 1e-6 relative there
        assert np.allclose(rms.cpu().numpy(), rr, rtol=1e-5, atol=1e-12), kind
        assert np.array_equal(grad.cpu().numpy(), rgr.astype(np.float32)), kind
    # an empty 1-d indexer skips the update
    data = data0.clone().to(gpu)
    _C.rmsprop_step(data, rms0.clone().to(gpu), grad0.clone().to(gpu),
                    torch.zeros(0, dtype=torch.int64, device=gpu), 0.95, 0.1, 1e-8, -1.5, 0.05)
    assert torch.equal(data.cpu(), data0)


def _interleaved(R, gen, dtype, dev):
    """(buffer (R, 32), density view (R, 1), sh view (R, 27)) of the package's
    voxel row layout, random contents."""
    buf = torch.randn(R, 32, generator=gen).to(dtype).to(dev)
    return buf, buf[:, 27:28], buf[:, :27]


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_rmsprop_and_sgd_on_interleaved_rows(gpu, dtype):
    """rmsprop_step and sgd_step on the strided views of the interleaved row
    layout (density: one column at offset 27; SH: 27 columns), every indexer
    kind, against the closed forms; a sparse mask spanning many 64-row
    windows; index entries out of range are skipped; pad columns untouched."""
    gen = torch.Generator().manual_seed(12)
    R = 70001
    mask = torch.rand(R, generator=gen) < 0.03
    rows_m = np.nonzero(mask.numpy())[0]
    idx = torch.cat([torch.nonzero(mask).flatten(), torch.tensor([-3, R, R + 7])])
    for view in ("density", "sh"):
        for kind, indexer, rows in (("all", torch.empty(()), np.arange(R)),
                                    ("mask", mask, rows_m), ("index", idx, rows_m)):
            for optim in ("rmsprop", "sgd"):
                data_b, dd, ds = _interleaved(R, gen, dtype, gpu)
                rms_b, rd_, rs_ = _interleaved(R, gen, torch.float32, gpu)
                rms_b.abs_()
                grad_b, gd_, gs_ = _interleaved(R, gen, torch.float32, gpu)
                data, rms, grad = (dd, rd_, gd_) if view == "density" else (ds, rs_, gs_)
                before = [t.cpu().clone() for t in (data_b, rms_b, grad_b)]
                d0, r0, g0 = (t.double().cpu().numpy() for t in (data, rms, grad))
                if optim == "rmsprop":
                    _C.rmsprop_step(data, rms, grad, indexer.to(gpu), 0.95, 0.1, 1e-8, -1.5, 0.05)
                    wd, wr, wg = O.rmsprop_step(d0, r0, g0, rows, 0.95, 0.1, 1e-8, -1.5, 0.05)
                else:
                    _C.sgd_step(data, grad, indexer.to(gpu), 0.1, 0.05)
                    wd, wg = O.sgd_step(d0, g0, rows, 0.1, 0.05)
                    wr = r0
                tol = 1e-6 if dtype == torch.float32 else 8e-3
                tag = (view, kind, optim)
                assert np.allclose(data.double().cpu().numpy(), wd, rtol=tol, atol=tol), tag
                assert np.allclose(rms.cpu().numpy(), wr, rtol=1e-5, atol=1e-12), tag
                assert np.array_equal(grad.cpu().numpy(), wg.astype(np.float32)), tag
                # columns outside the view are untouched
                cols = slice(27, 28) if view == "density" else slice(0, 27)
                other = np.ones(32, bool)
                other[cols] = False
                for b, b0 in zip((data_b, rms_b, grad_b), before):
                    assert torch.equal(b.cpu()[:, other], b0[:, other]), tag


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_rmsprop_grouped_equals_two_calls(gpu, dtype):
    """rmsprop_step_grouped over [SH | density] rows with separate masks and
    learning rates equals rmsprop_step on the SH view then on the density
    view, bit for bit; the facade's optim_density_sh_step equals
    optim_sh_step + optim_density_step."""
    gen = torch.Generator().manual_seed(13)
    R = 50000
    m_sh = (torch.rand(R, generator=gen) < 0.05).to(gpu)
    m_d = (torch.rand(R, generator=gen) < 0.05).to(gpu)
    data_b, dd, ds = _interleaved(R, gen, dtype, gpu)
    rms_b, rd_, rs_ = _interleaved(R, gen, torch.float32, gpu)
    rms_b.abs_().mul_(torch.rand(R, 32, generator=gen).to(gpu) > 0.5)
    grad_b, gd_, gs_ = _interleaved(R, gen, torch.float32, gpu)
    copies = [t.clone() for t in (data_b, rms_b, grad_b)]
    _C.rmsprop_step(ds, rs_, gs_, m_sh, 0.95, 1e-2, 1e-8, -1e9, 1e-2)
    _C.rmsprop_step(dd, rd_, gd_, m_d, 0.9, 30.0, 1e-8, -1e9, 30.0)
    cd, cr, cg = copies
    _C.rmsprop_step_grouped(cd[:, :28], cr[:, :28], cg[:, :28], 27, m_sh, m_d,
                            0.95, 1e-2, 0.9, 30.0, 1e-8, -1e9)
    for a, b in zip((data_b, rms_b, grad_b), copies):
        assert torch.equal(a, b)
    # the facade
    gen = torch.Generator().manual_seed(2)
    sc = scenes.make_blob_grid(32, scenes.BlobSpec(5, (0.1, 0.2), (5.0, 20.0)), gen)
    grids = [scenes.to_sparse_grid(sc, device=gpu, dtype=dtype) for _ in range(2)]
    Rg = grids[0].density_data.size(0)
    gg = torch.randn(Rg, 32, generator=gen).to(gpu)
    md = (torch.rand(Rg, generator=gen) < 0.3).to(gpu)
    ms = (torch.rand(Rg, generator=gen) < 0.3).to(gpu)
    for g in grids:
        gd, gs, _, _ = g._get_data_grads()
        gd.copy_(gg[:, 27:28])
        gs.copy_(gg[:, :27])
        g.sparse_grad_indexer, g.sparse_sh_grad_indexer = md.clone(), ms.clone()
    grids[0].optim_sh_step(1e-2, beta=0.95)
    grids[0].optim_density_step(30.0, beta=0.95)
    grids[1].optim_density_sh_step(30.0, 1e-2, beta=0.95)
    for a, b in ((grids[0].density_data, grids[1].density_data),
                 (grids[0].sh_data, grids[1].sh_data)):
        assert torch.equal(a.data, b.data) and torch.equal(a.grad, b.grad)
    assert torch.equal(grids[0].sh_rms, grids[1].sh_rms)
    assert torch.equal(grids[0].density_rms, grids[1].density_rms)


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_tv_grad_sparse_shapes_and_layouts(gpu, dtype):
    """tv_grad_sparse against the oracle beyond the blob grid: an odd grid
    (17 x 12 x 9, so the flat cell index splits by non-powers of two) with
    links mixing rows, -1 and skip levels; random cells with repeats; the
    SH columns and the density column of interleaved 32-wide rows; 40
    channels (more than one pass of a 32-lane half); and scale 0, where the
    gradient stays zero but rows are still marked wherever a difference is
    nonzero, as the reference's rule marks them."""
    rng = np.random.default_rng(21)
    shape = (17, 12, 9)
    cells_n = int(np.prod(shape))
    R = 600
    links = rng.integers(-4, R, size=shape).astype(np.int32)
    links[rng.random(shape) < 0.5] = -1
    cells = rng.integers(0, cells_n, size=900).astype(np.int32)
    buf = torch.from_numpy(rng.standard_normal((R, 32))).float().to(dtype)
    wide = torch.from_numpy(rng.standard_normal((R, 40))).float().to(dtype)
    cases = [("sh", 0, 27, 1e-3, True, False), ("density", 0, 1, 1e-2, False, True),
             ("sh", 3, 20, 0.0, False, False), ("wide", 0, 40, 0.5, True, False),
             ("wide", 5, 39, 2.0, False, True)]
    lk = torch.from_numpy(links).to(gpu)
    cells_d = torch.from_numpy(cells).to(gpu)
    for kind, lo, hi, scale, ie, lz in cases:
        cols = {"sh": slice(0, 27), "density": slice(27, 28)}.get(kind)
        if kind == "wide":
            host = wide
            dev_data = wide.to(gpu)
            grad = torch.zeros(R, 40, device=gpu)
        else:
            host = buf[:, cols]
            dev_data = buf.to(gpu)[:, cols]
            grad = torch.zeros(R, 32, device=gpu)[:, cols]
        mask = torch.zeros(R, dtype=torch.bool, device=gpu)
        _C.tv_grad_sparse(lk, dev_data, cells_d, mask, lo, hi, scale, False, 2.0, ie, lz,
                          -1.0, -1.0, grad)
        rg, rm = O.tv_grad_sparse(links, host.double().numpy(), cells, lo, hi, scale, ie, lz)
        g = grad.cpu().numpy()
        tag = (kind, lo, hi, scale, ie, lz)
        assert np.abs(g - rg).max() <= 1e-5 * max(np.abs(rg).max(), 1e-30), tag
        assert np.array_equal(mask.cpu().numpy(), rm), tag
        if scale == 0.0:
            assert not g.any() and rm.any(), tag
