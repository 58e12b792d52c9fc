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
