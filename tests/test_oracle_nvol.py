"""Pins of the nvol oracle (render_lerp_kernel_nvol.cu): its compositing
against a closed form on uniform fields, and its restated two-pass backward
against the autograd gradient of its compositing (with the sparsity term),
on grids where every sample's density exceeds the threshold; the one case
where the reference's backward is not that gradient (pass 2 also visits
samples at or below the threshold) is shown with its counterpart."""
import math

import numpy as np
import torch

import oracle.oracle as O
from svox2_amd import scenes


def _grid(N, dens, sh, basis_dim=9):
    links = np.arange(N ** 3, dtype=np.int32).reshape(N, N, N)
    Nf = np.array([N, N, N], np.float64)
    return O.Grid(links, np.asarray(dens, np.float64), np.asarray(sh, np.float64), basis_dim,
                  offset=0.5 * Nf - 0.5, scaling=0.5 * Nf)


def _rays(n, seed):
    g = torch.Generator().manual_seed(seed)
    o, d = scenes.sphere_rays(n, g, radius=1.7, target=0.5)
    return o.numpy(), d.numpy()


def test_nvol_uniform_closed_form():
    """Uniform density sigma and colour c: every sample adds a = 1 - e^{-ws
    sigma}, so A = min(M a, 1) for a ray of M samples and C = A c + (1 - A) bg;
    both a density whose rays mostly stay below A = 1 and one whose rays
    saturate."""
    N = 12
    dc = np.array([0.8, -0.2, 0.3])
    sh = np.zeros((N ** 3, 27))
    sh[:, [0, 9, 18]] = dc
    c = np.maximum(dc * O.SH_C0 + 0.5, 0.0)
    o, d = _rays(200, 1)
    opt = O.Opts(background_brightness=0.7, sigma_thresh=0.0)
    for sigma, expect_sat in ((0.05, False), (3.0, True)):
        g = _grid(N, np.full(N ** 3, sigma), sh)
        smp = O.nvol_samples(g, o, d, opt)
        rgb = O.nvol_fp64(g, o, d, opt, torch.tensor(g.density), torch.tensor(g.sh),
                          samples=smp).numpy()
        M = np.bincount(smp["ray"], minlength=len(o))
        ws = np.zeros(len(o))
        ws[smp["ray"]] = smp["ws"]
        A = np.minimum(M * (1.0 - np.exp(-ws * sigma)), 1.0)
        ref = A[:, None] * c[None, :] + (1.0 - A)[:, None] * 0.7
        assert np.abs(rgb - ref).max() < 1e-12
        hit = M > 0
        assert (A[hit] >= 1.0).mean() > 0.5 if expect_sat else (A[hit] < 1.0).mean() > 0.5


def _energy_terms(g, o, d, opt, smp, sparsity):
    """sum_r g_r . C_r + sparsity * sum log(1 + 2 sigma^2) over the samples the
    reference's pass 1 gives a density term (contributing, before the
    saturating one, not it), as a function of density and sh."""
    dens = torch.tensor(g.density, requires_grad=True)
    sh = torch.tensor(g.sh, requires_grad=True)
    rgb = O.nvol_fp64(g, o, d, opt, dens, sh, samples=smp)
    sigma, _, _, _, _, _, _ = O._nvol_sample_terms(g, d, smp, dens, sh)
    ws = smp["ws"].astype(np.float64)
    contrib = (smp["flag"] & 1).astype(bool)
    sel = np.zeros(len(ws), bool)
    bounds = np.concatenate([[0], np.cumsum(np.bincount(smp["ray"], minlength=len(o)))])
    sv = sigma.detach().numpy()
    for r in range(len(o)):
        alpha = 0.0
        for i in range(bounds[r], bounds[r + 1]):
            if not contrib[i]:
                continue
            alpha = min(alpha + 1.0 - math.exp(-ws[i] * sv[i]), 1.0)
            if alpha >= 1.0:
                break
            sel[i] = True
    sp = torch.log(1.0 + 2.0 * sigma[torch.from_numpy(sel)] ** 2).sum()
    return dens, sh, rgb, sp


def test_nvol_backward_is_autograd_gradient():
    """On a dense grid of positive densities (every sample contributes), the
    restated two-pass backward equals the autograd gradient of the compositing
    plus the Cauchy sparsity term, for rays that saturate and rays that do
    not; its mask is the set of rows that received a density term (here every
    row touched by a sample up to the saturating one)."""
    N = 10
    rng = np.random.default_rng(3)
    dens = rng.uniform(0.02, 0.8, N ** 3)
    sh = rng.normal(0.0, 0.4, (N ** 3, 27))
    g = _grid(N, dens, sh)
    o, d = _rays(120, 2)
    opt = O.Opts(background_brightness=0.6, sigma_thresh=1e-6)
    smp = O.nvol_samples(g, o, d, opt)
    assert (smp["flag"] & 1).all()
    gout = rng.normal(size=(len(o), 3))
    for sparsity in (0.0, 0.3):
        gd, gs, mask = O.nvol_grads_fp64(g, o, d, opt, gout, sparsity=sparsity, samples=smp)
        dt, st, rgb, sp = _energy_terms(g, o, d, opt, smp, sparsity)
        ((rgb * torch.tensor(gout)).sum() + sparsity * sp).backward()
        for a, b in ((gd, dt.grad.numpy()), (gs, st.grad.numpy())):
            assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
        assert mask[np.abs(gd) > 0].all() and mask.sum() > 100
    # some rays saturate, others reach the background
    sigma = O._nvol_sample_terms(g, d, smp, torch.tensor(dens), torch.tensor(sh))[0].numpy()
    bounds = np.concatenate([[0], np.cumsum(np.bincount(smp["ray"], minlength=len(o)))])
    sat = []
    for r in range(len(o)):
        alpha = 0.0
        for i in range(bounds[r], bounds[r + 1]):
            alpha = min(alpha + 1.0 - math.exp(-smp["ws"][i] * sigma[i]), 1.0)
        sat.append(alpha >= 1.0)
    assert 0 < sum(sat) < len(o)


def test_nvol_pass2_visits_subthreshold_samples():
    """Where samples before the saturating one have density at or below the
    threshold, the reference's pass 2 still gives them -ws e L (the autograd
    gradient of the compositing does not); lowering the threshold below every
    density makes both the same again."""
    N = 10
    rng = np.random.default_rng(5)
    dens = rng.uniform(0.2, 6.0, N ** 3)
    dens[: N ** 3 // 2] = 0.0  # half the grid (x < N / 2) empty
    sh = rng.normal(0.0, 0.4, (N ** 3, 27))
    g = _grid(N, dens, sh)
    o, d = _rays(80, 4)
    gout = rng.normal(size=(len(o), 3))
    opt = O.Opts(background_brightness=0.6, sigma_thresh=1e-6)
    smp = O.nvol_samples(g, o, d, opt)
    assert not (smp["flag"] & 1).all()
    gd, gs, _ = O.nvol_grads_fp64(g, o, d, opt, gout, samples=smp)
    dt, st, rgb, _ = _energy_terms(g, o, d, opt, smp, 0.0)
    (rgb * torch.tensor(gout)).sum().backward()
    assert np.abs(gs - st.grad.numpy()).max() <= 1e-10 * np.abs(gs).max()
    assert np.abs(gd - dt.grad.numpy()).max() > 1e-3 * np.abs(gd).max()
    opt_all = O.Opts(background_brightness=0.6, sigma_thresh=-1.0)
    smp = O.nvol_samples(g, o, d, opt_all)
    assert (smp["flag"] & 1).all()
    gd, gs, _ = O.nvol_grads_fp64(g, o, d, opt_all, gout, samples=smp)
    dt, st, rgb, _ = _energy_terms(g, o, d, opt_all, smp, 0.0)
    (rgb * torch.tensor(gout)).sum().backward()
    assert np.abs(gd - dt.grad.numpy()).max() <= 1e-10 * np.abs(gd).max()
    assert np.abs(gs - st.grad.numpy()).max() <= 1e-10 * np.abs(gs).max()
