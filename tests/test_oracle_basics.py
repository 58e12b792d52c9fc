"""Pins of the oracle's small restatements against closed forms, library
routines and brute force: SH basis, Morton codes, trilinear sampling,
accel_dist_prop, RMSprop."""
import math

import numpy as np
import pytest
import torch
import torch.nn.functional as F

import oracle.oracle as O
from svox2_amd import utils as U


# --------------------------------------------------------------------------- SH
def test_sh_constants_closed_form():
    pi = math.pi
    assert O.SH_C0 == pytest.approx(1 / (2 * math.sqrt(pi)), abs=1e-15)
    assert O.SH_C1 == pytest.approx(math.sqrt(3) / (2 * math.sqrt(pi)), abs=1e-15)
    c2 = [math.sqrt(15) / (2 * math.sqrt(pi)), -math.sqrt(15) / (2 * math.sqrt(pi)),
          math.sqrt(5) / (4 * math.sqrt(pi)), -math.sqrt(15) / (2 * math.sqrt(pi)),
          math.sqrt(15) / (4 * math.sqrt(pi))]
    for a, b in zip(O.SH_C2, c2):
        assert a == pytest.approx(b, abs=1e-15)


def _sphere_quadrature(n_theta=16, n_phi=32):
    """Gauss-Legendre in cos(theta) x uniform phi: exact for polynomials of
    degree < 2*n_theta on the sphere."""
    x, w = np.polynomial.legendre.leggauss(n_theta)
    phi = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    ct, ph = np.meshgrid(x, phi, indexing="ij")
    st = np.sqrt(1 - ct ** 2)
    dirs = np.stack([st * np.cos(ph), st * np.sin(ph), ct], -1).reshape(-1, 3)
    wts = (w[:, None] * np.full(n_phi, 2 * np.pi / n_phi)[None, :]).reshape(-1)
    return torch.from_numpy(dirs), torch.from_numpy(wts)


def test_sh_orthonormal():
    dirs, w = _sphere_quadrature()
    Y = O.eval_sh_bases(9, dirs)
    G = (Y * w[:, None]).T @ Y
    assert torch.allclose(G, torch.eye(9, dtype=torch.float64), atol=1e-12)


def test_sh_specific_values():
    d = torch.tensor([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], dtype=torch.float64)
    Y = O.eval_sh_bases(9, d)
    # Y_2^0 = sqrt(5/(16 pi)) (3 z^2 - 1) appears with the reference's sign/normalisation
    assert Y[0, 6] == pytest.approx(math.sqrt(5 / (16 * math.pi)) * 2, abs=1e-14)
    assert Y[1, 6] == pytest.approx(-math.sqrt(5 / (16 * math.pi)), abs=1e-14)
    assert Y[0, 2] == pytest.approx(O.SH_C1, abs=1e-15)
    assert Y[1, 3] == pytest.approx(-O.SH_C1, abs=1e-15)
    assert Y[2, 1] == pytest.approx(-O.SH_C1, abs=1e-15)
    assert Y[1, 8] == pytest.approx(O.SH_C2[4], abs=1e-15)


@pytest.mark.parametrize("bd", [1, 4, 9])
def test_sh_product_matches_oracle(bd):
    g = torch.Generator().manual_seed(1)
    d = torch.randn(100, 3, generator=g, dtype=torch.float64)
    d /= d.norm(dim=1, keepdim=True)
    assert torch.allclose(U.eval_sh_bases(bd, d), O.eval_sh_bases(bd, d), atol=1e-15)


# ----------------------------------------------------------------------- Morton
def test_morton_bijective_and_matches():
    D = 16
    ref = O.gen_morton(D)
    assert sorted(ref.flatten().tolist()) == list(range(D ** 3))
    prod = U.gen_morton(D).numpy()
    assert np.array_equal(ref, prod)
    for x, y, z in [(0, 0, 1), (0, 1, 0), (1, 0, 0), (3, 5, 7), (1023, 0, 511)]:
        c = O.morton_code_3(x, y, z)
        assert int(U.morton_code_3(torch.tensor(x), torch.tensor(y), torch.tensor(z))) == c
        X, Y, Z = U.inv_morton_code_3(torch.tensor(c))
        assert (int(X), int(Y), int(Z)) == (x, y, z)
    assert O.morton_code_3(1, 0, 0) == 4 and O.morton_code_3(0, 1, 0) == 2


# ---------------------------------------------------------------------- trilerp
def _dense_grid(N, C, seed):
    g = torch.Generator().manual_seed(seed)
    links = np.arange(N ** 3, dtype=np.int32).reshape(N, N, N)
    dens = torch.randn(N ** 3, generator=g, dtype=torch.float64)
    sh = torch.randn(N ** 3, C, generator=g, dtype=torch.float64)
    return links, dens, sh


def test_trilerp_matches_grid_sample():
    N, C = 9, 4
    links, dens, sh = _dense_grid(N, C, 0)
    g = torch.Generator().manual_seed(3)
    world = torch.rand(500, 3, generator=g, dtype=torch.float64) * 2.4 - 1.2  # includes border
    grid_pts = (world * (N / 2) + N / 2 - 0.5).numpy()
    d, s = O.sample_trilerp(links, dens, sh, grid_pts)
    vol = torch.cat([dens.view(N, N, N, 1), sh.view(N, N, N, C)], -1).permute(3, 0, 1, 2)[None]
    ref = F.grid_sample(vol, world[:, [2, 1, 0]].view(1, -1, 1, 1, 3), mode="bilinear",
                        padding_mode="border", align_corners=False)
    ref = ref.view(C + 1, -1).T
    assert torch.allclose(d[:, 0], ref[:, 0], atol=1e-12)
    assert torch.allclose(s, ref[:, 1:], atol=1e-12)


def test_trilerp_exact_on_linear_field_and_empty_corners():
    N = 6
    links = np.arange(N ** 3, dtype=np.int32).reshape(N, N, N)
    X, Y, Z = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    lin = torch.from_numpy((0.5 + 2 * X - 3 * Y + 0.25 * Z).reshape(-1).astype(np.float64))
    pts = np.random.default_rng(0).uniform(0, N - 1, size=(200, 3))
    d, _ = O.sample_trilerp(links, lin, lin[:, None].clone(), pts)
    expect = 0.5 + 2 * pts[:, 0] - 3 * pts[:, 1] + 0.25 * pts[:, 2]
    assert np.allclose(d[:, 0].numpy(), expect, atol=1e-12)
    # an empty corner reads 0: at a corner's own position the sample is 0
    links2 = links.copy()
    links2[2, 3, 4] = -1
    d2, _ = O.sample_trilerp(links2, lin, lin[:, None].clone(), np.array([[2.0, 3.0, 4.0]]))
    assert float(d2[0, 0]) == 0.0


# ------------------------------------------------------------------ accel_dist
def _accel_brute(links):
    """Per cell: -1 - (number of consecutive levels i >= 1 whose aligned block
    of 2^i cells around the cell is empty), by slicing."""
    occ = links >= 0
    S = occ.shape
    out = links.copy()
    nlev = 0
    s = list(S)
    while all(v > 1 for v in s):
        s = [(v + 1) // 2 for v in s]
        nlev += 1
    for x, y, z in zip(*np.nonzero(~occ)):
        k = 0
        for i in range(1, nlev + 1):
            b = 1 << i
            x0, y0, z0 = (x >> i) << i, (y >> i) << i, (z >> i) << i
            if occ[x0:x0 + b, y0:y0 + b, z0:z0 + b].any():
                break
            k += 1
        out[x, y, z] = -1 - k
    return out


@pytest.mark.parametrize("shape,frac,seed", [((12, 10, 9), 0.05, 0), ((16, 16, 16), 0.01, 1),
                                             ((7, 5, 3), 0.2, 2), ((8, 8, 8), 0.0, 3)])
def test_accel_dist_prop_brute_force(shape, frac, seed):
    rng = np.random.default_rng(seed)
    occ = rng.random(shape) < frac
    links = np.where(occ, np.arange(np.prod(shape)).reshape(shape), -1).astype(np.int32)
    a = O.accel_dist_prop(links)
    b = _accel_brute(links)
    assert np.array_equal(a, b)
    # idempotent on already-coded links
    assert np.array_equal(O.accel_dist_prop(a), a)


# ---------------------------------------------------------------------- rmsprop
def test_rmsprop_closed_form():
    data = np.array([[1.0, 2.0], [3.0, -4.0], [0.5, 0.5]])
    rms = np.array([[0.0, 0.25], [1.0, 0.0], [9.0, 9.0]])
    grad = np.array([[0.5, -1.0], [2.0, 0.1], [3.0, 3.0]])
    d, r, g = O.rmsprop_step(data, rms, grad, rows_sel=[0, 1], beta=0.9, lr=0.1, eps=1e-8,
                             minval=-3.5, lr_last=0.2)
    # row 0 col 0: rms==0 -> g^2 = 0.25; x = 1 - 0.1*0.5/(0.5+eps)
    assert r[0, 0] == 0.25 and d[0, 0] == pytest.approx(1 - 0.1 * 0.5 / (0.5 + 1e-8))
    # row 0 col 1 (last channel uses lr_last): rms = 0.9*0.25 + 0.1*1
    assert r[0, 1] == pytest.approx(0.325)
    assert d[0, 1] == pytest.approx(2 + 0.2 / (math.sqrt(0.325) + 1e-8))
    # clamp to minval
    assert d[1, 1] == pytest.approx(max(-4 - 0.2 * 0.1 / (0.1 + 1e-8), -3.5))
    assert (g[:2] == 0).all() and (g[2] == grad[2]).all() and (d[2] == data[2]).all()


def _tv_energy(links, data, cells, lo, hi, scale, ignore_edge, ignore_last_z):
    """Sparse TV energy whose data gradient the reference kernel evaluates on a
    256-cube (loss_kernel.cu:178-246, scale / n_cells from :597): per selected
    cell and channel, (scale / n) * sqrt(1e-9 + dx^2 + dy^2 + dz^2) of the
    forward differences, with the kernel's neighbour and edge rules (a missing
    +z/+y/+x neighbour past the grid reads row 0; an empty link reads 0, or
    v000 when ignore_edge)."""
    s0, s1, s2 = links.shape
    n = len(cells)
    e = data.sum() * 0.0
    for xyz in cells:
        z, y, x = xyz % s2, (xyz // s2) % s1, xyz // (s1 * s2)
        l000 = int(links[x, y, z])
        if ignore_edge and l000 == 0:
            continue
        if ignore_last_z and z == s2 - 2:
            continue
        l001 = int(links[x, y, z + 1]) if (z + 1 < s2 and (not ignore_last_z or z != s2 - 2)) else 0
        l010 = int(links[x, y + 1, z]) if y + 1 < s1 else 0
        l100 = int(links[x + 1, y, z]) if x + 1 < s0 else 0
        v000 = data[l000, lo:hi] if l000 >= 0 else torch.zeros(hi - lo, dtype=data.dtype)
        nul = v000 if ignore_edge else torch.zeros(hi - lo, dtype=data.dtype)
        nb = [data[lk, lo:hi] if lk >= 0 else nul for lk in (l100, l010, l001)]
        d2 = sum((v - v000) ** 2 for v in nb)
        e = e + (scale / n) * torch.sqrt(1e-9 + d2).sum()
    return e


def _tv_case(size, seed, block_at):
    """Links of a `size` cube: -1 except random occupancy (70%) in a 14^3 block
    at `block_at` and in the 6^3 block at the far corner; random rows."""
    g = np.random.default_rng(seed)
    links = np.full((size,) * 3, -1, np.int32)
    a = block_at
    occ = g.random((14, 14, 14)) < 0.7
    links[a:a + 14, a:a + 14, a:a + 14][occ] = 1  # marked, numbered below
    occ2 = g.random((6, 6, 6)) < 0.7
    links[size - 6:, size - 6:, size - 6:][occ2] = 1
    idx = np.flatnonzero(links.reshape(-1) >= 0)
    g.shuffle(idx)
    flat = links.reshape(-1)
    flat[idx] = np.arange(len(idx), dtype=np.int32)
    cells = []
    for x in range(a, a + 12):
        for y in range(a, a + 12):
            for z in range(a, a + 12):
                cells.append((x * size + y) * size + z)
    for x in range(size - 5, size):
        for y in range(size - 5, size):
            for z in range(size - 5, size):
                cells.append((x * size + y) * size + z)
    return links, len(idx), np.array(cells, np.int64)


@pytest.mark.parametrize("ignore_edge,ignore_last_z", [(False, False), (True, False),
                                                       (False, True)])
def test_tv_grad_is_energy_gradient_at_256(ignore_edge, ignore_last_z):
    """On a 256-cube the per-axis factor N/256 is 1 and the TV gradient the
    oracle forms is exactly the autograd gradient of the TV energy; its mask is
    the set of rows whose gradient is non-zero."""
    links, R, cells = _tv_case(256, 5, 100)
    g = np.random.default_rng(6)
    data = g.normal(size=(R, 5))
    lo, hi, scale = 1, 4, 0.37
    grad, mask = O.tv_grad_sparse(links, data, cells, lo, hi, scale, ignore_edge, ignore_last_z)
    t = torch.tensor(data, requires_grad=True)
    _tv_energy(links, t, cells, lo, hi, scale, ignore_edge, ignore_last_z).backward()
    ref = t.grad.numpy()
    assert np.abs(grad - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(mask, np.abs(grad).sum(1) > 0)
    assert mask.sum() > 100


def test_tv_grad_axis_factor():
    """The factor N/256 scales the differences after the normaliser is formed
    (loss_kernel.cu:220-225): the same neighbourhoods in a 128-cube give
    exactly half the 256-cube gradient (it would be equal if the normaliser
    used the scaled differences)."""
    l256, R, _ = _tv_case(256, 8, 60)
    l128 = np.full((128,) * 3, -1, np.int32)
    l128[60:74, 60:74, 60:74] = l256[60:74, 60:74, 60:74]
    cells = np.array([(x * 256 + y) * 256 + z for x in range(60, 72) for y in range(60, 72)
                      for z in range(60, 72)], np.int64)
    cells128 = np.array([(x * 128 + y) * 128 + z for x in range(60, 72) for y in range(60, 72)
                         for z in range(60, 72)], np.int64)
    data = np.random.default_rng(9).normal(size=(R, 2))
    g256, m256 = O.tv_grad_sparse(l256, data, cells, 0, 2, 1.0, False, False)
    g128, m128 = O.tv_grad_sparse(l128, data, cells128, 0, 2, 1.0, False, False)
    assert np.array_equal(m256, m128) and m256.sum() > 100
    assert np.array_equal(g128, 0.5 * g256)


def _lumi_energy(links, data, cells, B, sv, svu, scale, f):
    """Lumisphere TV energy whose coefficient gradient the reference kernel
    evaluates (loss_kernel.cu:336-446, scale / n_cells from :671): per listed
    cell and channel, (scale / n) * sqrt(1e-9 + dx^2 + dy^2 + dz^2 + du^2) of
    the colour differences A(neighbour) - A(v000) times size_axis / 256 and
    (U(v000) - A(v000)) * f, with A / U the coefficient sums weighted by sv /
    svu, over the first 32 coefficients; cells split by the extents size - 1,
    link-0 cells and x >= size0 - 1 skipped, missing v000 = 0, missing
    neighbour = v000."""
    s0, s1, s2 = links.shape
    n = len(cells)
    dim = min(data.shape[1], 32)
    e = data.sum() * 0.0
    zero = torch.zeros(dim, dtype=data.dtype)
    for xyz in cells:
        z, y, x = xyz % (s2 - 1), (xyz // (s2 - 1)) % (s1 - 1), xyz // ((s1 - 1) * (s2 - 1))
        if x >= s0 - 1 or int(links[x, y, z]) == 0:
            continue
        l000 = int(links[x, y, z])
        v000 = data[l000, :dim] if l000 >= 0 else zero
        nb = [data[int(lk), :dim] if int(lk) >= 0 else v000
              for lk in (links[x + 1, y, z], links[x, y + 1, z], links[x, y, z + 1])]
        for c0 in range(0, dim, B):
            c1 = min(c0 + B, dim)
            w, wu = sv[:c1 - c0], svu[:c1 - c0]
            a = (v000[c0:c1] * w).sum()
            dx = ((nb[0][c0:c1] * w).sum() - a) * (s0 / 256)
            dy = ((nb[1][c0:c1] * w).sum() - a) * (s1 / 256)
            dz = ((nb[2][c0:c1] * w).sum() - a) * (s2 / 256)
            du = ((v000[c0:c1] * wu).sum() - a) * f
            e = e + (scale / n) * torch.sqrt(1e-9 + dx * dx + dy * dy + dz * dz + du * du)
    return e


def _lumi_case(shape, seed, occ=0.6):
    """Random links of `shape` (occupied with probability occ, rows numbered
    in random order so that row 0 sits at some cell) and every flat index
    0 .. prod(shape) - 1 as the cell list (those past the (size - 1) extents
    included)."""
    g = np.random.default_rng(seed)
    links = np.where(g.random(shape) < occ, 0, -1).astype(np.int32)
    links[2, 2, 2] = 0
    idx = np.flatnonzero(links.reshape(-1) >= 0)
    g.shuffle(idx)
    links.reshape(-1)[idx] = np.arange(len(idx), dtype=np.int32)
    zero_at = tuple(np.argwhere(links == 0)[0])
    links[zero_at], links[2, 2, 2] = links[2, 2, 2], 0  # row 0 at an interior cell
    return links, len(idx), np.arange(int(np.prod(shape)), dtype=np.int64)


@pytest.mark.parametrize("B,D", [(4, 12), (9, 27), (16, 48), (25, 75)])
def test_lumisphere_grad_is_energy_gradient(B, D):
    """The lumisphere gradient the oracle forms is the autograd gradient of the
    lumisphere energy on an odd grid (its normaliser uses the scaled
    differences, so this holds at any size), including channels cut at the 32nd
    coefficient (B = 16, 25); its mask is the set of rows with a nonzero
    gradient; coefficients past the 32nd receive nothing."""
    links, R, cells = _lumi_case((7, 6, 8), 20 + B)
    g = np.random.default_rng(B)
    data = g.normal(size=(R, D))
    sv, svu = g.normal(size=B), g.normal(size=B)
    grad, mask = O.lumisphere_tv_grad_sparse(links, data, cells, B, sv, svu, 0.3, 0.7)
    t = torch.tensor(data, requires_grad=True)
    _lumi_energy(links, t, cells, B, torch.tensor(sv), torch.tensor(svu), 0.3, 0.7).backward()
    ref = t.grad.numpy()
    assert np.abs(grad - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(mask, np.abs(grad).sum(1) > 0)
    assert mask.sum() > R // 2
    assert (grad[:, 32:] == 0).all()
    # the cell at (2, 2, 2) holds row 0 and is skipped (its occupied neighbours
    # would give terms); the cell below it in z is not, and row 0 receives
    # its +z neighbour term
    g0, m0 = O.lumisphere_tv_grad_sparse(links, data, [(2 * 5 + 2) * 7 + 2], B, sv, svu, 1.0,
                                         0.7)
    assert not g0.any() and not m0.any()
    g1, m1 = O.lumisphere_tv_grad_sparse(links, data, [(2 * 5 + 2) * 7 + 1], B, sv, svu, 1.0,
                                         0.7)
    assert m1[0] and g1[0, :min(D, 32)].any()


def test_lumisphere_reduces_to_tv():
    """One basis function of value 1 for both directions (so du = 0) makes the
    lumisphere term the plain TV term of each channel with the ignore_edge rules;
    a one-hot basis of 9 makes it the plain TV term of coefficients 0, 9 and 18.
    Both against tv_grad_sparse (itself pinned to a TV energy) on a 256-cube,
    where the axis factor is 1, with the cell lists mapped between the size - 1
    and size splits."""
    links, R, _ = _tv_case(256, 11, 90)
    pts = [(x, y, z) for x in range(90, 102) for y in range(90, 102) for z in range(90, 102)]
    pts += [(x, y, z) for x in range(250, 255) for y in range(250, 255) for z in range(250, 255)]
    lumi_cells = np.array([(x * 255 + y) * 255 + z for x, y, z in pts], np.int64)
    tv_cells = np.array([(x * 256 + y) * 256 + z for x, y, z in pts], np.int64)
    g = np.random.default_rng(12)
    data = g.normal(size=(R, 3))
    grad, mask = O.lumisphere_tv_grad_sparse(links, data, lumi_cells, 1, [1.0], [1.0], 0.4, 0.9)
    ref, rmask = O.tv_grad_sparse(links, data, tv_cells, 0, 3, 0.4, True, False)
    assert np.abs(grad - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.array_equal(mask, rmask) and mask.sum() > 100
    data27 = g.normal(size=(R, 27))
    e0 = np.zeros(9)
    e0[0] = 1.0
    grad, mask = O.lumisphere_tv_grad_sparse(links, data27, lumi_cells, 9, e0, e0, 0.4, 0.9)
    ref3, rmask = O.tv_grad_sparse(links, data27[:, [0, 9, 18]], tv_cells, 0, 3, 0.4, True,
                                   False)
    assert np.abs(grad[:, [0, 9, 18]] - ref3).max() <= 1e-13 * np.abs(ref3).max()
    others = np.delete(np.arange(27), [0, 9, 18])
    assert (grad[:, others] == 0).all()
    assert np.array_equal(mask, rmask)


def test_sgd_closed_form_known_answer():
    """data - lr * g per selected row, lr_last on the last column, the
    gradient zeroed only where applied: a hand-computed case."""
    data = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]])
    grad = np.array([[0.5, 1.0, -2.0], [1.0, 1.0, 1.0]])
    d, g = O.sgd_step(data, grad, [0], 0.1, 0.25)
    assert np.allclose(d, [[0.95, 1.9, 3.5], [4.0, 5.0, 6.0]], rtol=0, atol=1e-15)
    assert np.array_equal(g, [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]])
