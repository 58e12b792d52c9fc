"""Pins of the MSI background oracle (render_lerp_kernel_cuvol.cu:386-607)
and of the MSI total variation oracle (loss_kernel.cu:248-333,606-641).

The background pins use a grid whose frame is the world frame (offset
N / 2 - 0.5, scaling N / 2: the grid's [-1, 1] sphere frame is the world's, so
a unit world direction has world step 1) and rays from points near the
centre, so every sample lies on the ray's own direction from the centre and
the first radius is 1.  Then sample i sits at radius r_i = n / (n - i - 0.5)
(n = n_steps), its step is 1 / r_{i-1} - 1 / r_i (= 1 / n after the first,
0.5 / n for the first), and its layer is (1 - 1 / r_i) nlayers - 0.5 =
(i + 0.5) nlayers / n - 0.5: closed forms independent of the oracle's
geometry code."""
import math

import numpy as np
import torch

import oracle.oracle as O


def _world_grid(N=8):
    links = np.full((N, N, N), -1, np.int32)
    Nf = np.array([N, N, N], np.float64)
    return O.Grid(links, np.zeros(0), np.zeros((0, 27)), 9, offset=0.5 * Nf - 0.5,
                  scaling=0.5 * Nf)


def _bg(reso, nl, fill):
    """Background with every column linked in raster order and data from
    fill(x, y, z) -> (r, g, b, sigma) (stored values)."""
    links = np.arange(2 * reso * reso, dtype=np.int32).reshape(2 * reso, reso)
    data = np.zeros((2 * reso * reso, nl, 4))
    for x in range(2 * reso):
        for y in range(reso):
            for z in range(nl):
                data[x * reso + y, z] = fill(x, y, z)
    return O.Background(links, data)


def _render(grid, bgm, o, d, opt, log_t=None):
    n = len(o)
    lt = torch.zeros(n, dtype=torch.float64) if log_t is None else torch.as_tensor(log_t)
    return O.background_fp64(grid, bgm, o, d, opt, torch.tensor(bgm.data), lt).numpy()


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def test_uniform_background_telescopes():
    """Uniform sigma and colour: the steps telescope, so the transmittance
    left is T0 exp(-sigma (1 - 1 / (2 n))) (the last sample is at radius 2 n)
    whatever the direction, and the colour is (T0 - T) c + T bg; rays whose
    grid log transmittance is below -25 get nothing."""
    grid = _world_grid()
    sig, rgb_raw = 0.7, np.array([0.4, -0.9, 1.3])
    bgm = _bg(6, 5, lambda x, y, z: (*rgb_raw, sig))
    rng = np.random.default_rng(0)
    d = _unit(rng.normal(size=(40, 3)))
    o = np.zeros((40, 3))  # from the centre
    opt = O.Opts(background_brightness=0.35, step_size=0.5, stop_thresh=0.0)
    n_steps = int(5 / 0.5) + 2
    L0 = np.log(rng.uniform(0.2, 1.0, 40))
    L0[:3] = -30.0
    out = _render(grid, bgm, o, d, opt, L0)
    T0 = np.exp(L0)
    T = T0 * np.exp(-sig * (1.0 - 1.0 / (2 * n_steps)))
    c = np.maximum(rgb_raw * O.SH_C0 + 0.5, 0.0)
    ref = (T0 - T)[:, None] * c + T[:, None] * 0.35
    ref[:3] = 0.0
    assert np.abs(out - ref).max() < 1e-12
    assert np.all(out[:3] == 0.0)
    # A line passing 1.5 from the centre starts at radius 1.5 + 1e-3: the
    # steps telescope from there.
    o2 = np.array([[0.0, 1.5, -0.4], [1.5, 0.3, 0.0]])
    d2 = np.array([[0.0, 0.0, 1.0], [0.0, 1.0, 0.0]])
    out2 = _render(grid, bgm, o2, d2, opt)
    T2 = np.exp(-sig * (1.0 / 1.501 - 1.0 / (2 * n_steps)))
    assert np.abs(out2 - ((1.0 - T2) * c + T2 * 0.35)).max() < 1e-12


def test_layer_schedule_closed_form():
    """Colour linear in the layer index, density uniform: the composite is
    sum_i T_{i-1} (1 - e^{-sigma s_i}) c(z_i), s_0 = 0.5 / n, s_i = 1 / n, z_i
    = clamp((i + 0.5) nlayers / n - 0.5, 0, nlayers - 1), c linear in z
    (trilinear is exact on a linear field)."""
    grid = _world_grid()
    nl, sig = 7, 2.5
    a, b = -1.2, 0.55
    bgm = _bg(4, nl, lambda x, y, z: (a + b * z, 0.0, 0.0, sig))
    d = _unit([[0.3, 0.2, 0.9], [-0.5, 0.1, -0.4], [0.0, 1.0, 0.0]])
    o = np.zeros((3, 3))
    for step in (0.5, 0.7):
        opt = O.Opts(background_brightness=0.0, step_size=step, stop_thresh=0.0)
        n = int(np.float32(nl) / np.float32(step)) + 2
        out = _render(grid, bgm, o, d, opt)
        T, acc = 1.0, 0.0
        for i in range(n):
            s = (0.5 if i == 0 else 1.0) / n
            z = min(max((i + 0.5) * nl / n - 0.5, 0.0), nl - 1.0)
            c = max((a + b * z) * O.SH_C0 + 0.5, 0.0)
            acc += T * (1.0 - math.exp(-sig * s)) * c
            T *= math.exp(-sig * s)
        assert np.abs(out[:, 0] - acc).max() < 1e-12, step


def test_equirect_orientation():
    """Each column's colour encodes its pixel (r = x / 2 reso, g = y / reso):
    a ray along +z reads pixel x = reso, y = reso / 2; +x reads x = 1.5 reso;
    -x reads x = 0.5 reso; +y (latitude pi / 2) reads row y = 0; -z
    (longitude pi) reads x = 2 reso, wrapped to column 0."""
    reso, nl = 8, 3
    grid = _world_grid()
    enc = lambda x, y: ((x / (2 * reso) - 0.5) / O.SH_C0, (y / reso - 0.5) / O.SH_C0)
    bgm = _bg(reso, nl, lambda x, y, z: (*enc(x, y), 0.0, 1.0))
    s45 = math.sqrt(0.5)
    cases = {(0, 0, 1): (reso, reso // 2), (1, 0, 0): (1.5 * reso, reso // 2),
             (-1, 0, 0): (0.5 * reso, reso // 2), (0, 1, 0): (reso, 0), (0, 0, -1): (0, reso // 2),
             (0, s45, s45): (reso, reso / 4)}  # latitude +45 degrees: a quarter down
    d = np.array(list(cases.keys()), np.float64)
    opt = O.Opts(background_brightness=0.0, step_size=0.5, stop_thresh=0.0)
    out = _render(grid, bgm, np.zeros_like(d), d, opt)
    alpha = 1.0 - math.exp(-(1.0 - 1.0 / (2 * (int(nl / 0.5) + 2))))
    for k, (px, py) in enumerate(cases.values()):
        assert abs(out[k, 0] / alpha - px / (2 * reso)) < 1e-9, (d[k], out[k])
        assert abs(out[k, 1] / alpha - py / reso) < 1e-9, (d[k], out[k])
    # Latitude -80 degrees falls between the last row and the first (the
    # latitude index wraps like the longitude's): y = reso (0.5 + 80 / 180).
    lat = math.radians(-80.0)
    dd = np.array([[0.0, math.sin(lat), math.cos(lat)]])
    y = reso * (0.5 + 80.0 / 180.0)
    f = y - (reso - 1)
    out = _render(grid, bgm, np.zeros_like(dd), dd, opt)
    assert abs(out[0, 1] / alpha - (1.0 - f) * (reso - 1) / reso) < 1e-9, out


def test_stop_and_empty_columns():
    """The march stops after the sample whose transmittance falls below
    stop_thresh (later samples add nothing), and empty columns (link -1) read
    as 0: a background whose every column is empty passes T0 bg through."""
    grid = _world_grid()
    bgm = _bg(4, 6, lambda x, y, z: (0.3, 0.3, 0.3, 40.0))
    d = _unit([[0.2, 0.3, 0.9]])
    o = np.zeros((1, 3))
    opt = O.Opts(background_brightness=0.5, step_size=0.5, stop_thresh=1e-3)
    out = _render(grid, bgm, o, d, opt)
    n = int(6 / 0.5) + 2
    T, acc = 1.0, 0.0
    c = 0.3 * O.SH_C0 + 0.5
    for i in range(n):
        s = (0.5 if i == 0 else 1.0) / n
        acc += T * (1.0 - math.exp(-40.0 * s)) * c
        T *= math.exp(-40.0 * s)
        if T < 1e-3:
            break
    assert i < n - 1
    assert abs(out[0, 0] - (acc + T * 0.5)) < 1e-12
    empty = O.Background(np.full_like(bgm.links, -1), bgm.data)
    out = _render(grid, empty, o, d, opt, np.array([math.log(0.6)]))
    assert np.abs(out - 0.6 * 0.5).max() < 1e-12


def test_stopped_grid_rays_see_no_background():
    """A ray the grid's stop test ends leaves log transmittance -1e3
    (render_lerp_kernel_cuvol.cu:103-105), so the background adds nothing to
    it; a ray that misses the grid sees the whole background."""
    N = 8
    links = np.arange(N ** 3, dtype=np.int32).reshape(N, N, N)
    Nf = np.array([N, N, N], np.float64)
    grid = O.Grid(links, np.full(N ** 3, 50.0), np.zeros((N ** 3, 27)), 9,
                  offset=0.5 * Nf - 0.5, scaling=0.5 * Nf)
    bgm = _bg(4, 3, lambda x, y, z: (1.0, 1.0, 1.0, 2.0))
    o = np.array([[0.0, 0.0, -3.0], [0.0, 3.0, -3.0]], np.float32)
    # (a direction component of exactly 0 skips that axis's slab test, so the
    # missing ray is slanted on every axis)
    d = np.array([[0.0, 0.0, 1.0], _unit([0.05, 0.1, 1.0])], np.float32)
    opt = O.Opts(background_brightness=0.5, step_size=0.5, stop_thresh=1e-4)
    dens, sh = torch.tensor(grid.density), torch.tensor(grid.sh)
    fg, log_t = O.render_fp64(grid, o, d, O.Opts(background_brightness=0.0, stop_thresh=1e-4),
                              dens, sh, return_log_t=True)
    assert log_t[0].item() == -1e3 and log_t[1].item() == 0.0
    rgb = O.render_bg_fp64(grid, bgm, o, d, opt, dens, sh, torch.tensor(bgm.data)).numpy()
    assert np.array_equal(rgb[0], fg[0].numpy())
    only_bg = _render(grid, bgm, o[1:], d[1:], opt)
    assert np.abs(rgb[1] - only_bg[0]).max() < 1e-12 and rgb[1, 0] > 0.5


def _msi_energy_grad(links, msi, cells, scale, scale_last):
    """The gradient the MSI TV kernel adds, by autograd: for each listed cell,
    idelta (detached) times the gradient of 0.5 (fx dx^2 + fy dy^2 + fz dz^2),
    the differences gathered with wrapping neighbours and the layer-edge rule,
    so autograd decides which values receive which terms."""
    links = torch.from_numpy(np.asarray(links, np.int64))
    v = torch.tensor(msi, requires_grad=True)
    rows, nl, C = msi.shape
    reso = links.shape[1]
    cells = torch.from_numpy(np.asarray(cells, np.int64))
    z = cells % nl
    x = (cells // nl) // reso
    y = (cells // nl) % reso
    nxt = lambda a, m: torch.where(a == m - 1, torch.zeros_like(a), a + 1)
    zero = torch.zeros((), dtype=torch.float64)

    def val(lk, zz, ok=None):
        good = lk >= 0 if ok is None else (lk >= 0) & ok
        r = torch.where(good, lk, torch.zeros_like(lk))
        zc = torch.clamp(zz, max=nl - 1)
        return torch.where(good[:, None], v[r, zc], zero)

    l00, l01, l10 = links[x, y], links[x, nxt(y, reso)], links[nxt(x, 2 * reso), y]
    v00 = val(l00, z)
    has_z = (z + 1 < nl)
    vz_in = val(l00, z + 1, has_z)
    edge = torch.stack([v00[:, c] if c < C - 1 else torch.zeros_like(v00[:, c])
                        for c in range(C)], 1)
    vz = torch.where((has_z & (l00 >= 0))[:, None], vz_in, edge)
    dx, dy, dz = val(l10, z) - v00, val(l01, z) - v00, vz - v00
    sc = torch.full((C,), scale, dtype=torch.float64)
    sc[-1] = scale_last
    idelta = (sc / len(cells) / torch.sqrt(1e-9 + dx * dx + dy * dy + dz * dz)).detach()
    f = (2 * reso / 256.0, reso / 256.0, nl / 256.0)
    q = 0.5 * (f[0] * dx * dx + f[1] * dy * dy + f[2] * dz * dz)
    (idelta * q).sum().backward()
    return v.grad.numpy()


def test_msi_tv_is_detached_normaliser_gradient():
    """The restated kernel equals idelta times the autograd gradient of the
    axis-weighted quadratic form of the differences, on a background with
    empty columns, cells at the image edges (wrap) and the top layer (the
    missing layer reads v00 for colours and 0 for sigma), with different
    colour and sigma scales; the mask is the set of layers with a nonzero
    term."""
    rng = np.random.default_rng(2)
    reso, nl, C = 4, 5, 4
    links = np.arange(2 * reso * reso, dtype=np.int32).reshape(2 * reso, reso)
    links[1, 2] = -1
    links[7, 0] = -1
    msi = rng.normal(size=(2 * reso * reso, nl, C))
    total = 2 * reso * reso * nl
    cells = np.concatenate([rng.choice(total, 60, replace=False),
                            [((2 * reso - 1) * reso + reso - 1) * nl + nl - 1,  # corner, top layer
                             (1 * reso + 2) * nl + 1]])  # the empty column
    g, mask = O.msi_tv_grad_sparse(links, msi, cells, 0.8, 0.3)
    ref = _msi_energy_grad(links, msi, cells, 0.8, 0.3)
    assert np.abs(g - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.array_equal(mask, np.abs(g).sum(-1) > 0)
    assert not mask[links[links >= 0].max() + 1:].any()
    # the sigma channel's top-layer rule: the missing layer counts as 0
    top = np.array([((2 * reso - 1) * reso + reso - 1) * nl + nl - 1])
    g1, _ = O.msi_tv_grad_sparse(links, msi, top, 0.0, 1.0)
    r = links[2 * reso - 1, reso - 1]
    assert g1[r, nl - 1, 3] != 0.0 and np.all(g1[..., :3] == 0.0)
