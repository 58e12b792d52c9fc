"""Pins of the svox1 oracle (render_svox1_kernel.cu): the FP32 walk
(ref_svox1_samples) against an exact float64 cell traversal, the FP64
compositing against closed forms, and the restated backward against the
autograd gradient of the compositing where no stop occurs."""
import math

import numpy as np
import torch

import oracle.oracle as O
from svox2_amd import scenes


def _dense(N, dens, sh, basis_dim=9):
    links = np.arange(N ** 3, dtype=np.int32).reshape(N, N, N)
    Nf = np.array([N, N, N], np.float64)
    return O.Grid(links, np.asarray(dens, np.float64), np.asarray(sh, np.float64), basis_dim,
                  offset=0.5 * Nf - 0.5, scaling=0.5 * Nf)


def _rays(n, seed, radius=1.7):
    g = torch.Generator().manual_seed(seed)
    o, d = scenes.sphere_rays(n, g, radius=radius, target=0.5)
    return o.numpy(), d.numpy()


def _exact_cells(grid, o, d):
    """Cells (flat raster index) a ray crosses in the walk's frame (grid
    coordinates + 0.5, box [0, size - 1]), in order, with their chord lengths
    (grid units), from t = max(0, entry): float64, plane crossings."""
    N = np.asarray(grid.size, np.float64)
    og = o.astype(np.float64) * grid.scaling + grid.offset + 0.5
    dg = d.astype(np.float64) * grid.scaling
    dg /= np.linalg.norm(dg)
    with np.errstate(divide="ignore"):
        inv = 1.0 / dg
    t1, t2 = (0.0 - og) * inv, (N - 1.0 - og) * inv
    t0, t_end = max(0.0, np.minimum(t1, t2).max()), np.maximum(t1, t2).min()
    if t0 >= t_end:
        return [], []
    ts = [t0, t_end]
    for a in range(3):
        for k in range(int(N[a]) + 1):
            t = (k - og[a]) * inv[a]
            if t0 < t < t_end:
                ts.append(t)
    ts = np.unique(ts)
    cells, chords = [], []
    for a, b in zip(ts[:-1], ts[1:]):
        p = og + 0.5 * (a + b) * dg
        c = np.minimum(np.floor(p).astype(int), np.asarray(grid.size) - 1)
        cells.append(int((c[0] * grid.size[1] + c[1]) * grid.size[2] + c[2]))
        chords.append(b - a)
    return cells, chords


def test_walk_follows_the_cells():
    """On a dense grid every sample is an occupied cell: the visited cells are
    the crossed cells in order, minus only cells whose chord is below the
    1e-2 nudge; a cell entered from the cell before it is entered 1e-2 past
    their shared face, so its step (grid units; world step = grid step /
    world scale) is the rest of its chord plus 1e-2, the whole chord."""
    N = 12
    grid = _dense(N, np.full(N ** 3, 0.01), np.zeros((N ** 3, 27)))
    o, d = _rays(60, 1)
    opt = O.Opts(sigma_thresh=0.0, stop_thresh=0.0)
    smp = O.march().svox1_samples(grid, o, d, opt)
    scale = np.linalg.norm(d * grid.scaling, axis=1)  # grid units per world unit
    checked = 0
    for r in range(len(o)):
        sel = smp["ray"] == r
        cells, chords = _exact_cells(grid, o[r], d[r])
        # The box is [0, size - 1]: a ray entering through a far face first
        # samples cell size - 1 at that face (a zero chord), then moves on.
        face = [c for c in smp["link"][sel] if (N - 1) in (c // (N * N), c // N % N, c % N)]
        assert len(face) <= 1
        visited = [c for c in smp["link"][sel] if c not in face]
        it = iter(cells)
        assert all(c in it for c in visited), r  # an ordered subsequence
        must = {c for c, ch in zip(cells, chords) if ch > 0.011}
        assert must <= set(visited), r
        steps = smp["dt"][sel].astype(np.float64) * scale[r]
        steps = steps[len(face):]
        for prev, c, s in zip(visited[:-1], visited[1:], steps[1:]):
            k = cells.index(c)
            if k == 0 or cells[k - 1] != prev:
                continue
            # entered 1e-2 past the shared face: the rest of the chord plus 1e-2
            assert abs(s - chords[k]) < 1e-4, (r, s, chords[k])
            checked += 1
    assert checked > 300


def test_walk_skips_empty_blocks():
    """On a sparse grid with accelerated links (empty blocks of 2^k cells,
    accel_dist_prop), the occupied cells visited are the occupied crossed
    cells in order, minus only those with a chord below the nudge; a cell
    entered after an empty cell is entered 1e-2 past its face (the skip ends
    1e-2 past the block's exit face), so its step is its whole chord too."""
    N = 16
    rng = np.random.default_rng(8)
    occ = np.zeros((N, N, N), bool)
    occ[4:9, 3:12, 5:10] = rng.random((5, 9, 5)) < 0.6
    occ[11:14, 10:14, 2:6] = True
    links = np.full((N, N, N), -1, np.int32)
    links[occ] = np.arange(occ.sum(), dtype=np.int32)
    acc = O.accel_dist_prop(links)
    assert acc.min() < -2  # multi-level blocks
    Nf = np.array([N, N, N], np.float64)
    rows = int(occ.sum())
    grid = O.Grid(acc, np.full(rows, 0.01), np.zeros((rows, 27)), 9, offset=0.5 * Nf - 0.5,
                  scaling=0.5 * Nf)
    o, d = _rays(200, 3)
    smp = O.march().svox1_samples(grid, o, d, O.Opts(sigma_thresh=0.0, stop_thresh=0.0))
    flat_links = links.reshape(-1)
    row_cell = np.flatnonzero(flat_links >= 0)  # row -> flat cell (rows in raster order)
    scale = np.linalg.norm(d * grid.scaling, axis=1)
    checked = 0
    for r in range(len(o)):
        sel = smp["ray"] == r
        cells, chords = _exact_cells(grid, o[r], d[r])
        visited = [int(row_cell[k]) for k in smp["link"][sel]]
        occ_cells = [c for c in cells if flat_links[c] >= 0]
        it = iter(occ_cells)
        assert all(c in it for c in visited), r
        assert {c for c, ch in zip(cells, chords) if flat_links[c] >= 0 and ch > 0.011} \
            <= set(visited), r
        steps = smp["dt"][sel].astype(np.float64) * scale[r]
        for j, (c, s) in enumerate(zip(visited, steps)):
            k = cells.index(c)
            if k == 0 or chords[k] < 0.011:
                continue
            prev = cells[k - 1]
            if chords[k - 1] < 0.011:  # a thinner cell before it: landed past it
                continue
            if flat_links[prev] < 0 or (j > 0 and visited[j - 1] == prev):
                assert abs(s - chords[k]) < 1e-4, (r, s, chords[k])
                checked += 1
    assert checked > 200


def test_uniform_compositing_closed_form():
    """Uniform density and colour: the weights telescope to 1 - T, T =
    exp(-sigma sum dt), so C = c (1 - T) + T bg; with the stop, the colour
    of a ray stopped at sample j is c (1 - T_j) / (1 - T_{j-1})."""
    N = 10
    dc = np.array([0.6, -0.4, 1.1])
    sh = np.zeros((N ** 3, 27))
    sh[:, [0, 9, 18]] = dc
    c = np.maximum(dc * O.SH_C0 + 0.5, 0.0)
    o, d = _rays(80, 2)
    for sigma, stop in ((0.3, 0.0), (4.0, 0.05)):
        g = _dense(N, np.full(N ** 3, sigma), sh)
        opt = O.Opts(background_brightness=0.45, sigma_thresh=0.0, stop_thresh=stop)
        smp = O.march().svox1_samples(g, o, d, opt)
        rgb = O.svox1_fp64(g, o, d, opt, torch.tensor(g.density), torch.tensor(g.sh),
                           samples=smp).numpy()
        n_stop = 0
        for r in range(len(o)):
            dts = smp["dt"][smp["ray"] == r].astype(np.float64)
            fl = smp["flag"][smp["ray"] == r]
            if dts.size == 0:
                assert np.allclose(rgb[r], 0.45, atol=0, rtol=0)
                continue
            D = np.cumsum(dts) * sigma
            if fl[-1] & 2:
                n_stop += 1
                ref = c * (1 - math.exp(-D[-1])) / (1 - math.exp(-(D[-2] if D.size > 1 else 0.0)))
            else:
                T = math.exp(-D[-1])
                ref = c * (1 - T) + T * 0.45
            assert np.abs(rgb[r] - ref).max() < 1e-12
        assert (n_stop > 10) == (stop > 0)
    # An opaque last sample (density + 1e9) takes all the transmittance left:
    # C = c, no background.
    g = _dense(N, np.full(N ** 3, 0.2), sh)
    opt = O.Opts(background_brightness=0.45, sigma_thresh=0.0, stop_thresh=0.0,
                 last_sample_opaque=True)
    smp = O.march().svox1_samples(g, o, d, opt)
    rgb = O.svox1_fp64(g, o, d, opt, torch.tensor(g.density), torch.tensor(g.sh), samples=smp)
    opaque = np.unique(smp["ray"][(smp["flag"] & 8) != 0])
    assert opaque.size > 40
    assert np.abs(rgb.numpy()[opaque] - c).max() < 1e-12


def test_backward_is_autograd_gradient_without_stop():
    """With no stop the restated backward equals the autograd gradient of the
    compositing (density and SH), for rays through a random grid with empty
    links (skipped blocks) and rays that reach the background; its mask is
    the rows of the counted samples."""
    gen = torch.Generator().manual_seed(4)
    sc = scenes.make_blob_grid(16, scenes.SPARSE256, gen)
    links = sc.links.numpy().astype(np.int32)
    rows = int(links.max()) + 1
    rng = np.random.default_rng(5)
    dens = rng.uniform(-0.2, 1.5, rows)
    sh = rng.normal(0, 0.5, (rows, 27))
    Nf = np.array(links.shape, np.float64)
    g = O.Grid(links, dens, sh, 9, offset=0.5 * Nf - 0.5, scaling=0.5 * Nf)
    o, d = _rays(150, 6)
    opt = O.Opts(background_brightness=0.7, sigma_thresh=0.05, stop_thresh=0.0)
    smp = O.march().svox1_samples(g, o, d, opt)
    assert (smp["flag"] & 1).any() and not (smp["flag"] & 1).all()
    dt = torch.tensor(g.density, requires_grad=True)
    st = torch.tensor(g.sh, requires_grad=True)
    rgb = O.svox1_fp64(g, o, d, opt, dt, st, samples=smp)
    gout = rng.normal(size=(len(o), 3))
    (rgb * torch.tensor(gout)).sum().backward()
    gd, gs, mask = O.svox1_grads_fp64(g, o, d, opt, gout, rgb.detach().numpy(), samples=smp)
    assert np.abs(gd - dt.grad.numpy()).max() <= 1e-10 * np.abs(gd).max()
    assert np.abs(gs - st.grad.numpy()).max() <= 1e-10 * np.abs(gs).max()
    assert np.array_equal(np.flatnonzero(mask),
                          np.unique(smp["link"][(smp["flag"] & 1) != 0]))
