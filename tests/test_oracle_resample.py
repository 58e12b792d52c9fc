"""Pins of the resample oracle pieces (dilate, mask_to_links, resample_axes,
MarchRef.grid_weight_render; svox2.py:1223-1424, misc_kernel.cu:20-48,184-334)
by known answers, a library routine and closed forms."""
import numpy as np
import pytest
import scipy.ndimage as ndi

import oracle.oracle as O


def test_dilate_known_answers_and_library():
    m = np.zeros((5, 6, 7), bool)
    m[2, 3, 3] = True
    assert O.dilate(m).sum() == 27  # interior voxel: the full 3x3x3 box
    m[:] = False
    m[0, 0, 6] = True
    assert O.dilate(m).sum() == 8   # corner voxel: the box clipped to 2x2x2
    rng = np.random.default_rng(0)
    for shape, p in (((9, 8, 12), 0.05), ((3, 17, 4), 0.2), ((1, 5, 5), 0.1)):
        r = rng.random(shape) < p
        want = ndi.binary_dilation(r, structure=np.ones((3, 3, 3), bool), border_value=0)
        assert np.array_equal(O.dilate(r), want)
        # two passes = one box of radius 2 (clipped boxes compose)
        want2 = ndi.binary_dilation(r, structure=np.ones((5, 5, 5), bool), border_value=0)
        assert np.array_equal(O.dilate(O.dilate(r)), want2)


def test_mask_to_links_orders():
    rng = np.random.default_rng(1)
    m = rng.random((8, 8, 8)) < 0.3
    links, kept = O.mask_to_links(m)
    assert np.array_equal(kept, np.flatnonzero(m))            # raster order
    assert np.array_equal(links.reshape(-1)[kept], np.arange(kept.size))
    assert np.all(links[~m] == -1)
    lz, kz = O.mask_to_links(m, z_order=True)
    codes = np.array([O.morton_code_3(*np.unravel_index(c, m.shape)) for c in kz])
    assert np.all(np.diff(codes) > 0) and set(kz) == set(kept)  # Morton order
    assert np.array_equal(lz.reshape(-1)[kz], np.arange(kz.size))
    assert np.all(lz[~m] == -1)


def test_resample_axes_closed_forms():
    # same size: cell centres map to themselves
    assert np.array_equal(O.resample_axes((6, 6, 6), (6, 6, 6)), np.tile(np.arange(6.0), 3))
    # halving: new cell i covers old cells 2i, 2i+1 -> centre 2i + 1/2
    ax = O.resample_axes((8, 8, 8), (4, 4, 4))
    assert np.array_equal(ax[:4], [0.5, 2.5, 4.5, 6.5])
    # doubling: new centres at (2i + 1)/4 - 1/2 in old coordinates
    ax = O.resample_axes((4, 4, 4), (8, 8, 8))
    assert np.allclose(ax[:8], (2 * np.arange(8) + 1) / 4 - 0.5, atol=1e-6)
    pts = O.resample_points(ax, (8, 8, 8), np.array([0, 7, 8 * 8 * 8 - 1]))
    assert np.allclose(pts, [[-0.25] * 3, [-0.25, -0.25, 3.25], [3.25] * 3], atol=1e-6)


@pytest.mark.parametrize("sigma,stop", [(0.5, 1e-7), (4.0, 0.2)])
def test_grid_weight_render_axis_ray(sigma, stop):
    """One pixel whose ray runs along +x through a uniform volume, world =
    volume coordinates: samples at x_k = -1/2 + k/2 (clamped to 0), world step
    1/2, weights a^k (1 - a), a = exp(-sigma/2), the ray stopping after the
    first sample with a^(k+1) < stop.  Each sample raises the 8 corners of
    its cell (l = min(trunc x, 6), 3, 4); weights fall with k, so each corner
    holds the weight of the first sample whose cell touches it."""
    N = 8
    data = np.full((N, N, N), sigma, np.float32)
    # camera: +z (forward) -> world +x, +x -> world +y, +y -> world +z
    c2w = np.array([[0, 0, 1, -3.0], [1, 0, 0, 3.25], [0, 1, 0, 4.5]], np.float32)
    out = O.march().grid_weight_render(data, c2w, 1.0, 1.0, 0.5, 0.5, 1, 1, 0.5, stop, False,
                                       [0, 0, 0], [1, 1, 1])
    a = np.exp(-sigma * 0.5)
    want = np.zeros((N, N, N), np.float64)
    for k in range(2 * N + 1):  # x from -0.5 to 7.5
        x = max(-0.5 + 0.5 * k, 0.0)
        lx = min(int(x), N - 2)
        w = a ** k * (1 - a)
        for cx in (lx, lx + 1):
            for cy in (3, 4):
                for cz in (4, 5):
                    want[cx, cy, cz] = max(want[cx, cy, cz], w)
        if a ** (k + 1) < stop:
            break
    assert np.allclose(out, want, rtol=2e-6, atol=0)
    # last_sample_opaque: the last sample's weight is 1 at its cell's corners
    out2 = O.march().grid_weight_render(data, c2w, 1.0, 1.0, 0.5, 0.5, 1, 1, 0.5, 0.0, True,
                                        [0, 0, 0], [1, 1, 1])
    assert np.all(out2[N - 2:, 3:5, 4:6] == 1.0)
