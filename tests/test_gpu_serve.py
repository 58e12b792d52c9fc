"""Serving loop on the GPU: every request's image is the same alone (one
volume_render_image call) and in the mix (tiles of several requests, grids and
storage kinds per graph-replayed launch), bit for bit: each pixel runs the same
march code.  Tolerance: exact (0)."""
import numpy as np
import pytest
import torch

from svox2_amd import scenes
from svox2_amd.serve import RenderRequest, RenderServer, serve_one_per_launch

pytestmark = pytest.mark.gpu


def _grids(dev):
    sc64, _, _, _ = scenes.workload_dense64(seed=0)
    sc96, cams, gen = scenes.workload_blobs(96, scenes.SPARSE256, seed=1, n_cameras=8,
                                            device=dev)
    out = []
    for sc, dt in ((sc64, torch.float32), (sc96, torch.bfloat16), (sc96, torch.float32)):
        g = scenes.to_sparse_grid(sc, device=dev, dtype=dt)
        g.opt.sigma_thresh, g.opt.stop_thresh = 1e-8, 1e-7
        out.append(g)
    return out, cams


def _requests(cams, n, seed):
    g = np.random.default_rng(seed)
    sides = [(17, 9), (64, 64), (100, 37), (128, 128), (5, 200)]
    reqs = []
    for i in range(n):
        w, h = sides[g.integers(len(sides))]
        cam = cams.camera(int(g.integers(cams.c2w.shape[0])), width=w, height=h)
        reqs.append(RenderRequest(grid=int(g.integers(3)), camera=cam, arrival=0.0))
    return reqs


@pytest.mark.parametrize("use_graph", [True, False])
def test_mix_equals_alone(gpu, use_graph):
    grids, cams = _grids(gpu)
    reqs = _requests(cams, 14, seed=4)
    # A small budget, few slots, small tiles and few tiles per launch force
    # waiting admissions, multi-tile requests, partial tiles and mixing.
    server = RenderServer(grids, hbm_budget=200_000, tile_pixels=700, max_tiles=5,
                          max_active=4, use_graph=use_graph)
    stats = server.run(reqs, realtime=False)
    assert stats["requests"] == len(reqs)
    assert server.steps > len(reqs) // 4
    for r in reqs:
        with torch.no_grad():
            alone = grids[r.grid].volume_render_image(r.camera)
        assert torch.equal(r.image, alone), (r.grid, r.camera.width, r.camera.height)


@pytest.mark.parametrize("concurrent", [False, True])
def test_idle_requests_go_direct(gpu, concurrent):
    """Requests that arrive on an idle server are rendered by one direct
    image launch each (no tile step), bit for bit as alone; with the tile
    steps switched off for them (idle_direct=False) the same requests go
    through tiles and still match."""
    grids, cams = _grids(gpu)
    reqs = _requests(cams, 5, seed=6)
    for i, r in enumerate(reqs):
        r.arrival = 0.03 * i  # far apart: each finds the device idle
    server = RenderServer(grids, tile_pixels=700, max_tiles=5, concurrent=concurrent)
    server.run(reqs, realtime=True)
    assert server.direct_launches == len(reqs) == server.steps
    tiled = RenderServer(grids, tile_pixels=700, max_tiles=5, concurrent=concurrent,
                         idle_direct=False)
    again = _requests(cams, 5, seed=6)
    tiled.run(again, realtime=False)
    assert tiled.direct_launches == 0
    for r, q in zip(reqs, again):
        with torch.no_grad():
            alone = grids[r.grid].volume_render_image(r.camera)
        assert torch.equal(r.image, alone) and torch.equal(q.image, alone)


def test_one_per_launch_baseline(gpu):
    grids, cams = _grids(gpu)
    reqs = _requests(cams, 4, seed=5)
    stats = serve_one_per_launch(grids, reqs, realtime=False)
    assert stats["requests"] == 4 and stats["p99_ms"] >= stats["p50_ms"] > 0
