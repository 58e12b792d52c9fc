"""Scheduling logic of the serving loop (no GPU): both tile plans cover every
pixel of every request exactly once, in order, within the per-launch tile cap;
FIFO serves requests in admission order, round-robin lets small requests
finish before large ones; the array-built plan the server uses equals the
per-tile plan."""
import pytest
import math

import numpy as np

from svox2_amd.serve import RenderRequest, plan_runs, plan_tiles


class _Cam:
    def __init__(self, w, h):
        self.width, self.height = w, h


def _req(w, h):
    return RenderRequest(grid=0, camera=_Cam(w, h))


@pytest.mark.parametrize("policy", ["fifo", "round_robin"])
def test_plan_covers_each_pixel_once(policy):
    reqs = [_req(100, 37), _req(5, 5), _req(64, 64), _req(1, 1)]
    seen = {id(r): [] for r in reqs}
    steps = 0
    while any(r.next_pixel < r.pixels for r in reqs):
        plan = plan_tiles(reqs, max_tiles=3, tile_pixels=500, policy=policy)
        assert 0 < len(plan) <= 3
        for r, p0, n in plan:
            assert 0 < n <= 500
            seen[id(r)].append((p0, n))
        steps += 1
        assert steps < 100
    for r in reqs:
        spans = seen[id(r)]
        assert spans[0][0] == 0
        for (a, n), (b, _) in zip(spans, spans[1:]):
            assert b == a + n
        assert spans[-1][0] + spans[-1][1] == r.pixels


def test_round_robin_lets_small_requests_finish_first():
    big, small = _req(1024, 1024), _req(128, 128)
    done_at = {}
    step = 0
    while big.next_pixel < big.pixels or small.next_pixel < small.pixels:
        plan_tiles([big, small], max_tiles=4, tile_pixels=4096, policy="round_robin")
        step += 1
        for name, r in (("big", big), ("small", small)):
            if r.next_pixel >= r.pixels and name not in done_at:
                done_at[name] = step
    assert done_at["small"] == math.ceil(128 * 128 / 4096 / 2)
    assert done_at["big"] > done_at["small"]
    assert done_at["big"] == math.ceil((1024 * 1024 + 128 * 128) / 4096 / 4)


def test_fifo_serves_in_admission_order():
    first, second = _req(300, 10), _req(10, 10)
    plan = plan_tiles([first, second], max_tiles=3, tile_pixels=1000)
    assert [(r is first, p0, n) for r, p0, n in plan] == [(True, 0, 1000), (True, 1000, 1000),
                                                          (True, 2000, 1000)]
    plan = plan_tiles([first, second], max_tiles=3, tile_pixels=1000)
    assert [(r is first, n) for r, _, n in plan] == [(False, 100)]


def test_tile_record_matches_c_layout():
    import ctypes
    from svox2_amd import _C, serve
    t = _C._Tile
    assert serve._TILE_DTYPE.itemsize == ctypes.sizeof(t)
    for name in ("grid", "camera", "pixel0", "n_pixels", "rgb_out"):
        assert serve._TILE_DTYPE.fields[name][1] == getattr(t, name).offset


@pytest.mark.parametrize("policy", ["fifo", "round_robin"])
def test_plan_runs_equal_plan_tiles(policy):
    """The array-built plan (runs of one request's consecutive tiles) is the
    per-tile plan, with the same request progress."""
    g = np.random.default_rng(5)
    for trial in range(40):
        sides = g.integers(1, 900, size=g.integers(1, 6))
        a = [_req(int(s), int(s)) for s in sides]
        b = [_req(int(s), int(s)) for s in sides]
        for r, q in zip(a, b):
            r.next_pixel = q.next_pixel = int(g.integers(0, r.pixels))
        tp, mt = int(g.integers(100, 5000)), int(g.integers(1, 40))
        plan = plan_tiles(a, max_tiles=mt, tile_pixels=tp, policy=policy)
        runs = plan_runs(b, max_tiles=mt, tile_pixels=tp, policy=policy)
        flat = [(b.index(r), int(p), int(n)) for r, p0, nn in runs for p, n in zip(p0, nn)]
        assert flat == [(a.index(r), p, n) for r, p, n in plan]
        assert [r.next_pixel for r in a] == [r.next_pixel for r in b]
