"""Scheduling logic of the serving loop (no GPU): the round-robin tile plan
covers every pixel of every request exactly once, in order, within the
per-launch tile cap, and lets small requests finish before large ones."""
import math

from svox2_amd.serve import RenderRequest, plan_tiles


class _Cam:
    def __init__(self, w, h):
        self.width, self.height = w, h


def _req(w, h):
    return RenderRequest(grid=0, camera=_Cam(w, h))


def test_plan_covers_each_pixel_once():
    reqs = [_req(100, 37), _req(5, 5), _req(64, 64), _req(1, 1)]
    seen = {id(r): [] for r in reqs}
    steps = 0
    while any(r.next_pixel < r.pixels for r in reqs):
        plan = plan_tiles(reqs, max_tiles=3, tile_pixels=500)
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
        plan_tiles([big, small], max_tiles=4, tile_pixels=4096)
        step += 1
        for name, r in (("big", big), ("small", small)):
            if r.next_pixel >= r.pixels and name not in done_at:
                done_at[name] = step
    assert done_at["small"] == math.ceil(128 * 128 / 4096 / 2)
    assert done_at["big"] > done_at["small"]
    assert done_at["big"] == math.ceil((1024 * 1024 + 128 * 128) / 4096 / 4)
