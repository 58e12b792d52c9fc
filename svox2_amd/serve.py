"""Serving loop: several grids resident in HBM, image requests of uneven size
admitted by a memory budget, and their pixels batched into multi-request
launches that replay one captured HIP graph per launch group.

No counterpart in the reference, whose only render entry point is the
one-image-per-call ``SparseGrid.volume_render_image`` (svox2/svox2.py:1144-1179).

How a request flows:

* ``submit`` queues it (FIFO).
* Admission: it is admitted when a camera slot is free and its output image
  (H x W x 3 float32) fits in the HBM budget next to the images of the
  requests in flight.
* Continuous batching: every ``step`` builds one tile table per launch group
  (grids of one storage kind and basis) from all admitted requests. It hands
  out one tile of ``tile_pixels`` pixels per request per round, until the
  table holds ``max_tiles`` tiles. A request admitted during a step joins the
  next step, and small requests finish in a few steps even while large ones
  are in flight.
* Each step copies the tables to fixed device buffers and replays the group's
  graph. The graph is one ``svox2_volume_render_cuvol_tiles`` launch captured
  once, whose shape depends only on ``max_tiles`` and ``tile_pixels``.
* A request completes once all its tiles have run. Its image is then
  bit-identical to ``volume_render_image`` of the same grid and camera,
  because every pixel runs the same march code.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _C
from .svox2 import Camera, RenderOptions, SparseGrid


@dataclass(eq=False)
class RenderRequest:
    """One image request: grid index into the server's grids, a camera, and
    an arrival time (seconds on the caller's clock, for latency).  Requests
    compare by identity."""

    grid: int
    camera: Camera
    arrival: float = 0.0
    image: Optional[torch.Tensor] = None
    admitted: float = math.nan
    done: float = math.nan
    # scheduler state
    slot: int = -1
    next_pixel: int = 0

    @property
    def pixels(self) -> int:
        return int(self.camera.width) * int(self.camera.height)

    @property
    def image_bytes(self) -> int:
        return self.pixels * 3 * 4

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def plan_tiles(active: Sequence[RenderRequest], max_tiles: int,
               tile_pixels: int) -> List[tuple]:
    """Round-robin tile plan for one launch group: (request, pixel0, n)
    entries, one tile per request per round in admission order, at most
    max_tiles in total.  Advances each request's next_pixel."""
    plan: List[tuple] = []
    live = [r for r in active if r.next_pixel < r.pixels]
    while live and len(plan) < max_tiles:
        nxt = []
        for r in live:
            if len(plan) >= max_tiles:
                break
            n = min(tile_pixels, r.pixels - r.next_pixel)
            plan.append((r, r.next_pixel, n))
            r.next_pixel += n
            if r.next_pixel < r.pixels:
                nxt.append(r)
        live = nxt
    return plan


class _Group:
    """Grids of one storage dtype and basis_dim, rendered by one launch."""

    def __init__(self, storage: int, basis_dim: int, grid_ids: List[int]):
        self.storage = storage
        self.basis_dim = basis_dim
        self.grid_ids = grid_ids  # server grid index -> position in this group's table
        self.local: Dict[int, int] = {g: i for i, g in enumerate(grid_ids)}
        self.grids_dev: Optional[torch.Tensor] = None
        self.tiles_host: Optional[torch.Tensor] = None
        self.tiles_dev: Optional[torch.Tensor] = None
        self.count_host: Optional[torch.Tensor] = None
        self.count_dev: Optional[torch.Tensor] = None
        self.graph = None


class RenderServer:
    """Continuous-batching renderer over resident grids (see module doc).

    :param grids: resident SparseGrids (on the GPU; their cell maps are built
        once here).
    :param hbm_budget: bytes of output images in flight at most.
    :param tile_pixels: pixels per tile.
    :param max_tiles: tiles per launch and group.
    :param max_active: camera slots (requests in flight) at most.
    :param use_graph: replay a captured graph per launch group (else launch
        directly).
    """

    def __init__(self, grids: Sequence[SparseGrid], opt: Optional[RenderOptions] = None,
                 hbm_budget: int = 1 << 30, tile_pixels: int = 8192, max_tiles: int = 64,
                 max_active: int = 64, use_graph: bool = True):
        if not grids:
            raise ValueError("RenderServer needs at least one grid")
        self.grids = list(grids)
        self.device = self.grids[0].density_data.device
        if self.device.type != "cuda":
            raise RuntimeError("RenderServer: grids must be on the GPU")
        self.opt = opt or self.grids[0].opt
        self.hbm_budget = int(hbm_budget)
        self.tile_pixels = int(tile_pixels)
        self.max_tiles = int(max_tiles)
        self.max_active = int(max_active)
        self.use_graph = use_graph
        self.waiting: List[RenderRequest] = []
        self.active: List[RenderRequest] = []
        self.free_slots = list(range(self.max_active))
        self.bytes_in_flight = 0
        self.steps = 0
        self.rays_rendered = 0
        cam_size = ctypes.sizeof(_C._Cam)
        self.cams_dev = torch.zeros(self.max_active * cam_size, dtype=torch.uint8,
                                    device=self.device)
        self.cam_host = torch.zeros(cam_size, dtype=torch.uint8).pin_memory()
        # Launch groups by (storage, basis_dim).
        groups: Dict[tuple, List[int]] = {}
        for i, g in enumerate(self.grids):
            if g.density_data.device != self.device:
                raise RuntimeError("RenderServer: all grids must be on one device")
            key = (_C._storage_of(g.density_data), int(g.basis_dim))
            groups.setdefault(key, []).append(i)
        self.groups = [_Group(k[0], k[1], ids) for k, ids in groups.items()]
        self.group_of = {gid: grp for grp in self.groups for gid in grp.grid_ids}
        tile_size = ctypes.sizeof(_C._Tile)
        # The device grid tables hold raw pointers: the server keeps every
        # tensor they point at, the cell maps included, alive.
        self._maps = []
        for grp in self.groups:
            structs = []
            for i in grp.grid_ids:
                spec = self.grids[i]._to_cpp()
                spec.cell_map = _C.build_cell_map(spec.links, spec.density_data.size(0))
                self._maps.append(spec.cell_map)
                structs.append(_C.grid_struct(spec))
            grp.grids_dev = _C.pack_structs(structs).to(self.device)
            grp.tiles_host = torch.zeros(self.max_tiles * tile_size, dtype=torch.uint8).pin_memory()
            grp.tiles_dev = torch.zeros(self.max_tiles * tile_size, dtype=torch.uint8,
                                        device=self.device)
            grp.count_host = torch.zeros(1, dtype=torch.int32).pin_memory()
            grp.count_dev = torch.zeros(1, dtype=torch.int32, device=self.device)
        # The kernel is loaded and run once (empty tables) before any capture.
        for grp in self.groups:
            self._launch(grp)
        torch.cuda.synchronize(self.device)
        if self.use_graph:
            for grp in self.groups:
                graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(graph):
                    self._launch(grp)
                grp.graph = graph
            torch.cuda.synchronize(self.device)

    # -- device work ---------------------------------------------------------
    def _launch(self, grp: _Group) -> None:
        _C.volume_render_cuvol_tiles(grp.grids_dev, len(grp.grid_ids), grp.storage,
                                     grp.basis_dim, self.cams_dev, self.max_active,
                                     self.opt._to_cpp(), grp.tiles_dev, grp.count_dev,
                                     self.max_tiles, self.tile_pixels)

    # -- scheduling ------------------------------------------------------------
    def submit(self, req: RenderRequest) -> None:
        if not 0 <= req.grid < len(self.grids):
            raise ValueError(f"request names grid {req.grid} of {len(self.grids)}")
        if req.image_bytes > self.hbm_budget:
            raise ValueError("request image exceeds the whole HBM budget")
        self.waiting.append(req)

    def _admit(self, now: float) -> None:
        while self.waiting and self.free_slots:
            req = self.waiting[0]
            if self.bytes_in_flight + req.image_bytes > self.hbm_budget:
                break
            self.waiting.pop(0)
            req.slot = self.free_slots.pop(0)
            req.image = torch.empty((int(req.camera.height), int(req.camera.width), 3),
                                    dtype=torch.float32, device=self.device)
            req.next_pixel = 0
            req.admitted = now
            self.bytes_in_flight += req.image_bytes
            cam = _C.camera_struct(req.camera._to_cpp())
            ctypes.memmove(self.cam_host.data_ptr(), ctypes.addressof(cam), ctypes.sizeof(cam))
            size = ctypes.sizeof(cam)
            self.cams_dev[req.slot * size:(req.slot + 1) * size].copy_(self.cam_host,
                                                                       non_blocking=True)
            # The staging buffer is reused: wait for this copy to leave it.
            torch.cuda.current_stream(self.device).synchronize()
            self.active.append(req)

    def step(self, now: float) -> List[RenderRequest]:
        """Admit what fits, run one multi-request launch per group, wait for
        it; returns the requests that completed (their ``done`` set to the
        clock after the wait, via ``clock`` in :meth:`run`)."""
        self._admit(now)
        if not self.active:
            return []
        tile_size = ctypes.sizeof(_C._Tile)
        for grp in self.groups:
            mine = [r for r in self.active if self.group_of[r.grid] is grp]
            plan = plan_tiles(mine, self.max_tiles, self.tile_pixels)
            structs = [_C.tile_struct(grp.local[r.grid], r.slot, p0, n, r.image)
                       for r, p0, n in plan]
            grp.count_host[0] = len(plan)
            if plan:
                blob = _C.pack_structs(structs)
                grp.tiles_host[:blob.numel()].copy_(blob)
                grp.tiles_dev[:blob.numel()].copy_(grp.tiles_host[:blob.numel()],
                                                   non_blocking=True)
                self.rays_rendered += sum(n for _, _, n in plan)
            grp.count_dev.copy_(grp.count_host, non_blocking=True)
            if plan:
                if grp.graph is not None:
                    grp.graph.replay()
                else:
                    self._launch(grp)
        self.steps += 1
        torch.cuda.current_stream(self.device).synchronize()
        finished = [r for r in self.active if r.next_pixel >= r.pixels]
        for r in finished:
            self.active.remove(r)
            self.free_slots.append(r.slot)
            self.bytes_in_flight -= r.image_bytes
        return finished

    def run(self, requests: Sequence[RenderRequest], realtime: bool = True,
            clock=time.perf_counter) -> dict:
        """Serves ``requests`` (sorted by arrival); with ``realtime`` each is
        submitted once the clock passes its arrival (seconds from the start),
        else all are submitted at once.  Returns latency and throughput
        figures."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = clock()
        i = 0
        done: List[RenderRequest] = []
        while len(done) < len(reqs):
            now = clock() - t0
            while i < len(reqs) and (not realtime or reqs[i].arrival <= now):
                self.submit(reqs[i])
                i += 1
            if not self.active and not self.waiting:
                if i < len(reqs):
                    time.sleep(max(0.0, min(reqs[i].arrival - now, 1e-3)))
                continue
            fin = self.step(now)
            t = clock() - t0
            for r in fin:
                r.done = t
            done.extend(fin)
        return summarize(done, clock() - t0)


def summarize(done: Sequence[RenderRequest], wall: float) -> dict:
    lat = np.array([r.latency for r in done]) * 1e3
    rays = sum(r.pixels for r in done)
    return dict(requests=len(done), rays=int(rays), wall_s=wall, rays_per_s=rays / wall,
                p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()))


def serve_one_per_launch(grids: Sequence[SparseGrid], requests: Sequence[RenderRequest],
                         realtime: bool = True, clock=time.perf_counter) -> dict:
    """The plain alternative: requests in arrival order, each rendered alone by
    one ``volume_render_image`` launch and waited for."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    t0 = clock()
    for r in reqs:
        now = clock() - t0
        if realtime and r.arrival > now:
            time.sleep(r.arrival - now)
        with torch.no_grad():
            r.image = grids[r.grid].volume_render_image(r.camera)
        torch.cuda.synchronize()
        r.done = clock() - t0
    return summarize(reqs, clock() - t0)
