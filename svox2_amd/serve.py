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
* Direct dispatch: while at most ``direct_max`` fresh requests are pending,
  each goes out at once as one image-kernel launch, in arrival order on one
  stream, with no tile table to plan or upload; a longer queue is batched.
* Continuous batching: every ``step`` builds one tile table per launch group
  (grids of one storage kind and basis) from all admitted requests: their
  remaining tiles of ``tile_pixels`` pixels in admission order, until the
  table holds ``max_tiles`` tiles (``policy="round_robin"`` instead hands out
  one tile per request per round). A request admitted during a step joins the
  next step, and many requests share one launch.
* Each step copies the tables to fixed device buffers and replays the group's
  graph. The graph is one ``svox2_volume_render_cuvol_tiles`` launch captured
  once, whose shape depends only on ``max_tiles`` and ``tile_pixels``.
* Steps are pipelined over ``depth`` table slots, each with its own buffers
  and graphs: the host plans and uploads step k+1 while step k runs, and it
  learns of completions from per-step events. By default each slot has a
  stream of its own, so consecutive steps share the device
  (``concurrent=False`` runs them in order on one stream instead).
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
    cam_struct: Optional[object] = None

    @property
    def pixels(self) -> int:
        return int(self.camera.width) * int(self.camera.height)

    @property
    def image_bytes(self) -> int:
        return self.pixels * 3 * 4

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def plan_tiles(active: Sequence[RenderRequest], max_tiles: int, tile_pixels: int,
               policy: str = "fifo") -> List[tuple]:
    """Tile plan for one launch group: (request, pixel0, n) entries, at most
    max_tiles, advancing each request's next_pixel.  ``fifo`` gives the
    requests their remaining tiles in admission order (a request waits
    behind earlier ones, as with one launch per request, but many requests
    share a launch); ``round_robin`` gives one tile per request per round."""
    plan: List[tuple] = []
    live = [r for r in active if r.next_pixel < r.pixels]
    if policy == "fifo":
        for r in live:
            while r.next_pixel < r.pixels and len(plan) < max_tiles:
                n = min(tile_pixels, r.pixels - r.next_pixel)
                plan.append((r, r.next_pixel, n))
                r.next_pixel += n
            if len(plan) >= max_tiles:
                break
        return plan
    if policy != "round_robin":
        raise ValueError(f"unknown tile policy {policy!r}")
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


def plan_runs(active: Sequence[RenderRequest], max_tiles: int, tile_pixels: int,
              policy: str = "fifo") -> List[tuple]:
    """The tile plan of ``plan_tiles`` as runs of consecutive tiles of one
    request: (request, pixel0 array, n_pixels array) entries whose
    concatenation is plan_tiles' list, with the same next_pixel updates.
    The FIFO plan is built with array arithmetic per request rather than a
    Python step per tile (it is on the serving loop's critical path)."""
    if policy != "fifo":
        runs: List[tuple] = []
        for r, p0, n in plan_tiles(active, max_tiles, tile_pixels, policy):
            runs.append((r, np.array([p0], np.int64), np.array([n], np.int64)))
        return runs
    runs = []
    budget = max_tiles
    for r in active:
        if budget <= 0:
            break
        rem = r.pixels - r.next_pixel
        if rem <= 0:
            continue
        k = min(budget, -(-rem // tile_pixels))
        p0 = r.next_pixel + tile_pixels * np.arange(k, dtype=np.int64)
        n = np.minimum(tile_pixels, r.pixels - p0)
        runs.append((r, p0, n))
        r.next_pixel = int(p0[-1] + n[-1])
        budget -= k
    return runs


# svox2_tile_t as a numpy record (int32 grid, camera, pixel0, n_pixels; float* rgb_out).
_TILE_DTYPE = np.dtype([("grid", "<i4"), ("camera", "<i4"), ("pixel0", "<i4"),
                        ("n_pixels", "<i4"), ("rgb_out", "<u8")])


class _Group:
    """Grids of one storage dtype and basis_dim, rendered by one launch."""

    def __init__(self, storage: int, basis_dim: int, grid_ids: List[int]):
        self.storage = storage
        self.basis_dim = basis_dim
        self.grid_ids = grid_ids  # server grid index -> position in this group's table
        self.local: Dict[int, int] = {g: i for i, g in enumerate(grid_ids)}
        self.grids_dev: Optional[torch.Tensor] = None


class _Slot:
    """One in-flight step's tables: a device buffer [count | tiles | cameras]
    per launch group, its pinned staging copy, captured graphs and stream."""

    def __init__(self, device, stream=None):
        # default priority; steps share one stream unless the server runs
        # them concurrently
        self.stream = stream if stream is not None else torch.cuda.Stream(device=device)
        self.cam_version: List[int] = []  # per launch group: camera table version copied
        self.host: List[torch.Tensor] = []
        self.host_np: List[np.ndarray] = []
        self.dev: List[torch.Tensor] = []
        self.graphs: list = []  # per launch group: {tile-count class: graph}
        self.event = None
        self.busy = False
        self.finishing: List[RenderRequest] = []


class RenderServer:
    """Continuous-batching renderer over resident grids (see module doc).

    Steps are pipelined: ``depth`` table slots, each with its own captured
    graphs and stream, so the host plans and uploads step k+1 while step k
    runs; a request completes when the event of the step that ran its last
    tile has fired.

    :param grids: resident SparseGrids (on the GPU; their cell maps are built
        once here).
    :param hbm_budget: bytes of output images in flight at most.
    :param tile_pixels: pixels per tile.
    :param max_tiles: tiles per launch and group.
    :param max_active: camera slots (requests in flight) at most.
    :param use_graph: replay captured graphs (else launch directly).
    :param concurrent: run the table slots' steps on streams of their own
        (the default), so that consecutive steps share the device; with False
        they run in order on one stream (first come, first served on the
        device, as one launch per request is).  Under overload the concurrent
        steps give the lower latencies.
    :param idle_direct: launch short queues directly (below); False batches
        every request into tile steps.
    :param direct_max: while at most this many requests are pending and none
        has started, each is rendered by one direct image-kernel launch, in
        arrival order on one stream, instead of a tile step: no batching delay
        and no tile table when little queues; longer queues are batched.
    """

    def __init__(self, grids: Sequence[SparseGrid], opt: Optional[RenderOptions] = None,
                 hbm_budget: int = 1 << 30, tile_pixels: int = 4096, max_tiles: int = 128,
                 max_active: int = 64, use_graph: bool = True, policy: str = "fifo",
                 depth: int = 3, concurrent: bool = True, idle_direct: bool = True,
                 direct_max: int = 4):
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
        self.policy = policy
        if depth < 1:
            raise ValueError("depth must be at least 1")
        self.depth = int(depth)
        self.waiting: List[RenderRequest] = []
        self.active: List[RenderRequest] = []
        self.free_cams = list(range(self.max_active))
        self.bytes_in_flight = 0
        self.steps = 0
        self.direct_launches = 0  # steps that were one direct image launch
        self.rays_rendered = 0
        self.cam_size = ctypes.sizeof(_C._Cam)
        self.tile_size = ctypes.sizeof(_C._Tile)
        self.cams = bytearray(self.max_active * self.cam_size)  # host master copy
        self.cam_version = 0
        # Table layout: [count (16 B) | tiles | cameras].
        self.tiles_off = 16
        self.cams_off = self.tiles_off + self.max_tiles * self.tile_size
        self.table_bytes = self.cams_off + self.max_active * self.cam_size
        groups: Dict[tuple, List[int]] = {}
        for i, g in enumerate(self.grids):
            if g.density_data.device != self.device:
                raise RuntimeError("RenderServer: all grids must be on one device")
            key = (_C._storage_of(g.density_data), int(g.basis_dim))
            groups.setdefault(key, []).append(i)
        self.groups = [_Group(k[0], k[1], ids) for k, ids in groups.items()]
        self.group_of = {gid: grp for grp in self.groups for gid in grp.grid_ids}
        # The device grid tables hold raw pointers: the server keeps every
        # tensor they point at, the cell maps included, alive.
        self._maps = []
        self._grid_structs: Dict[int, object] = {}  # host descriptors, for direct launches
        self._opt_struct = _C._opt_struct(self.opt._to_cpp())
        for grp in self.groups:
            structs = []
            for i in grp.grid_ids:
                spec = self.grids[i]._to_cpp()
                spec.cell_map = _C.build_cell_map(spec.links, spec.density_data.size(0))
                self._maps.append(spec.cell_map)
                structs.append(_C.grid_struct(spec))
                self._grid_structs[i] = structs[-1]
            grp.grids_dev = _C.pack_structs(structs).to(self.device)
        self.concurrent = bool(concurrent)
        self.direct_max = int(direct_max) if idle_direct else 0
        self.direct_stream = torch.cuda.Stream(device=self.device)  # default priority
        shared = None if self.concurrent else torch.cuda.Stream(device=self.device)
        self.slots = [_Slot(self.device, shared) for _ in range(self.depth)]
        for slot in self.slots:
            for _ in self.groups:
                slot.cam_version.append(-1)
                slot.host.append(torch.zeros(self.table_bytes, dtype=torch.uint8).pin_memory())
                slot.host_np.append(slot.host[-1].numpy())
                slot.dev.append(torch.zeros(self.table_bytes, dtype=torch.uint8,
                                            device=self.device))
        # Launch sizes: a step with few tiles replays a graph with a small
        # grid, since every workgroup of the grid is dispatched even when its
        # tile is not live.
        self.size_classes = sorted({c for c in (8, 32, self.max_tiles) if c <= self.max_tiles})
        # The kernel is loaded and run once (empty tables) before any capture.
        for slot in self.slots:
            with torch.cuda.stream(slot.stream):
                for gi, grp in enumerate(self.groups):
                    self._launch(grp, slot.dev[gi], self.max_tiles)
        torch.cuda.synchronize(self.device)
        if self.use_graph:
            for slot in self.slots:
                for gi, grp in enumerate(self.groups):
                    per_size = {}
                    for size in self.size_classes:
                        graph = torch.cuda.CUDAGraph()
                        with torch.cuda.graph(graph):
                            self._launch(grp, slot.dev[gi], size)
                        per_size[size] = graph
                    slot.graphs.append(per_size)
            torch.cuda.synchronize(self.device)
        self._next_slot = 0
        self._inflight: List[_Slot] = []
        self._clock = time.perf_counter

    # -- device work ---------------------------------------------------------
    def _launch(self, grp: _Group, table: torch.Tensor, size: int) -> None:
        _C.volume_render_cuvol_tiles(grp.grids_dev, len(grp.grid_ids), grp.storage,
                                     grp.basis_dim, table[self.cams_off:], self.max_active,
                                     self.opt._to_cpp(), table[self.tiles_off:self.cams_off],
                                     table[:4].view(torch.int32), size, self.tile_pixels)

    # -- scheduling ------------------------------------------------------------
    def submit(self, req: RenderRequest) -> None:
        if not 0 <= req.grid < len(self.grids):
            raise ValueError(f"request names grid {req.grid} of {len(self.grids)}")
        if req.image_bytes > self.hbm_budget:
            raise ValueError("request image exceeds the whole HBM budget")
        self.waiting.append(req)

    def _admit(self, now: float) -> None:
        while self.waiting and self.free_cams:
            req = self.waiting[0]
            if self.bytes_in_flight + req.image_bytes > self.hbm_budget:
                break
            self.waiting.pop(0)
            req.slot = self.free_cams.pop(0)
            req.image = torch.empty((int(req.camera.height), int(req.camera.width), 3),
                                    dtype=torch.float32, device=self.device)
            req.next_pixel = 0
            req.admitted = now
            self.bytes_in_flight += req.image_bytes
            cam = _C.camera_struct(req.camera._to_cpp())
            req.cam_struct = cam
            o = req.slot * self.cam_size
            self.cams[o:o + self.cam_size] = bytes(memoryview(cam))
            self.cam_version += 1
            self.active.append(req)

    def _retire(self, slot: _Slot) -> List[RenderRequest]:
        slot.event.synchronize()
        t = self._clock()
        slot.busy = False
        done = slot.finishing
        slot.finishing = []
        for r in done:
            r.done = t
            self.active.remove(r)
            self.free_cams.append(r.slot)
            self.bytes_in_flight -= r.image_bytes
        return done

    def poll(self) -> List[RenderRequest]:
        """Requests whose last step has finished on the device (no waiting)."""
        done: List[RenderRequest] = []
        while self._inflight and self._inflight[0].event.query():
            done += self._retire(self._inflight.pop(0))
        return done

    def step(self, now: float) -> List[RenderRequest]:
        """Admit what fits and launch one multi-request step (waiting first
        for the oldest in-flight step if both table slots are busy).
        Returns the requests that completed meanwhile."""
        done = self.poll()
        self._admit(now)
        pending = [r for r in self.active if r.next_pixel < r.pixels]
        if not pending:
            return done
        slot = self.slots[self._next_slot]
        if slot.busy:
            done += self._retire(self._inflight.pop(0))
        self._next_slot = (self._next_slot + 1) % self.depth
        finishing: List[RenderRequest] = []
        # Steps run on the slots' stream (one stream shared by all slots
        # unless the server is concurrent, then a stream per slot, so a step
        # can start while the previous one drains).  Steps share nothing a
        # step writes: each slot has its own tables, tiles of one image are
        # disjoint, and camera slots and images are reused only after the
        # event of the step that last read them.  The slot stream follows the
        # caller's stream, where the grids were written.
        # A short queue of fresh requests goes out as direct image launches,
        # in arrival order on one stream (first come, first served, as one
        # launch per request); a longer queue is batched into a tile step.
        direct = len(pending) <= self.direct_max and all(r.next_pixel == 0 for r in pending)
        stream = self.direct_stream if direct else slot.stream
        stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(stream):
            if direct:
                for r in pending:
                    self._launch_direct(r, stream, finishing)
            else:
                self._fill_and_launch(slot, pending, finishing)
        slot.finishing = finishing
        if slot.event is None:
            slot.event = torch.cuda.Event()
        slot.event.record(stream)  # the slot's previous step has been retired
        slot.busy = True
        self._inflight.append(slot)
        self.steps += 1
        return done

    def _launch_direct(self, req: RenderRequest, stream, finishing: List[RenderRequest]) -> None:
        """A request of a short queue: its whole image in one launch of the
        image kernel (the same per-pixel march as the tiles kernel), with no
        tile table to plan or upload."""
        _C._call("svox2_volume_render_cuvol_image", ctypes.byref(self._grid_structs[req.grid]),
                 ctypes.byref(req.cam_struct), ctypes.byref(self._opt_struct),
                 ctypes.c_void_p(req.image.data_ptr()), None,
                 ctypes.c_void_p(stream.cuda_stream))
        req.next_pixel = req.pixels
        self.rays_rendered += req.pixels
        self.direct_launches += 1
        finishing.append(req)

    def _fill_and_launch(self, slot: _Slot, pending: List[RenderRequest],
                         finishing: List[RenderRequest]) -> None:
        for gi, grp in enumerate(self.groups):
            mine = [r for r in pending if self.group_of[r.grid] is grp]
            runs = plan_runs(mine, self.max_tiles, self.tile_pixels, self.policy)
            if not runs:
                continue  # nothing launched: this group's table is not read
            host = slot.host_np[gi]
            n_tiles = sum(len(p0) for _, p0, _ in runs)
            host[:4].view(np.int32)[0] = n_tiles
            tiles = host[self.tiles_off:self.tiles_off + n_tiles * self.tile_size].view(_TILE_DTYPE)
            at = 0
            for r, p0, n in runs:
                k = len(p0)
                t = tiles[at:at + k]
                t["grid"] = grp.local[r.grid]
                t["camera"] = r.slot
                t["pixel0"] = p0
                t["n_pixels"] = n
                t["rgb_out"] = r.image.data_ptr()
                at += k
                self.rays_rendered += int(n.sum())
                if r.next_pixel >= r.pixels and r not in finishing:
                    finishing.append(r)
            nbytes = self.tiles_off + n_tiles * self.tile_size
            if slot.cam_version[gi] != self.cam_version:
                host[self.cams_off:] = np.frombuffer(self.cams, dtype=np.uint8)
                nbytes = self.table_bytes
                slot.cam_version[gi] = self.cam_version
            slot.dev[gi][:nbytes].copy_(slot.host[gi][:nbytes], non_blocking=True)
            size = next(c for c in self.size_classes if c >= n_tiles)
            if self.use_graph:
                slot.graphs[gi][size].replay()
            else:
                self._launch(grp, slot.dev[gi], size)

    def drain(self) -> List[RenderRequest]:
        done: List[RenderRequest] = []
        while self._inflight:
            done += self._retire(self._inflight.pop(0))
        return done

    def run(self, requests: Sequence[RenderRequest], realtime: bool = True,
            clock=time.perf_counter) -> dict:
        """Serves ``requests`` (sorted by arrival); with ``realtime`` each is
        submitted once the clock passes its arrival (seconds from the start),
        else all are submitted at once.  The loop never blocks on the device
        while requests can still be admitted: it launches whenever a table
        slot is free and polls step events otherwise.  Returns latency and
        throughput figures."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = clock()
        self._clock = lambda: clock() - t0
        i = 0
        n_done = 0
        done_all: List[RenderRequest] = []
        while n_done < len(reqs):
            now = clock() - t0
            while i < len(reqs) and (not realtime or reqs[i].arrival <= now):
                self.submit(reqs[i])
                i += 1
            fin = self.poll()
            has_work = bool(self.waiting) or any(r.next_pixel < r.pixels for r in self.active)
            if has_work and not self.slots[self._next_slot].busy:
                fin += self.step(now)
            elif not self._inflight and not has_work and i < len(reqs):
                # Idle: sleep until the next arrival.
                time.sleep(max(0.0, reqs[i].arrival - (clock() - t0)))
            elif not fin:
                time.sleep(0)  # yield while steps run
            n_done += len(fin)
            done_all += fin
        return summarize(done_all, clock() - t0)


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
