"""Benchmark: fused forward+backward rays/s of the sparse-voxel renderer on the
512^3 (about 3% occupied, 4.26M rows) SH-2 grid, 65,536-ray batches, float32
storage, on one MI355X.

A step is one SparseGrid.volume_render_fused call (render + MSE gradient into
density_data.grad / sh_data.grad + touched-row mask) over one batch of random
camera rays already resident in HBM.  Prints one JSON line with the
throughput, a roofline entry for the fused kernel (algorithmic bytes per
launch / average launch time from HIP events on the launch stream, against the
8 TB/s HBM peak), side measurements (BF16 storage, 5,000-ray batches, the
million-ray autograd step, the serving loop, the 256^3 training step with
the TV, RMSprop, accel_dist_prop and sample kernels against their own
algorithmic bytes, an 800x800 forward image) and CPU baselines on the host's
cores: the C port of the reference's fused forward+backward
(oracle/march_ref.c) on the same 512^3 grid, and the restated reference CPU
path (_volume_render_gradcheck_lerp + autograd) on the 64^3 dense workload.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
  python bench.py --only {small,side,image,serve,stress,next}   (one side measurement)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402
import torch  # noqa: E402

METRIC = "fused fwd+bwd rays/s @512^3 SH2 (GB/s, % HBM peak); serve p50/p99 ms"
HBM_PEAK_GBS = 8000.0
BATCH = 65536
POOL = 8


PROFILE_SUMMARY = os.path.join(ROOT, "profiles", "fused_512_f32_65536.md")


def measured_traffic():
    """HBM bytes per fused call of the headline workload from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE passes of
    `bench.py --pmc`, each scaled by what the counter reads for a 1 GiB copy),
    or None if the summary is absent."""
    try:
        text = open(PROFILE_SUMMARY).read()
    except OSError:
        return None
    import re
    m = re.search(r"^traffic_bytes_per_call: (\d+)$", text, re.M)
    return int(m.group(1)) if m else None


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(stats: np.ndarray, s_sigma: int, s_sh: int) -> dict:
    """Logical bytes of the reference algorithm for one fused launch: per ray
    rays/gt/colour traffic, 4 B link read per march iteration (I), 7 further
    links and 8 densities per evaluated sample (E), 8 SH rows per contributing
    sample (H) in the forward and again in the backward, and in the backward
    216 + 8 float atomics (8 B read-modify-write each) + 8 mask bytes per
    contributing sample."""
    Q = stats.shape[0]
    I = int(stats[:, 0].astype(np.int64).sum())
    E = int(stats[:, 1].astype(np.int64).sum())
    H = int(stats[:, 2].astype(np.int64).sum())
    fwd = 36 * Q + 4 * I + (28 + 8 * s_sigma) * E + 216 * s_sh * H
    bwd = 48 * Q + 4 * I + (28 + 8 * s_sigma) * E + (216 * s_sh + 1728 + 64 + 8) * H
    return dict(Q=Q, I=I, E=E, H=H, fwd=fwd, bwd=bwd, total=fwd + bwd)


class Workload:
    def __init__(self, dev, dtype=torch.float32, seed=0):
        from svox2_amd import scenes
        t = time.time()
        self.scene, self.cams, self.gen = scenes.workload_blobs(512, scenes.SPARSE512, seed=seed,
                                                                device=dev)
        self.dev = dev
        self.grids = {}
        log(f"512^3 grid: {self.scene.density.shape[0]} rows "
            f"({self.scene.occupancy:.2%}) in {time.time() - t:.1f}s")

    def grid(self, dtype):
        from svox2_amd import scenes
        if dtype not in self.grids:
            g = scenes.to_sparse_grid(self.scene, device=self.dev, dtype=dtype)
            g.opt.sigma_thresh = 1e-8
            g.opt.stop_thresh = 1e-7
            self.grids[dtype] = g
        return self.grids[dtype]

    def batches(self, n_rays, count):
        out = []
        for _ in range(count):
            o, d = self.cams.random_batch(n_rays, self.gen)
            gt = torch.rand(n_rays, 3, generator=self.gen)
            out.append((o.to(self.dev), d.to(self.dev), gt.to(self.dev)))
        return out


def run_steps(grid, batches, steps, warmup):
    """Warm-up, then `steps` timed facade steps; returns (elapsed s, per-step
    GPU ms from events)."""
    import svox2_amd
    rays = [(svox2_amd.Rays(o, d), gt) for o, d, gt in batches]
    for i in range(warmup):
        r, gt = rays[i % len(rays)]
        grid.volume_render_fused(r, gt)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        r, gt = rays[i % len(rays)]
        ev[i][0].record()
        grid.volume_render_fused(r, gt)
        ev[i][1].record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    return elapsed, [a.elapsed_time(b) for a, b in ev]


def kernel_times(grid, batches, reps):
    """Average duration (ms) of the fused kernel alone, HIP events on the
    stream it is launched on (torch's current stream)."""
    from svox2_amd import _C
    import svox2_amd
    spec = grid._to_cpp()
    opt = grid.opt._to_cpp()
    gd, gs, _, _ = grid._get_data_grads()
    holder = _C.GridOutputGrads()
    holder.grad_density_out, holder.grad_sh_out = gd, gs
    holder.mask_out = torch.zeros(grid.density_data.size(0), dtype=torch.bool, device=gd.device)
    outs = [torch.empty_like(o) for o, _, _ in batches]
    rays = [svox2_amd.Rays(o, d)._to_cpp() for o, d, _ in batches]
    stream = torch.cuda.current_stream()
    times = []
    for i in range(reps):
        k = i % len(batches)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        _C.volume_render_cuvol_fused(spec, rays[k], opt, batches[k][2], 0.0, 0.0, outs[k], holder)
        b.record(stream)
        times.append((a, b, k))
    torch.cuda.synchronize()
    return [(a.elapsed_time(b), k) for a, b, k in times]


def batch_stats(grid, batches):
    from svox2_amd import _C
    import svox2_amd
    return [_C.volume_render_cuvol_stats(grid._to_cpp(), svox2_amd.Rays(o, d)._to_cpp(),
                                         grid.opt._to_cpp()).cpu().numpy()
            for o, d, _ in batches]


def measure(grid, batches, steps, warmup, s_sigma, s_sh):
    elapsed, step_ms = run_steps(grid, batches, steps, warmup)
    stats = batch_stats(grid, batches)
    kt = kernel_times(grid, batches, max(steps, len(batches)))
    byts = [algorithmic_bytes(s, s_sigma, s_sh) for s in stats]
    k_ms = float(np.mean([t for t, _ in kt]))
    k_bytes = float(np.mean([byts[k]["total"] for _, k in kt]))
    Q = batches[0][0].shape[0]
    return dict(
        rays_per_s=Q * steps / elapsed,
        ms_per_step=1e3 * elapsed / steps,
        step_gpu_ms=float(np.mean(step_ms)),
        kernel_ms=k_ms,
        alg_bytes_per_launch=k_bytes,
        alg_gbps=k_bytes / (k_ms * 1e-3) / 1e9,
        counters={k: int(np.mean([b[k] for b in byts])) for k in ("I", "E", "H")},
    )


def _event_ms(fn, reps: int) -> float:
    """Average ms of fn() over reps launches, HIP events on the current stream
    (the stream the kernels are launched on)."""
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def _roof(ms: float, nbytes: float) -> dict:
    gbps = nbytes / (ms * 1e-3) / 1e9
    return dict(ms=ms, alg_bytes=int(nbytes), alg_gbps=gbps, frac=gbps / HBM_PEAK_GBS)


def _tv_pairs(links: torch.Tensor, cells: torch.Tensor, rows: int, ignore_edge: bool):
    """(cells inside the grid, (cell, neighbour) pairs whose link addresses a
    row) of one tv_grad_sparse launch: the cell and its +z, +y, +x
    neighbours, an outside neighbour reading link 0, and with ignore_edge the
    cells whose own link is 0 skipped (loss_kernel.cu:178-246).  Each pair is
    one data read and, for a nonzero difference, one gradient atomic and one
    mask store per channel."""
    s0, s1, s2 = links.shape
    n = s0 * s1 * s2
    flat = links.reshape(-1).long()
    c = cells.long()
    z, y, x = c % s2, (c // s2) % s1, c // (s1 * s2)
    inside = (c >= 0) & (x < s0)
    at = lambda i: flat[i.clamp(0, n - 1)]
    zero = torch.zeros_like(c)
    l000 = at(c)
    l001 = torch.where(z + 1 < s2, at(c + 1), zero)
    l010 = torch.where(y + 1 < s1, at(c + s2), zero)
    l100 = torch.where(x + 1 < s0, at(c + s1 * s2), zero)
    ok = lambda l: ((l >= 0) & (l < rows)).long()
    live = inside & ~((l000 == 0) & ignore_edge)
    pairs = (ok(l000) + ok(l001) + ok(l010) + ok(l100)) * live.long()
    return int(inside.sum()), int(pairs.sum())


def _valid_corners(grid, pts: torch.Tensor) -> int:
    """Corners, over all points, whose link addresses a row: the trilinear
    cell of each point (grid coordinates clamped to the grid, lower corner at
    most size - 2 per axis) and its 8 corners."""
    n = torch.tensor(grid.links.shape, device=pts.device)
    g = grid.world2grid(pts).clamp(min=0)
    g = torch.minimum(g, (n - 1).to(g.dtype))
    lo = torch.minimum(g.floor().long(), n - 2)
    flat = grid.links.reshape(-1).long()
    rows = grid.density_data.size(0)
    total = 0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                i = ((lo[:, 0] + dx) * n[1] + lo[:, 1] + dy) * n[2] + lo[:, 2] + dz
                lk = flat[i]
                total += int(((lk >= 0) & (lk < rows)).sum())
    return total


def side_kernels(wl, steps: int, warmup: int, reps: int = 20) -> dict:
    """The kernels around the march, on the workloads they serve.

    train_256: the full training step on the 256^3 grid (about 8% occupied,
    SH-2, float32): a 5000-ray fused forward+backward, TV of the density
    (lambda 1e-5) and of the SH (lambda 1e-3) on a contiguous 1% of the cells,
    then RMSprop on both (beta 0.95, lr 30 and 1e-2) over the touched rows,
    through SparseGrid as a training script calls it; ms per step is host
    wall time with the device synchronised, as a training loop sees it.

    Kernel-only times (HIP events, launches repeated on fixed inputs) with
    algorithmic bytes per launch against the 8 TB/s HBM peak:
    - tv_grad_sparse: 16 B of links per cell, and per (cell, neighbour) pair
      whose link addresses a row, per channel, a data read, an 8 B gradient
      atomic and a 1 B mask store;
    - rmsprop_step: 24 B per updated element (data, state and gradient, each
      read and written) plus the indexer (8 B per row index);
    - accel_dist_prop on the 512^3 links: two reads of the links, a write of
      every empty link, and the pyramid cleared, then read and rewritten by
      the node pass;
    - sample (trilinear density + SH at 1M random points of the 512^3 grid):
      per point 12 B in, 8 links and 112 B out, plus a 112 B row per corner
      that has one; its backward reads the point, 8 links and 112 B of
      output gradient and adds into the gradient row of each corner that has
      one (28 float atomics of 8 B)."""
    import ctypes

    import svox2_amd
    from svox2_amd import _C, scenes
    out = {}
    sc, cams, _ = scenes.workload_blobs(256, scenes.SPARSE256, seed=1, device=wl.dev)
    grid = scenes.to_sparse_grid(sc, device=wl.dev)
    grid.opt.sigma_thresh, grid.opt.stop_thresh = 1e-8, 1e-7
    gen = torch.Generator().manual_seed(5)
    batches = []
    for _ in range(POOL):
        o, d = cams.random_batch(5000, gen)
        batches.append((svox2_amd.Rays(o.to(wl.dev), d.to(wl.dev)),
                        torch.rand(5000, 3, generator=gen).to(wl.dev)))
    np.random.seed(0)  # the TV run's random start, drawn with np.random as the reference

    def step(i):
        rays, gt = batches[i % POOL]
        grid.volume_render_fused(rays, gt)
        grid.inplace_tv_grad(grid.density_data.grad, scaling=1e-5, sparse_frac=0.01)
        grid.inplace_tv_color_grad(grid.sh_data.grad, scaling=1e-3, sparse_frac=0.01)
        grid.optim_density_step(30.0, beta=0.95)
        grid.optim_sh_step(1e-2, beta=0.95)

    for i in range(warmup):
        step(i)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for i in range(steps):
        step(i)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t) * 1e3 / steps
    out["train_256"] = dict(ms_per_step=ms, rays_per_s=5000 / (ms * 1e-3), rows=int(grid.density_data.size(0)))

    # Kernel-only: the inputs of one step, reused.
    rays, gt = batches[0]
    grid.volume_render_fused(rays, gt)
    rows = grid.density_data.size(0)
    cells = grid._get_rand_cells(0.01, contiguous=True)
    mask = grid._get_sparse_grad_indexer()
    gd, gs = grid.density_data.grad, grid.sh_data.grad
    for name, data, grad, lo, hi, edge, lam in (
            ("tv_density", grid.density_data, gd, 0, 1, False, 1e-5),
            ("tv_sh", grid.sh_data, gs, 0, grid.sh_data.size(1), True, 1e-3)):
        inside, pairs = _tv_pairs(grid.links, cells, rows, edge)
        s_data = data.element_size()
        nbytes = 16 * inside + (hi - lo) * pairs * (s_data + 8 + 1)
        ms = _event_ms(lambda: _C.tv_grad_sparse(grid.links, data, cells, mask, lo, hi, lam,
                                                 False, 2.0, edge, False, -1.0, -1.0, grad),
                       reps)
        out[name] = dict(_roof(ms, nbytes), cells=inside, pairs=pairs, channels=hi - lo)
    grid._ensure_rms()
    idx = torch.nonzero(mask.flatten(), as_tuple=False).flatten()
    for name, data, rms, grad, lr in (
            ("rmsprop_density", grid.density_data, grid.density_rms, gd, 30.0),
            ("rmsprop_sh", grid.sh_data, grid.sh_rms, gs, 1e-2)):
        dim = data.size(1)
        nbytes = 24 * idx.numel() * dim + 8 * idx.numel()
        ms = _event_ms(lambda: _C.rmsprop_step(data.data, rms, grad, idx, 0.95, lr, 1e-8,
                                               -1e9, lr), reps)
        out[name] = dict(_roof(ms, nbytes), rows_updated=int(idx.numel()), channels=dim)

    g512 = wl.grid(torch.float32)
    links = g512.links.clone()
    n = links.numel()
    empty = int((links < 0).sum())
    size3 = (ctypes.c_int32 * 3)(*links.shape)
    ws_bytes = _C.load_library().svox2_accel_dist_prop_workspace_bytes(size3)
    nbytes = 8 * n + 4 * empty + 3 * ws_bytes
    ms = _event_ms(lambda: _C.accel_dist_prop(links), 5)
    out["accel_dist_prop_512"] = dict(_roof(ms, nbytes), cells=n, empty=empty)
    assert torch.equal(links, g512.links)  # idempotent on accelerated links

    pts = (torch.rand(1 << 20, 3, generator=torch.Generator().manual_seed(3)) * 2 - 1).to(wl.dev)
    npts = pts.size(0)
    corners = _valid_corners(g512, pts)
    with torch.no_grad():
        ms = _event_ms(lambda: g512.sample(pts), reps)
    out["sample_512_1m"] = dict(_roof(ms, npts * (12 + 32 + 112) + corners * 112), points=npts,
                                corners_with_rows=corners)
    gd, gs, _, _ = g512._get_data_grads()
    gen_g = torch.Generator().manual_seed(4)
    god = torch.randn(npts, 1, generator=gen_g).to(wl.dev)
    gos = torch.randn(npts, g512.sh_data.size(1), generator=gen_g).to(wl.dev)
    spec = g512._to_cpp()
    ms = _event_ms(lambda: _C.sample_grid_backward(spec, pts, god, gos, gd, gs, True), reps)
    out["sample_backward_512_1m"] = dict(_roof(ms, npts * (12 + 32 + 112) + corners * 28 * 8),
                                         points=npts, corners_with_rows=corners)
    return out


def image_forward(grid, cams, reps=5):
    cam = cams.camera(0)
    with torch.no_grad():
        grid.volume_render_image(cam)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            grid.volume_render_image(cam)
        b.record()
        torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    return dict(ms=ms, rays_per_s=cam.width * cam.height / (ms * 1e-3))


def stress_1m(wl: Workload, grid, steps: int = 3):
    """Million-ray differentiation step on the 512^3 grid: 1,048,576 random
    camera rays with per-ray (t_near ~ U[0, 0.5], t_far ~ U[2, 4]) clips, the
    autograd path (SparseGrid.volume_render, MSE, backward re-marching every
    ray), timed with events; peak HBM is the allocator's high-water mark
    during the steps above what was allocated before them."""
    import svox2_amd
    from svox2_amd import scenes
    n = 1 << 20
    gen = torch.Generator().manual_seed(77)
    o, d = wl.cams.random_batch(n, gen)
    clip = scenes.config_clip(n, gen)
    gt = torch.rand(n, 3, generator=gen)
    rays = svox2_amd.Rays(o.to(wl.dev), d.to(wl.dev), clip.to(wl.dev))
    gt = gt.to(wl.dev)

    def step():
        grid.density_data.grad = None
        grid.sh_data.grad = None
        rgb = grid.volume_render(rays)
        loss = torch.nn.functional.mse_loss(rgb, gt)
        loss.backward()

    step()
    torch.cuda.synchronize()
    grid.density_data.grad = None
    grid.sh_data.grad = None
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        step()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    peak = torch.cuda.max_memory_allocated() - base
    grid.density_data.grad = None
    grid.sh_data.grad = None
    return dict(rays=n, ms_per_step=ms, rays_per_s=n / (ms * 1e-3), peak_hbm_bytes=int(peak),
                peak_hbm_gb=peak / 1e9,
                note="autograd volume_render + MSE backward, per-ray clips; peak above the "
                     "resident grid")


SERVE_SIDES = (128, 256, 512, 800, 1024)


def serve_requests(cam_sets, n: int, qps: float, seed: int):
    """Poisson arrivals at `qps`; each request a uniform random grid of four,
    a uniform random camera of that grid's hemisphere set and a uniform side
    from SERVE_SIDES (square images, the 800-pixel focal length scaled)."""
    from svox2_amd.serve import RenderRequest
    g = np.random.default_rng(seed)
    t = np.cumsum(g.exponential(1.0 / qps, size=n))
    reqs = []
    for i in range(n):
        grid = int(g.integers(4))
        cams = cam_sets[grid]
        side = int(SERVE_SIDES[g.integers(len(SERVE_SIDES))])
        cam = cams.camera(int(g.integers(cams.c2w.shape[0])), width=side, height=side)
        reqs.append(RenderRequest(grid=grid, camera=cam, arrival=float(t[i])))
    return reqs


def serve_bench(wl: Workload, qps_list=(200.0, 600.0, 1000.0), n: int = 600):
    """The serving loop on four resident grids (256^3 ~8% and 512^3 ~3%
    occupied, one of each in bf16): at each rate, the same Poisson request
    stream served by the continuous-batching server and by one
    volume_render_image launch per request."""
    from svox2_amd import scenes
    from svox2_amd.serve import RenderServer, serve_one_per_launch
    sc256, cams256, _ = scenes.workload_blobs(256, scenes.SPARSE256, seed=1, device=wl.dev)
    g256 = {}
    for dt in (torch.float32, torch.bfloat16):
        g = scenes.to_sparse_grid(sc256, device=wl.dev, dtype=dt)
        g.opt.sigma_thresh, g.opt.stop_thresh = 1e-8, 1e-7
        g256[dt] = g
    grids = [g256[torch.float32], g256[torch.bfloat16], wl.grid(torch.float32),
             wl.grid(torch.bfloat16)]
    cam_sets = [cams256, cams256, wl.cams, wl.cams]
    server = RenderServer(grids, hbm_budget=1 << 30, tile_pixels=4096, max_tiles=128,
                          max_active=64, policy="fifo", depth=3)
    server.run(serve_requests(cam_sets, 16, 100.0, seed=100), realtime=False)  # warm-up
    serve_one_per_launch(grids, serve_requests(cam_sets, 8, 100.0, seed=101), realtime=False)
    out = {"requests": n, "sides": list(SERVE_SIDES), "tile_pixels": 4096, "max_tiles": 128,
           "policy": "fifo", "depth": 3,
           "grids": "256^3 f32, 256^3 bf16, 512^3 f32, 512^3 bf16", "rates": []}
    for qps in qps_list:
        steps0 = server.steps
        batched = server.run(serve_requests(cam_sets, n, qps, seed=7))
        batched["steps"] = server.steps - steps0
        plain = serve_one_per_launch(grids, serve_requests(cam_sets, n, qps, seed=7))
        out["rates"].append({"qps": qps, "batched": batched, "one_per_launch": plain})
    return out


def host_threads() -> int:
    """Threads the host leg may use: OMP_NUM_THREADS when set (the GPU box
    sets it to its CPU share), else the CPUs this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl: Workload, budget_s: float = 12.0, runs: int = 5):
    """CPU baselines on this host's cores, median of `runs` timed runs after
    one warm-up:

    - the headline figure: the C port of the reference's fused forward +
      MSE backward (oracle/march_ref.c ref_fused, the CUDA kernel's algorithm
      ray by ray, rays split over OpenMP threads, gradient updates atomic) on
      65536-ray batches of the benchmark's 512^3 grid;
    - `reference_path`: the restatement of the reference's own CPU path,
      SparseGrid._volume_render_gradcheck_lerp plus autograd for the MSE
      gradient (svox2.py:658-884), float32, torch on all threads, at full size
      on the 64^3 dense grid with 4096 rays.  At 512^3 that path is out of
      reach on a CPU: every march step gathers rows with dense indexing whose
      backward zero-fills a rows x 27 gradient table per corner (8 x 460 MB
      per step at 4.26M rows, some 3000 steps per batch), and its autograd
      state for a 65536-ray batch would be about 200 GB."""
    import oracle.oracle as O
    O.build(force=False)
    threads = host_threads()
    grid = wl.grid(torch.float32)
    N = np.array(grid.links.shape, np.float64)
    hg = O.Grid(grid.links.cpu().numpy(), grid.density_data.data.cpu().double().numpy().reshape(-1),
                grid.sh_data.data.cpu().double().numpy(), 9, offset=0.5 * N - 0.5,
                scaling=0.5 * N)
    opt = O.Opts(sigma_thresh=1e-8, stop_thresh=1e-7)
    march = O.march()
    gen = torch.Generator().manual_seed(123)
    rates, spent = [], 0.0
    for i in range(runs + 1):
        o, d = wl.cams.random_batch(BATCH, gen)
        gt = torch.rand(BATCH, 3, generator=gen)
        t = time.perf_counter()
        march.fused(hg, o.numpy(), d.numpy(), opt, gt.numpy(), threads=threads)
        dt = time.perf_counter() - t
        if i > 0:
            rates.append(BATCH / dt)
            spent += dt
        log(f"CPU baseline (C port, {threads} threads): run {i} {BATCH / dt:.4g} rays/s")
        if i >= runs and spent > budget_s:
            break
    del hg, grid
    # the reference's own CPU path at full size on the 64^3 dense workload
    from svox2_amd import scenes
    torch.set_num_threads(threads)
    sc, o, d, gt = scenes.workload_dense64(seed=0)
    links = sc.links.long()
    ref_rates = []
    for i in range(runs + 1):
        dens = sc.density.float().clone().requires_grad_(True)
        sh = sc.sh.float().clone().requires_grad_(True)
        t = time.perf_counter()
        rgb = O.volume_render_gradcheck_lerp(links, dens, sh, 9, 1.0, 0.0, o, d,
                                             dtype=torch.float32)
        ((rgb - gt.float()) ** 2).mean().backward()
        dt = time.perf_counter() - t
        if i > 0:
            ref_rates.append(o.shape[0] / dt)
        log(f"CPU baseline (reference path, 64^3): run {i} {o.shape[0] / dt:.4g} rays/s")
    model = cpu_model()
    return dict(
        value=float(np.median(rates)), unit="rays/s", cores=threads, kind="port",
        cpu=model,
        sample=f"median of {len(rates)} runs (after 1 warm-up) of random 65536-ray batches "
               f"on the same 512^3 grid: C port of the reference's fused forward+backward, "
               f"float32, {threads} OpenMP threads",
        reference_path=dict(
            value=float(np.median(ref_rates)), unit="rays/s", cores=threads, kind="port",
            sample=f"median of {len(ref_rates)} runs (after 1 warm-up): restated "
                   f"_volume_render_gradcheck_lerp + autograd MSE gradient, float32, "
                   f"torch with {threads} threads, 64^3 dense grid, 4096 rays (full size)"))


def pmc_run(reps: int = 3):
    """A short run for hardware-counter passes: `reps` fused calls on one
    65536-ray batch of the 512^3 grid, then a 1 GiB device copy whose
    FETCH_SIZE / WRITE_SIZE calibrate the counters (2^30 bytes each way)."""
    dev = torch.device("cuda:0")
    wl = Workload(dev)
    batches = wl.batches(BATCH, 1)
    g32 = wl.grid(torch.float32)
    kt = kernel_times(g32, batches, reps)
    x = torch.empty(2 ** 28, dtype=torch.float32, device=dev)
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"pmc_run": {"fused_calls": reps, "fused_ms": [t for t, _ in kt],
                                  "copy_bytes": 2 ** 30}}), flush=True)


ONLY_PARTS = ("small", "side", "image", "serve", "stress", "next")


def next_components(wl, reps: int = 10) -> dict:
    """The components built around the march, timed on the headline workload
    (the 512^3 SH-2 float32 grid, one 65536-ray batch) with HIP events on
    fixed inputs, ms per call:
    - depth_expected / depth_sigma: volume_render_depth in both modes;
    - cuvol_fused, nvol_fused, svox1_fused: the three backends' fused MSE
      steps through SparseGrid.volume_render_fused;
    - cuvol_fused_background: the fused step with a 64-layer MSI background
      of 512 x 256 pixels (131072 columns, 134 MB of float32);
    - msi_tv: the background's TV over 1% of its cells (a contiguous run),
      with algorithmic bytes per cell: 3 links (12 B) and, per channel, 4
      values read (16 B), up to 4 gradient atomics (32 B) and a mask byte;
    - lumisphere_tv: the lumisphere TV kernel over 1% of the grid's cells
      with fixed basis values (cells per second);
    - sgd: sgd_step over every row of density and SH (16 B per element: data
      and gradient each read and written; lr 0 leaves the grid unchanged).
    The grid is left as it was found (no background, cuvol backend)."""
    import svox2_amd
    from svox2_amd import _C
    from torch import nn
    g = wl.grid(torch.float32)
    dev = wl.dev
    o, d, gt = wl.batches(BATCH, 1)[0]
    rays = svox2_amd.Rays(o, d)
    n = o.size(0)

    def rate(ms):
        return dict(ms=ms, rays_per_s=n / (ms * 1e-3))

    out = {}
    with torch.no_grad():
        out["depth_expected"] = rate(_event_ms(lambda: g.volume_render_depth(rays), reps))
        out["depth_sigma"] = rate(_event_ms(lambda: g.volume_render_depth(rays, sigma_thresh=1.0),
                                            reps))
    out["cuvol_fused"] = rate(_event_ms(lambda: g.volume_render_fused(rays, gt), reps))
    try:
        for backend in ("nvol", "svox1"):
            g.opt.backend = backend
            out[f"{backend}_fused"] = rate(_event_ms(lambda: g.volume_render_fused(rays, gt),
                                                     reps))
    finally:
        g.opt.backend = "cuvol"

    reso, nl = 256, 64
    gen = torch.Generator(device=dev).manual_seed(11)
    data = torch.empty(2 * reso * reso, nl, 4, device=dev)
    data[..., :3] = torch.randn(2 * reso * reso, nl, 3, generator=gen, device=dev) * 0.5
    data[..., 3] = torch.rand(2 * reso * reso, nl, generator=gen, device=dev) * 2.0
    try:
        g.background_nlayers, g.background_reso = nl, reso
        g.background_links = torch.arange(2 * reso * reso, dtype=torch.int32,
                                          device=dev).view(2 * reso, reso)
        g.background_data = nn.Parameter(data)
        out["cuvol_fused_background"] = rate(_event_ms(lambda: g.volume_render_fused(rays, gt),
                                                       reps))
        cells = g._get_rand_cells_background(0.01, contiguous=True)
        ms = _event_ms(lambda: _C.msi_tv_grad_sparse(
            g.background_links, g.background_data.data, cells, g.sparse_background_indexer,
            1e-3, 1e-3, g.background_data.grad), reps)
        out["msi_tv"] = dict(_roof(ms, cells.numel() * (12 + 4 * (16 + 32 + 1))),
                             cells=int(cells.numel()))
    finally:
        g.background_nlayers, g.background_links = 0, None
        g.background_data = nn.Parameter(torch.empty(0, 0, 0, device=dev), requires_grad=False)
        g.sparse_background_indexer = None

    cells = g._get_rand_cells(0.01, contiguous=True)
    basis = torch.full((g.basis_dim,), 0.3, device=dev)
    basis_u = torch.full((g.basis_dim,), 0.25, device=dev)
    holder = _C.GridOutputGrads()
    holder.grad_sh_out = g.sh_data.grad
    spec = g._to_cpp()
    ms = _event_ms(lambda: _C.lumisphere_tv_grad_sparse(spec, cells, basis, basis_u, 1e-3, -1.0,
                                                        -1.0, 1.0, holder), reps)
    out["lumisphere_tv"] = dict(ms=ms, cells=int(cells.numel()),
                                cells_per_s=cells.numel() / (ms * 1e-3))
    every = torch.empty((), device=dev)
    elems = g.density_data.numel() + g.sh_data.numel()
    ms = _event_ms(lambda: (_C.sgd_step(g.density_data.data, g.density_data.grad, every, 0.0,
                                        0.0),
                            _C.sgd_step(g.sh_data.data, g.sh_data.grad, every, 0.0, 0.0)), reps)
    out["sgd"] = dict(_roof(ms, 16 * elems), elements=int(elems))
    return out


def run_only(part: str, dev, steps: int, warmup: int) -> dict:
    """One side measurement of the full run, alone: `small` the 5000-ray
    fused steps (float32 and bfloat16), `side` the 256^3 training step and
    the TV / RMSprop / accel_dist_prop / sample kernels, `image` the 800x800
    forward images, `serve` the serving loop at 200 requests/s, `stress` the
    million-ray autograd step, `next` the components around the march."""
    wl = Workload(dev)
    if part == "small":
        small = wl.batches(5000, POOL)
        return {"fp32_5000": measure(wl.grid(torch.float32), small, steps, warmup, 4, 4),
                "bf16_5000": measure(wl.grid(torch.bfloat16), small, steps, warmup, 2, 2)}
    if part == "side":
        return side_kernels(wl, steps, warmup)
    if part == "image":
        return {"image_800_fp32": image_forward(wl.grid(torch.float32), wl.cams),
                "image_800_bf16": image_forward(wl.grid(torch.bfloat16), wl.cams)}
    if part == "serve":
        return serve_bench(wl, qps_list=(200.0,), n=200)
    if part == "stress":
        return stress_1m(wl, wl.grid(torch.float32))
    if part == "next":
        return next_components(wl)
    raise ValueError(part)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline measurement only")
    ap.add_argument("--pmc", action="store_true",
                    help="short run for rocprofv3 counter passes (3 fused calls + 1 GiB copy)")
    ap.add_argument("--only", choices=ONLY_PARTS,
                    help="run one side measurement alone (for per-workload rocprofv3 "
                         "summaries) and print its result as one JSON line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the renderer runs on a single device "
              "(multi-GPU sharding is a design note, not implemented); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    from svox2_amd import _C
    _C.load_library()
    if args.pmc:
        pmc_run()
        return
    dev = torch.device("cuda:0")
    if args.only:
        print(json.dumps({"only": args.only, "result": run_only(args.only, dev, args.steps,
                                                                 args.warmup)}), flush=True)
        return

    wl = Workload(dev)
    batches = wl.batches(BATCH, POOL)
    g32 = wl.grid(torch.float32)
    log("fp32 65536-ray fused steps")
    head = measure(g32, batches, args.steps, args.warmup, 4, 4)
    log(f"fp32: {head['rays_per_s']:.4g} rays/s, kernel {head['kernel_ms']:.3f} ms, "
        f"{head['alg_gbps']:.1f} GB/s algorithmic, counters {head['counters']}")
    extra = {}
    if not args.quick:
        gb = wl.grid(torch.bfloat16)
        log("bf16 65536-ray fused steps")
        extra["bf16_65536"] = measure(gb, batches, args.steps, args.warmup, 2, 2)
        extra["bf16_speedup"] = extra["bf16_65536"]["rays_per_s"] / head["rays_per_s"]
        log(f"bf16: {extra['bf16_65536']['rays_per_s']:.4g} rays/s "
            f"({extra['bf16_speedup']:.2f}x fp32)")
        small = wl.batches(5000, POOL)
        log("5000-ray fused steps")
        extra["fp32_5000"] = measure(g32, small, args.steps, args.warmup, 4, 4)
        extra["bf16_5000"] = measure(gb, small, args.steps, args.warmup, 2, 2)
        log(f"5000-ray batches: fp32 {extra['fp32_5000']['rays_per_s']:.4g}, bf16 "
            f"{extra['bf16_5000']['rays_per_s']:.4g} rays/s")
        log("1M-ray autograd steps with per-ray clips")
        extra["stress_1m_fp32"] = stress_1m(wl, g32)
        log(f"1M rays: {extra['stress_1m_fp32']['ms_per_step']:.2f} ms/step, peak HBM "
            f"{extra['stress_1m_fp32']['peak_hbm_gb']:.2f} GB")
        log("serving loop")
        extra["serve"] = serve_bench(wl)
        for rt in extra["serve"]["rates"]:
            b, p1 = rt["batched"], rt["one_per_launch"]
            log(f"serve {rt['qps']:.0f} QPS: batched p50 {b['p50_ms']:.2f} ms p99 "
                f"{b['p99_ms']:.2f} ms {b['rays_per_s']:.3g} rays/s; one-per-launch p50 "
                f"{p1['p50_ms']:.2f} ms p99 {p1['p99_ms']:.2f} ms {p1['rays_per_s']:.3g} rays/s")
        log("256^3 training step and the kernels around the march")
        extra["side"] = side_kernels(wl, args.steps, args.warmup)
        sd = extra["side"]
        log(f"256^3 training step (5000 rays, TV, RMSprop): {sd['train_256']['ms_per_step']:.3f} ms")
        for k in ("tv_density", "tv_sh", "rmsprop_density", "rmsprop_sh", "accel_dist_prop_512",
                  "sample_512_1m", "sample_backward_512_1m"):
            log(f"{k}: {sd[k]['ms']:.3f} ms, {sd[k]['alg_gbps']:.0f} GB/s algorithmic "
                f"({100 * sd[k]['frac']:.1f}% of HBM peak)")
        log("800x800 image forward")
        extra["image_800_fp32"] = image_forward(g32, wl.cams)
        extra["image_800_bf16"] = image_forward(gb, wl.cams)
        log(f"800x800 image: fp32 {extra['image_800_fp32']['ms']:.3f} ms, bf16 "
            f"{extra['image_800_bf16']['ms']:.3f} ms")
        log("components around the march (depth, nvol, svox1, background, MSI / lumisphere TV, SGD)")
        extra["next"] = next_components(wl)
        for k, v in extra["next"].items():
            log(f"{k}: {v['ms']:.3f} ms")
    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline")
        cpu = cpu_baseline(wl)
    achieved = head["alg_gbps"]
    line = {
        "metric": METRIC,
        "value": head["rays_per_s"],
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "fused fwd+bwd (SparseGrid.volume_render_fused) on a 512^3 sparse "
                        "SH-2 grid, float32 storage, 65536 random camera rays per step",
            "grid": 512,
            "rows": int(wl.scene.density.shape[0]),
            "occupancy": round(wl.scene.occupancy, 4),
            "batch_rays": BATCH,
            "storage": "f32",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(),
            "traffic_source": "profiles/fused_512_f32_65536.md (rocprofv3 --pmc FETCH_SIZE, "
                              "WRITE_SIZE of `bench.py --pmc`, calibrated on a 1 GiB copy)",
            "kernel": "fused call (float32 storage, SH-2): forward ray order "
                      "(k_ray_length_key, k_cost_hist, k_cost_scatter), k_render_rays_rec "
                      "(recording the samples), k_rec_colours, backward ray order "
                      "(k_cost_hist, k_cost_scatter), k_render_backward_replay, "
                      "k_render_backward_mse (rays whose records overflowed)",
            "kernel_ms": head["kernel_ms"],
            "alg_bytes_per_launch": head["alg_bytes_per_launch"],
            "counters_per_launch": head["counters"],
        },
        "serve": ([{"qps": rt["qps"], "p50_ms": rt["batched"]["p50_ms"],
                    "p99_ms": rt["batched"]["p99_ms"], "rays_per_s": rt["batched"]["rays_per_s"],
                    "one_per_launch_p50_ms": rt["one_per_launch"]["p50_ms"],
                    "one_per_launch_p99_ms": rt["one_per_launch"]["p99_ms"]}
                   for rt in extra["serve"]["rates"]] if "serve" in extra else None),
        "cpu_baseline": cpu,
        "detail": {"fp32_65536": head, **extra},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
