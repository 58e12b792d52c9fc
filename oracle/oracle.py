"""TEST INFRASTRUCTURE ONLY - the oracle.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py
may import this module, and only as the checker; the product (svox2_amd/)
never does.

CPU restatements of the reference algorithm for the hot path, each citing the
reference lines it follows:

* ``MarchRef`` - ctypes wrapper of march_ref.c, the FP32 restatement of the
  kernel's index arithmetic (sample positions, skips, density, stop test) and
  per-ray counters; bit-exact with the GPU kernel by construction.
* ``render_fp64`` - the reference's compositing (svox2/svox2.py:658-884 plus the
  kernel's thresholds, render_lerp_kernel_cuvol.cu:30-120) evaluated in FP64
  with PyTorch autograd over the exact FP32 sample list of ``MarchRef``: the
  oracle for colours and gradients.
* ``volume_render_gradcheck_lerp`` - a line-by-line, dtype-generic restatement
  of SparseGrid._volume_render_gradcheck_lerp (svox2/svox2.py:658-884) with the
  grid passed explicitly; the CPU baseline and a cross-check of render_fp64.
* ``depth_fp64`` (render_lerp_kernel_cuvol.cu:122-225), the depth
  renders over the same sample list.
* ``nvol_fp64`` / ``nvol_grads_fp64`` (render_lerp_kernel_nvol.cu:31-292),
  the nvol backend's compositing (differentiable) and its two-pass backward
  (restated), over the kernel's sample list without the transmittance stop.
* ``background_fp64`` / ``render_bg_fp64`` (render_lerp_kernel_cuvol.cu:
  386-607), the MSI background behind the grid's log transmittance, and
  ``msi_tv_grad_sparse`` (loss_kernel.cu:248-333,606-641), its TV.
* ``svox1_fp64`` / ``svox1_grads_fp64`` (render_svox1_kernel.cu:50-275),
  the svox1 backend's nearest-neighbour compositing (differentiable) and its
  backward (restated), over ``ref_svox1_samples``'s FP32 walk.
* ``sample_trilerp`` (svox2.py:604-653), ``eval_sh_bases`` (utils.py:115-162),
  ``morton_code_3``/``gen_morton`` (utils.py:45-79), ``tv_grad_sparse``
  (loss_kernel.cu:178-246,558-604), ``lumisphere_tv_grad_sparse``
  (loss_kernel.cu:336-446,643-678), ``rmsprop_step`` (optim_kernel.cu:15-93),
  ``sgd_step`` (optim_kernel.cu:97-147), ``accel_dist_prop``
  (misc_kernel.cu:109-181,356-416).

Parity pins (tests/test_oracle_*.py): closed-form compositing of uniform
fields, background-only cases, F.grid_sample, exactness on linear fields, the
SH closed form and orthonormality, Morton bijectivity, FP64 finite
differences, autograd of a TV value, brute-force integer pyramids, and
agreement of the two march restatements with each other.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(HERE, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "libmarch_ref.so")

SH_C0 = 0.28209479177387814
SH_C1 = 0.4886025119029199
SH_C2 = (1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
         -1.0925484305920792, 0.5462742152960396)


def build(force: bool = True) -> str:
    """Compiles march_ref.c with gcc (FP32, no contraction) into _build/."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    if not force and os.path.exists(LIB_PATH):
        return LIB_PATH
    src = os.path.join(HERE, "march_ref.c")
    tmp = LIB_PATH + ".tmp"
    cmd = ["gcc", "-O2", "-std=c99", "-fPIC", "-shared", "-ffp-contract=off",
           "-fno-fast-math", "-fopenmp", "-o", tmp, src, "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


# ---------------------------------------------------------------------------
# C restatement (march_ref.c)
# ---------------------------------------------------------------------------
class _RefGrid(ctypes.Structure):
    _fields_ = [
        ("links", ctypes.c_void_p), ("density", ctypes.c_void_p), ("sh", ctypes.c_void_p),
        ("rows", ctypes.c_int64), ("size", ctypes.c_int32 * 3), ("basis_dim", ctypes.c_int32),
        ("offset", ctypes.c_float * 3), ("scaling", ctypes.c_float * 3),
    ]


class _RefOpt(ctypes.Structure):
    _fields_ = [
        ("bg", ctypes.c_float), ("step", ctypes.c_float), ("sigma_thresh", ctypes.c_float),
        ("stop_thresh", ctypes.c_float), ("near_clip", ctypes.c_float),
        ("spheric", ctypes.c_int32), ("last_opaque", ctypes.c_int32), ("skip", ctypes.c_int32),
    ]


@dataclass
class Opts:
    """Render options with the reference's meaning (data_spec.hpp:115-135)."""
    background_brightness: float = 1.0
    step_size: float = 0.5
    sigma_thresh: float = 1e-10
    stop_thresh: float = 1e-7
    near_clip: float = 0.0
    use_spheric_clip: bool = False
    last_sample_opaque: bool = False


@dataclass
class Grid:
    """Host copy of a grid: links (X,Y,Z) int32, density (R,) and sh (R,C) as
    float64 numpy arrays holding exactly the values the GPU stores, plus the
    world->grid affine (p_grid = p * scaling + offset)."""
    links: np.ndarray
    density: np.ndarray
    sh: np.ndarray
    basis_dim: int
    offset: np.ndarray
    scaling: np.ndarray

    @property
    def size(self):
        return self.links.shape


def _c_ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class MarchRef:
    """The FP32 C restatement bound through ctypes."""

    def __init__(self):
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        G, O = ctypes.POINTER(_RefGrid), ctypes.POINTER(_RefOpt)
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        lib.ref_render.argtypes = [G, vp, vp, vp, i64, O, vp, vp, vp]
        lib.ref_render.restype = ctypes.c_int
        lib.ref_samples.argtypes = [G, vp, vp, vp, i64, O, i64, vp, vp, vp, vp, vp, vp]
        lib.ref_samples.restype = ctypes.c_int64
        lib.ref_svox1_samples.argtypes = [G, vp, vp, i64, O, i64, vp, vp, vp, vp]
        lib.ref_svox1_samples.restype = ctypes.c_int64
        lib.ref_ray_setup.argtypes = [G, vp, vp, vp, i64, O, vp]
        lib.ref_ray_setup.restype = ctypes.c_int
        lib.ref_camera_rays.argtypes = [vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_int32, ctypes.c_int32, vp, vp]
        lib.ref_camera_rays.restype = ctypes.c_int
        lib.ref_fused.argtypes = [G, vp, vp, vp, i64, O, vp, ctypes.c_float, ctypes.c_float,
                                  vp, vp, vp, vp, ctypes.c_int32]
        lib.ref_fused.restype = ctypes.c_int
        f32 = ctypes.c_float
        lib.ref_grid_weight_render.argtypes = [vp, vp, vp, f32, f32, f32, f32, ctypes.c_int32,
                                               ctypes.c_int32, f32, f32, ctypes.c_int32, vp, vp,
                                               vp]
        lib.ref_grid_weight_render.restype = ctypes.c_int
        self.lib = lib

    @staticmethod
    def _grid(g: Grid, keep: list) -> _RefGrid:
        links = np.ascontiguousarray(g.links, dtype=np.int32)
        dens = np.ascontiguousarray(g.density, dtype=np.float32).reshape(-1)
        sh = None if g.sh is None else np.ascontiguousarray(g.sh, dtype=np.float32)
        keep += [links, dens, sh]
        s = _RefGrid()
        s.links, s.density, s.sh = (links.ctypes.data, dens.ctypes.data,
                                    None if sh is None else sh.ctypes.data)
        s.rows = dens.shape[0]
        for i in range(3):
            s.size[i] = links.shape[i]
            s.offset[i] = float(np.float32(g.offset[i]))
            s.scaling[i] = float(np.float32(g.scaling[i]))
        s.basis_dim = g.basis_dim
        return s

    @staticmethod
    def _opt(o: Opts, skip: bool = True) -> _RefOpt:
        s = _RefOpt()
        s.bg, s.step, s.sigma_thresh = o.background_brightness, o.step_size, o.sigma_thresh
        s.stop_thresh, s.near_clip = o.stop_thresh, o.near_clip
        s.spheric, s.last_opaque, s.skip = int(o.use_spheric_clip), int(o.last_sample_opaque), int(skip)
        return s

    @staticmethod
    def _rays(origins, dirs, clip, keep):
        o = np.ascontiguousarray(origins, dtype=np.float32)
        d = np.ascontiguousarray(dirs, dtype=np.float32)
        c = None if clip is None else np.ascontiguousarray(clip, dtype=np.float32)
        keep += [o, d, c]
        return o, d, c

    def render(self, g: Grid, origins, dirs, opt: Opts, clip=None, skip: bool = True):
        """FP32 colours (n,3), final log transmittance (n,), counters (n,4)."""
        keep: list = []
        gs = self._grid(g, keep)
        o, d, c = self._rays(origins, dirs, clip, keep)
        n = o.shape[0]
        rgb = np.zeros((n, 3), np.float32)
        lt = np.zeros((n,), np.float32)
        st = np.zeros((n, 4), np.int32)
        self.lib.ref_render(ctypes.byref(gs), _c_ptr(o), _c_ptr(d), _c_ptr(c), n,
                            ctypes.byref(self._opt(opt, skip)), _c_ptr(rgb), _c_ptr(lt),
                            _c_ptr(st))
        return rgb, lt, st

    def samples(self, g: Grid, origins, dirs, opt: Opts, clip=None, skip: bool = True):
        """Every evaluated sample, in ray order: dict of ray (S,), t (S,),
        l (S,3) int, p (S,3) float32, ws (S,), flag (S,) (bit 0: FP32 density
        above sigma_thresh; bit 1: the FP32 march has not stopped before it)."""
        parts = list(self.sample_blocks(g, origins, dirs, opt, clip=clip, skip=skip))
        out = {k: np.concatenate([q[2][k] for q in parts]) for k in parts[0][2]}
        out["ray"] = np.concatenate([q[2]["ray"] + q[0] for q in parts])
        return out

    def svox1_samples(self, g: Grid, origins, dirs, opt: Opts) -> dict:
        """The svox1 walk's occupied-cell samples (ref_svox1_samples): dict of
        ray, link, dt (world step), flag (1 counted, 2 forward stop, 4
        backward stop, 8 opaque last)."""
        keep: list = []
        gs = self._grid(g, keep)
        o, d, _ = self._rays(origins, dirs, None, keep)
        ops = self._opt(opt)
        cap = max(1024, o.shape[0] * 64)
        while True:
            out = dict(ray=np.empty(cap, np.int32), link=np.empty(cap, np.int32),
                       dt=np.empty(cap, np.float32), flag=np.empty(cap, np.uint8))
            cnt = self.lib.ref_svox1_samples(ctypes.byref(gs), _c_ptr(o), _c_ptr(d), o.shape[0],
                                             ctypes.byref(ops), cap,
                                             *(_c_ptr(out[k]) for k in ("ray", "link", "dt",
                                                                         "flag")))
            if cnt <= cap:
                return {k: v[:cnt] for k, v in out.items()}
            cap = int(cnt)

    def sample_blocks(self, g: Grid, origins, dirs, opt: Opts, clip=None, skip: bool = True,
                      block: int = 2048, post=None):
        """The samples of ``samples`` block by block: yields (r0, r1, samples of
        rays [r0, r1) with ray ids relative to r0) in ray order, or (r0, r1,
        *post(samples)) with ``post``.  Rays are independent, so blocks are
        marched on worker threads (ctypes releases the GIL), ``post`` too."""
        keep: list = []
        gs = self._grid(g, keep)
        o, d, c = self._rays(origins, dirs, clip, keep)
        n = o.shape[0]
        ops = self._opt(opt, skip)

        def run(b):
            r0, r1 = b
            nb = r1 - r0
            oc, dc = _c_ptr(o[r0:r1]), _c_ptr(d[r0:r1])
            cc = None if c is None else _c_ptr(c[r0:r1])
            cap = max(1024, nb * 64)
            while True:
                out = dict(ray=np.empty(cap, np.int32), t=np.empty(cap, np.float32),
                           l=np.empty((cap, 3), np.int32), p=np.empty((cap, 3), np.float32),
                           ws=np.empty(cap, np.float32), flag=np.empty(cap, np.uint8))
                cnt = self.lib.ref_samples(ctypes.byref(gs), oc, dc, cc, nb, ctypes.byref(ops),
                                           cap, *(_c_ptr(out[k]) for k in
                                                  ("ray", "t", "l", "p", "ws", "flag")))
                if cnt <= cap:
                    break
                cap = int(cnt)
            out = {k: v[:cnt] for k, v in out.items()}
            return (r0, r1, out) if post is None else (r0, r1, *post(out))

        bounds = [(r0, min(n, r0 + block)) for r0 in range(0, n, block)] or [(0, 0)]
        if len(bounds) == 1:
            yield run(bounds[0])
            return
        from concurrent.futures import ThreadPoolExecutor
        workers = max(1, min(16, len(os.sched_getaffinity(0))))
        with ThreadPoolExecutor(workers) as ex:
            yield from ex.map(run, bounds)

    def ray_setup(self, g: Grid, origins, dirs, opt: Opts, clip=None) -> np.ndarray:
        keep: list = []
        gs = self._grid(g, keep)
        o, d, c = self._rays(origins, dirs, clip, keep)
        out = np.zeros((o.shape[0], 9), np.float32)
        self.lib.ref_ray_setup(ctypes.byref(gs), _c_ptr(o), _c_ptr(d), _c_ptr(c), o.shape[0],
                               ctypes.byref(self._opt(opt)), _c_ptr(out))
        return out

    def fused(self, g: Grid, origins, dirs, opt: Opts, rgb_gt, clip=None, beta_loss=0.0,
              sparsity_loss=0.0, threads: int = 1):
        """FP32 port of the fused forward + MSE backward: rgb (n,3),
        grad_density (R,), grad_sh (R,C), touched mask (R,); rays split over
        `threads` OpenMP threads (1: sequential, deterministic)."""
        keep: list = []
        gs = self._grid(g, keep)
        o, d, c = self._rays(origins, dirs, clip, keep)
        gt = np.ascontiguousarray(rgb_gt, dtype=np.float32)
        n = o.shape[0]
        R = np.asarray(g.density).reshape(-1).shape[0]
        C = 3 * g.basis_dim
        rgb = np.zeros((n, 3), np.float32)
        gd = np.zeros(R, np.float32)
        gsh = np.zeros((R, C), np.float32)
        m = np.zeros(R, np.uint8)
        self.lib.ref_fused(ctypes.byref(gs), _c_ptr(o), _c_ptr(d), _c_ptr(c), n,
                           ctypes.byref(self._opt(opt)), _c_ptr(gt), beta_loss, sparsity_loss,
                           _c_ptr(rgb), _c_ptr(gd), _c_ptr(gsh), _c_ptr(m), int(threads))
        return rgb, gd, gsh, m.astype(bool)

    def camera_rays(self, c2w, fx, fy, cx, cy, width, height):
        m = np.ascontiguousarray(np.asarray(c2w, dtype=np.float32)[:3, :4])
        o = np.zeros((width * height, 3), np.float32)
        d = np.zeros((width * height, 3), np.float32)
        self.lib.ref_camera_rays(_c_ptr(m), fx, fy, cx, cy, width, height, _c_ptr(o), _c_ptr(d))
        return o, d

    def grid_weight_render(self, data, c2w, fx, fy, cx, cy, width, height, step_size,
                           stop_thresh, last_opaque, offset, scaling, out=None):
        """Raises out (float32, data's shape; zeros if None) to the per-voxel
        maximum ray weight of one camera through the dense volume data."""
        data = np.ascontiguousarray(data, np.float32)
        out = np.zeros_like(data) if out is None else out
        assert out.dtype == np.float32 and out.shape == data.shape and out.flags.c_contiguous
        size = np.array(data.shape, np.int32)
        m = np.ascontiguousarray(np.asarray(c2w, dtype=np.float32)[:3, :4])
        off = np.asarray(offset, np.float32).reshape(3).copy()
        sc = np.asarray(scaling, np.float32).reshape(3).copy()
        self.lib.ref_grid_weight_render(_c_ptr(data), _c_ptr(size), _c_ptr(m), fx, fy, cx, cy,
                                        width, height, step_size, stop_thresh, int(last_opaque),
                                        _c_ptr(off), _c_ptr(sc), _c_ptr(out))
        return out


_MARCH: Optional[MarchRef] = None


def march() -> MarchRef:
    global _MARCH
    if _MARCH is None:
        _MARCH = MarchRef()
    return _MARCH


# ---------------------------------------------------------------------------
# SH basis, Morton codes (svox2/utils.py)
# ---------------------------------------------------------------------------
def eval_sh_bases(basis_dim: int, dirs: torch.Tensor) -> torch.Tensor:
    """Real SH basis, degree <= 2 (svox2/utils.py:115-162)."""
    out = torch.empty((*dirs.shape[:-1], basis_dim), dtype=dirs.dtype, device=dirs.device)
    out[..., 0] = SH_C0
    if basis_dim > 1:
        x, y, z = dirs.unbind(-1)
        out[..., 1] = -SH_C1 * y
        out[..., 2] = SH_C1 * z
        out[..., 3] = -SH_C1 * x
        if basis_dim > 4:
            out[..., 4] = SH_C2[0] * x * y
            out[..., 5] = SH_C2[1] * y * z
            out[..., 6] = SH_C2[2] * (2.0 * z * z - x * x - y * y)
            out[..., 7] = SH_C2[3] * x * z
            out[..., 8] = SH_C2[4] * (x * x - y * y)
    return out


def morton_code_3(x: int, y: int, z: int) -> int:
    """Bit interleave, x most significant per triple (svox2/utils.py:46-66),
    written as a plain loop over bits."""
    code = 0
    for b in range(10):
        code |= ((x >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b)
    return code


def gen_morton(D: int) -> np.ndarray:
    a = np.arange(D)
    X, Y, Z = np.meshgrid(a, a, a, indexing="ij")
    out = np.zeros((D, D, D), np.int64)
    for b in range(10):
        out |= ((X >> b) & 1) << (3 * b + 2)
        out |= ((Y >> b) & 1) << (3 * b + 1)
        out |= ((Z >> b) & 1) << (3 * b)
    return out


# ---------------------------------------------------------------------------
# FP64 renderer over the exact sample list
# ---------------------------------------------------------------------------
_CORNERS = [(dx, dy, dz) for dx in (0, 1) for dy in (0, 1) for dz in (0, 1)]


def corner_gather(links: np.ndarray, l: np.ndarray) -> np.ndarray:
    """(S, 8) corner links of cells l (S, 3), corner order (dx, dy, dz)."""
    links = np.ascontiguousarray(links)
    _, n1, n2 = links.shape
    base = (l[:, 0].astype(np.int64) * n1 + l[:, 1]) * n2 + l[:, 2]
    offs = np.array([(dx * n1 + dy) * n2 + dz for dx, dy, dz in _CORNERS], np.int64)
    return links.reshape(-1)[base[:, None] + offs[None, :]]


def corner_weights(p: torch.Tensor) -> torch.Tensor:
    """(S, 8) trilinear weights of fractional offsets p (S, 3)."""
    ws = []
    for dx, dy, dz in _CORNERS:
        wx = p[:, 0] if dx else 1 - p[:, 0]
        wy = p[:, 1] if dy else 1 - p[:, 1]
        wz = p[:, 2] if dz else 1 - p[:, 2]
        ws.append(wx * wy * wz)
    return torch.stack(ws, dim=1)


def _composite(grid: Grid, dirs: np.ndarray, opt: Opts, density: torch.Tensor,
               sh: torch.Tensor, smp: dict, n: int, thresholds: bool,
               extra: Optional[dict] = None) -> torch.Tensor:
    """Colours of n rays from their samples (ray ids 0..n-1, grouped by ray);
    ``extra`` (if given) receives per-ray optical depth and end transmittance
    and per-sample sigma and counted flags."""
    B = grid.basis_dim
    dt = torch.float64
    bg = float(opt.background_brightness)
    ray = torch.from_numpy(smp["ray"].astype(np.int64))
    S = ray.numel()
    if S == 0:
        if extra is not None:
            z = torch.zeros(n, dtype=dt)
            extra.update(depth=z, t_end=torch.ones(n, dtype=dt), sigma=torch.zeros(0, dtype=dt),
                         counted=torch.zeros(0, dtype=torch.bool),
                         stopped=torch.zeros(n, dtype=torch.bool))
        return torch.full((n, 3), bg, dtype=dt) + 0 * (density.sum() + sh.sum())
    basis = eval_sh_bases(B, torch.from_numpy(dirs).to(dt))  # at the raw dirs, as the kernel
    lk = smp["lk"] if "lk" in smp else corner_gather(grid.links, smp["l"])
    lk = torch.from_numpy(lk.astype(np.int64))
    rows = density.shape[0]
    valid = (lk >= 0) & (lk < rows)
    lk0 = torch.where(valid, lk, torch.zeros_like(lk))
    w = corner_weights(torch.from_numpy(smp["p"]).to(dt)) * valid.to(dt)
    sigma = (w * density[lk0]).sum(1)
    sh_s = (w[..., None] * sh[lk0]).sum(1)
    col = (sh_s.view(S, 3, B) * basis[ray][:, None, :]).sum(-1) + 0.5
    col = torch.clamp_min(col, 0.0)
    ws = torch.from_numpy(smp["ws"].astype(np.float64))
    if thresholds:
        contrib = (sigma > opt.sigma_thresh).detach()
        pcnt = torch.where(contrib, ws * sigma, torch.zeros_like(sigma))
    else:
        contrib = torch.ones(S, dtype=torch.bool)
        pcnt = ws * torch.relu(sigma)
    # exclusive per-ray prefix of optical depth, summed inside each ray only
    first = torch.ones(S, dtype=torch.bool)
    first[1:] = ray[1:] != ray[:-1]
    start_idx = torch.nonzero(first).flatten()
    seg = torch.cumsum(first.long(), 0) - 1
    pos = torch.arange(S) - start_idx[seg]
    L = int(pos.max()) + 1
    pad = torch.zeros((start_idx.numel(), L), dtype=dt)
    pad = pad.index_put((seg, pos), pcnt)
    excl = (torch.cumsum(pad, 1) - pad)[seg, pos]
    log_before = -excl
    log_after = log_before - pcnt
    if thresholds and opt.stop_thresh > 0:
        stop = (contrib & (torch.exp(log_after) < opt.stop_thresh)).detach()
        sb = torch.cumsum(stop.long(), 0) - stop.long()
        alive = (sb - sb[start_idx][seg]) == 0
        ray_stopped = torch.zeros(n, dtype=torch.long).index_add_(0, ray, stop.long()) > 0
    else:
        alive = torch.ones(S, dtype=torch.bool)
        ray_stopped = torch.zeros(n, dtype=torch.bool)
    weight = torch.exp(log_before) * (1.0 - torch.exp(-pcnt))
    cw = weight * (alive & contrib).to(dt)
    rgb = torch.zeros((n, 3), dtype=dt).index_add(0, ray, cw[:, None] * col)
    depth = torch.zeros(n, dtype=dt).index_add(0, ray, pcnt * alive.to(dt))
    t_end = torch.exp(-depth) * (~ray_stopped).to(dt)
    if extra is not None:
        extra.update(depth=depth, t_end=t_end, sigma=sigma, counted=(alive & contrib),
                     stopped=ray_stopped)
    return rgb + t_end[:, None] * bg


def _ray_chunks(smp: dict, n: int, max_samples: int):
    """Splits rays [0, n) into ranges holding at most ~max_samples samples."""
    counts = np.bincount(smp["ray"], minlength=n)
    bounds = np.concatenate([[0], np.cumsum(counts)])
    r0 = 0
    while r0 < n:
        r1 = r0 + 1
        while r1 < n and bounds[r1 + 1] - bounds[r0] <= max_samples:
            r1 += 1
        s0, s1 = bounds[r0], bounds[r1]
        sub = {k: v[s0:s1] for k, v in smp.items()}
        sub["ray"] = sub["ray"] - r0
        yield r0, r1, sub
        r0 = r1


def _compact_rows(grid: Grid, smp: dict, rows: int):
    """Drops the samples none of whose 8 corners has a row and renumbers the
    others' corner links to the rows the chunk uses.  A dropped sample has
    sigma = 0 exactly, so with sigma_thresh >= 0 it neither contributes nor
    moves the transmittance: the chunk's colours and gradients are unchanged,
    and autograd works on the used rows instead of the whole grid.  Returns
    the compacted samples (with "lk", local links, -1 for none) and the used
    global rows."""
    lk = corner_gather(grid.links, smp["l"]).astype(np.int64)
    valid = (lk >= 0) & (lk < rows)
    keep = valid.any(1)
    smp = {k: v[keep] for k, v in smp.items()}
    lk, valid = lk[keep], valid[keep]
    used, inv = np.unique(lk[valid], return_inverse=True)
    local = np.full(lk.shape, -1, np.int64)
    local[valid] = inv.reshape(-1)
    smp["lk"] = local
    return smp, torch.from_numpy(used)


def render_fp64(grid: Grid, origins, dirs, opt: Opts, density: torch.Tensor,
                sh: torch.Tensor, clip=None, samples=None, thresholds: bool = True,
                max_samples: int = 2_000_000, return_log_t: bool = False):
    """FP64 colours (n, 3) of the reference compositing over the kernel's exact
    sample positions; differentiable in ``density`` (R,) and ``sh`` (R, C).

    Compositing (svox2.py:767-785 and render_lerp_kernel_cuvol.cu:85-107):
    w_i = T_i (1 - exp(-ws_i sigma_i)), colour_i = max(sh_i . basis + 0.5, 0),
    C = sum_i w_i colour_i + T_end bg.  With ``thresholds`` (kernel semantics)
    samples with sigma <= sigma_thresh are skipped and a ray stops after the
    first sample whose transmittance falls below stop_thresh (T_end = 0);
    without, sigma is clamped at 0 and every sample counts (the reference's
    pure-PyTorch semantics).  With ``return_log_t`` also each ray's final log
    transmittance (n,): minus its optical depth, or -1e3 for a ray the stop
    test ended (render_lerp_kernel_cuvol.cu:103-105), differentiable likewise.
    """
    origins = np.asarray(origins, np.float32)
    dirs = np.asarray(dirs, np.float32)
    n = origins.shape[0]
    compact = thresholds and opt.sigma_thresh >= 0
    outs, log_ts = [], []
    for b0, b1, bsub, used in _blocks(grid, origins, dirs, opt, clip, samples, max_samples,
                                      density.shape[0], compact):
        dens, shb = (density, sh) if used is None else (density[used], sh[used])
        for c0, c1, sub in _ray_chunks(bsub, b1 - b0, max_samples):
            r0, r1 = b0 + c0, b0 + c1
            extra = {}
            outs.append(_composite(grid, dirs[r0:r1], opt, dens, shb, sub, r1 - r0, thresholds,
                                   extra))
            log_ts.append(torch.where(extra["stopped"], torch.full_like(extra["depth"], -1e3),
                                      -extra["depth"]))
    rgb = torch.cat(outs, 0) if outs else torch.zeros((0, 3), dtype=torch.float64)
    if return_log_t:
        return rgb, (torch.cat(log_ts, 0) if log_ts else torch.zeros(0, dtype=torch.float64))
    return rgb


def _blocks(grid: Grid, origins, dirs, opt: Opts, clip, samples, max_samples: int, rows: int,
            compact: bool):
    """(r0, r1, samples of rays [r0, r1) relative to r0, used rows) in ray
    order: marched block by block when ``samples`` is None, else cut from it;
    compacted by ``_compact_rows`` when ``compact``, else with used rows None
    (every row, in the caller's own numbering)."""
    post = (lambda smp: _compact_rows(grid, smp, rows)) if compact else (lambda smp: (smp, None))
    if samples is None:
        yield from march().sample_blocks(grid, origins, dirs, opt, clip=clip, post=post)
    else:
        for r0, r1, sub in _ray_chunks(samples, np.asarray(origins).shape[0], max_samples):
            yield (r0, r1, *post(sub))


def nvol_samples(grid: Grid, origins, dirs, opt: Opts, clip=None) -> dict:
    """The nvol kernel's samples: the colour march's placement to tmax
    (render_lerp_kernel_nvol.cu:52-71), without an opaque last sample; flag bit
    0 is the FP32 decision sigma > sigma_thresh."""
    import dataclasses
    return march().samples(grid, origins, dirs, dataclasses.replace(opt, last_sample_opaque=False),
                           clip=clip)


def _nvol_sample_terms(grid: Grid, dirs, smp: dict, density, sh):
    """Per-sample FP64 density (S,), colour + 0.5 before clamping (S, 3),
    corner links (S, 8) with their validity and trilinear weights, and the SH
    basis per ray."""
    B = grid.basis_dim
    lk = corner_gather(grid.links, smp["l"]).astype(np.int64)
    rows = density.shape[0]
    valid = (lk >= 0) & (lk < rows)
    lk0 = np.where(valid, lk, 0)
    dt = density.dtype if isinstance(density, torch.Tensor) else torch.float64
    w = corner_weights(torch.from_numpy(smp["p"]).to(torch.float64)) * \
        torch.from_numpy(valid).to(torch.float64)
    basis = eval_sh_bases(B, torch.from_numpy(np.asarray(dirs, np.float32)).to(torch.float64))
    ray = torch.from_numpy(smp["ray"].astype(np.int64))
    lkt = torch.from_numpy(lk0)
    sigma = (w * density[lkt]).sum(1)
    shs = (w[..., None] * sh[lkt]).sum(1)
    col = (shs.view(-1, 3, B) * basis[ray][:, None, :]).sum(-1) + 0.5
    return sigma, col, lk, valid, w, basis, dt


def nvol_fp64(grid: Grid, origins, dirs, opt: Opts, density: torch.Tensor, sh: torch.Tensor,
              clip=None, samples=None) -> torch.Tensor:
    """FP64 colours (n, 3) of the nvol compositing (render_lerp_kernel_nvol.cu:
    31-107) over the kernel's exact samples, differentiable in ``density``
    (R,) and ``sh`` (R, C): over the samples with FP32 sigma > sigma_thresh,
    A = min(A + 1 - exp(-ws sigma), 1), weight = the increase of A, colour =
    max(sh . basis + 0.5, 0); after A reaches 1 nothing is added; the
    background adds (1 - A) bg."""
    origins = np.asarray(origins, np.float32)
    dirs = np.asarray(dirs, np.float32)
    n = origins.shape[0]
    smp = nvol_samples(grid, origins, dirs, opt, clip) if samples is None else samples
    bg = float(opt.background_brightness)
    S = smp["ray"].size
    if S == 0:
        return torch.full((n, 3), bg, dtype=torch.float64) + 0 * (density.sum() + sh.sum())
    sigma, col, _, _, _, _, _ = _nvol_sample_terms(grid, dirs, smp, density, sh)
    ws = torch.from_numpy(smp["ws"].astype(np.float64))
    contrib = torch.from_numpy((smp["flag"] & 1).astype(bool))
    a = torch.where(contrib, 1.0 - torch.exp(-ws * sigma), torch.zeros_like(sigma))
    ray = torch.from_numpy(smp["ray"].astype(np.int64))
    first = torch.ones(S, dtype=torch.bool)
    first[1:] = ray[1:] != ray[:-1]
    start = torch.nonzero(first).flatten()
    seg = torch.cumsum(first.long(), 0) - 1
    pos = torch.arange(S) - start[seg]
    L = int(pos.max()) + 1
    rays_with = ray[start]
    A = torch.zeros((start.numel(), L), dtype=torch.float64).index_put((seg, pos), a)
    alpha = torch.zeros(start.numel(), dtype=torch.float64)
    weights = []
    for k in range(L):
        # a saturated ray takes no more samples (the kernel breaks out of its
        # loop): leave its opacity as it is rather than clamp it again
        sat = (alpha >= 1.0).detach()
        na = torch.where(sat, alpha, torch.clamp(alpha + A[:, k], max=1.0))
        weights.append(na - alpha)
        alpha = na
    W = torch.stack(weights, 1)[seg, pos]
    rgb = torch.zeros((n, 3), dtype=torch.float64).index_add(
        0, ray, (W * contrib.to(torch.float64))[:, None] * torch.clamp_min(col, 0.0))
    alpha_r = torch.zeros(n, dtype=torch.float64).index_add(0, rays_with, alpha)
    return rgb + (1.0 - alpha_r)[:, None] * bg


def nvol_grads_fp64(grid: Grid, origins, dirs, opt: Opts, grad_out, clip=None,
                    sparsity: float = 0.0, samples=None):
    """(grad_density (R,), grad_sh (R, C), mask (R,)) of the reference's nvol
    backward (trace_ray_nvol_backward, render_lerp_kernel_nvol.cu:109-292)
    restated in FP64, sample by sample.  Pass 1 over the contributing samples
    up to the saturating one: SH coefficient j of channel c gets
    weight [colour_c >= 0] g_c basis_j (:171-199); unless the sample saturates,
    the density gets ws e T + sparsity 4 sigma / (1 + 2 sigma^2) with e =
    exp(-ws sigma), T = (t_0 + t_2) + t_1, t_c = g_c max(colour_c, 0) (:201-220);
    the saturating sample's T, or bg ((g_0 + g_1) + g_2) if none saturates, is
    L (:220-231).  Pass 2 if L != 0, over every sample up to the saturating
    one: its own opacity accumulates, the pass stops at the sample that
    reaches 1, and the others' densities get -ws e L (:232-291).  Every
    density term marks its corners' rows."""
    origins = np.asarray(origins, np.float32)
    dirs = np.asarray(dirs, np.float32)
    n = origins.shape[0]
    density = torch.as_tensor(np.asarray(grid.density, np.float64).reshape(-1))
    sh = torch.as_tensor(np.asarray(grid.sh, np.float64))
    R, C = sh.shape
    B = grid.basis_dim
    gd = np.zeros(R)
    gs = np.zeros((R, C))
    mask = np.zeros(R, bool)
    smp = nvol_samples(grid, origins, dirs, opt, clip) if samples is None else samples
    if smp["ray"].size == 0:
        return gd, gs, mask
    sigma, col, lk, valid, w, basis, _ = _nvol_sample_terms(grid, dirs, smp, density, sh)
    sigma, col, w, basis = sigma.numpy(), col.numpy(), w.numpy(), basis.numpy()
    ws = smp["ws"].astype(np.float64)
    contrib = (smp["flag"] & 1).astype(bool)
    g = np.asarray(grad_out, np.float64)
    bg = float(opt.background_brightness)
    bounds = np.concatenate([[0], np.cumsum(np.bincount(smp["ray"], minlength=n))])

    def add_density(i, v):
        for k in range(8):
            if valid[i, k]:
                gd[lk[i, k]] += w[i, k] * v
                mask[lk[i, k]] = True

    for r in range(n):
        idx = range(bounds[r], bounds[r + 1])
        alpha, last, sat = 0.0, 0.0, None
        for i in idx:
            if not contrib[i]:
                continue
            e = math.exp(-ws[i] * sigma[i])
            na = min(alpha + 1.0 - e, 1.0)
            wgt = na - alpha
            alpha = na
            cl = np.maximum(col[i], 0.0)
            tc = cl * g[r]
            tot = (tc[0] + tc[2]) + tc[1]
            gcom = wgt * (cl == col[i]) * g[r]
            val = (gcom[:, None] * basis[r][None, :]).reshape(C)
            for k in range(8):
                if valid[i, k]:
                    gs[lk[i, k]] += w[i, k] * val
            if na < 1.0:
                v = ws[i] * e * tot
                if sparsity > 0.0:
                    v += sparsity * (4.0 * sigma[i] / (1.0 + 2.0 * sigma[i] * sigma[i]))
                add_density(i, v)
            else:
                last, sat = tot, i
                break
        if sat is None:
            last = bg * ((g[r][0] + g[r][1]) + g[r][2])
        if last != 0.0:
            alpha2 = 0.0
            for i in idx:
                if sat is not None and i > sat:
                    break
                e = math.exp(-ws[i] * sigma[i])
                alpha2 = min(alpha2 + 1.0 - e, 1.0)
                if alpha2 >= 1.0:
                    break
                add_density(i, -ws[i] * e * last)
    return gd, gs, mask


def depth_fp64(grid: Grid, origins, dirs, opt: Opts, sigma_thresh: Optional[float] = None,
               clip=None) -> np.ndarray:
    """FP64 depth of each ray (render_lerp_kernel_cuvol.cu:122-225) over the
    FP32 march's sample list, whose FP32 decisions (which samples are
    evaluated, which contribute, where the ray stops) are the kernel's:
      sigma_thresh None: sum over counted samples of
        w (t / step) ws, w = exp(-sum of earlier ws sigma) (1 - exp(-ws sigma)),
        sigma trilinear in FP64, with last_sample_opaque off (the reference's
        depth march has none);
      sigma_thresh s: (t / step) ws of the first sample whose FP32 density
        exceeds s, marched without the transmittance stop; 0 if none.
    Rays that miss the grid get 0."""
    from dataclasses import replace
    origins = np.asarray(origins, np.float32)
    dirs = np.asarray(dirs, np.float32)
    n = origins.shape[0]
    out = np.zeros(n, np.float64)
    step = float(np.float32(opt.step_size))
    if sigma_thresh is not None:
        o2 = replace(opt, sigma_thresh=float(sigma_thresh), stop_thresh=0.0,
                     last_sample_opaque=False)
        smp = march().samples(grid, origins, dirs, o2, clip=clip)
        hit = np.nonzero(smp["flag"] & 1)[0]
        ray = smp["ray"][hit]
        first = np.ones(hit.size, bool)
        first[1:] = ray[1:] != ray[:-1]
        k = hit[first]
        out[smp["ray"][k]] = (smp["t"][k].astype(np.float64) / step) * \
            smp["ws"][k].astype(np.float64)
        return out
    o2 = replace(opt, last_sample_opaque=False)
    smp = march().samples(grid, origins, dirs, o2, clip=clip)
    counted = (smp["flag"] & 3) == 3
    ray = smp["ray"][counted].astype(np.int64)
    if ray.size == 0:
        return out
    lk = corner_gather(grid.links, smp["l"][counted]).astype(np.int64)
    dens = np.asarray(grid.density, np.float64).reshape(-1)
    valid = (lk >= 0) & (lk < dens.shape[0])
    w8 = corner_weights(torch.from_numpy(smp["p"][counted]).double()).numpy() * valid
    sigma = (w8 * dens[np.where(valid, lk, 0)]).sum(1)
    ws = smp["ws"][counted].astype(np.float64)
    od = ws * sigma
    # exclusive per-ray cumulative optical depth (samples are grouped by ray)
    csum = np.cumsum(od)
    first = np.ones(ray.size, bool)
    first[1:] = ray[1:] != ray[:-1]
    start = np.maximum.accumulate(np.where(first, np.arange(ray.size), 0))
    before = csum - od - (csum[start] - od[start])
    w = np.exp(-before) * (1.0 - np.exp(-od))
    t = smp["t"][counted].astype(np.float64)
    np.add.at(out, ray, w * (t / step) * ws)
    return out


def touched_rows(grid: Grid, samples: dict) -> np.ndarray:
    """Bool (R,) mask of rows that are a corner of a contributing sample of the
    FP32 march (the kernel's mask_out, render_util.cuh:137-140)."""
    sel = (samples["flag"] & 3) == 3
    lk = corner_gather(grid.links, samples["l"][sel]).reshape(-1)
    R = np.asarray(grid.density).reshape(-1).shape[0]
    lk = lk[(lk >= 0) & (lk < R)]
    m = np.zeros(R, bool)
    m[lk] = True
    return m


def mse_grads_fp64(grid: Grid, origins, dirs, opt: Opts, rgb_gt, clip=None, samples=None,
                   max_samples: int = 2_000_000, beta_loss: float = 0.0,
                   sparsity_loss: float = 0.0):
    """FP64 colours and the gradient, by autograd and ray chunk by ray chunk, of
    mean((rgb - gt)^2) plus the two terms the fused kernel folds in
    (render_lerp_kernel_cuvol.cu:257-262, :345-348):
      beta:     sum_r b_r log T_r, b_r = beta/n (1 - T_r / (1 - T_r + 1e-3)) held
                constant, T_r the ray's final transmittance (0 once stopped);
      sparsity: sparsity * sum of log(1 + 2 sigma^2) over contributing samples."""
    origins = np.asarray(origins, np.float32)
    dirs = np.asarray(dirs, np.float32)
    n = origins.shape[0]
    gt = torch.as_tensor(np.asarray(rgb_gt, np.float64))
    dens_all = torch.from_numpy(np.asarray(grid.density, np.float64).reshape(-1))
    sh_all = torch.from_numpy(np.asarray(grid.sh, np.float64))
    gd = torch.zeros_like(dens_all)
    gs = torch.zeros_like(sh_all)
    rgbs = []
    for b0, b1, bsub, used in _blocks(grid, origins, dirs, opt, clip, samples, max_samples,
                                      dens_all.shape[0], opt.sigma_thresh >= 0):
        dens = (dens_all.clone() if used is None else dens_all[used]).requires_grad_(True)
        sh = (sh_all.clone() if used is None else sh_all[used]).requires_grad_(True)
        for c0, c1, sub in _ray_chunks(bsub, b1 - b0, max_samples):
            r0, r1 = b0 + c0, b0 + c1
            extra = {}
            rgb = _composite(grid, dirs[r0:r1], opt, dens, sh, sub, r1 - r0, True, extra)
            loss = ((rgb - gt[r0:r1]) ** 2).sum() / (3.0 * n)
            if beta_loss > 0:
                t_end = extra["t_end"].detach()
                b = beta_loss / n * (1.0 - t_end / (1.0 - t_end + 1e-3))
                loss = loss - (b * extra["depth"]).sum()
            if sparsity_loss > 0:
                s_ = extra["sigma"][extra["counted"]]
                loss = loss + sparsity_loss * torch.log1p(2.0 * s_ * s_).sum()
            loss.backward()
            rgbs.append(rgb.detach())
        if dens.grad is not None:
            gd.index_add_(0, used, dens.grad) if used is not None else gd.add_(dens.grad)
        if sh.grad is not None:
            gs.index_add_(0, used, sh.grad) if used is not None else gs.add_(sh.grad)
    rgb = torch.cat(rgbs, 0) if rgbs else torch.zeros((0, 3), dtype=torch.float64)
    return rgb.numpy(), gd.numpy(), gs.numpy()


# ---------------------------------------------------------------------------
# Literal restatement of _volume_render_gradcheck_lerp (svox2.py:658-884)
# ---------------------------------------------------------------------------
def volume_render_gradcheck_lerp(links: torch.Tensor, density: torch.Tensor, sh: torch.Tensor,
                                 basis_dim: int, radius, center, origins: torch.Tensor,
                                 dirs: torch.Tensor, step_size: float = 0.5,
                                 near_clip: float = 0.0, background_brightness: float = 1.0,
                                 dtype=torch.float32):
    """The reference's pure-PyTorch renderer, restated with the grid passed in
    and every tensor in ``dtype`` (the reference hard-codes float32 in
    _fetch_links, svox2.py:562-569, and its accumulators, :695-696).
    density (R, 1), sh (R, C), links (X, Y, Z) long/int."""
    radius = torch.as_tensor(radius, dtype=dtype).expand(3)
    center = torch.as_tensor(center, dtype=dtype).expand(3)
    gsz = torch.tensor(links.shape, dtype=dtype)
    offset = 0.5 * (1.0 - center / radius)
    scaling = 0.5 / radius
    # world2grid (svox2.py:1496-1509)
    origins = torch.addcmul(offset * gsz - 0.5, origins.to(dtype), scaling * gsz)
    dirs = dirs.to(dtype)
    dirs = dirs / torch.norm(dirs, dim=-1, keepdim=True)
    viewdirs = dirs
    B = dirs.size(0)
    dirs = dirs * (scaling * gsz)
    delta_scale = 1.0 / dirs.norm(dim=1)
    dirs = dirs * delta_scale.unsqueeze(-1)
    sh_mult = eval_sh_bases(basis_dim, viewdirs)
    invdirs = 1.0 / dirs
    t1 = (-0.5 - origins) * invdirs
    t2 = (gsz - 0.5 - origins) * invdirs
    t = torch.min(t1, t2)
    t[dirs == 0] = -1e9
    t = torch.max(t, dim=-1).values.clamp_min_(near_clip)
    tmax = torch.max(t1, t2)
    tmax[dirs == 0] = 1e9
    tmax = torch.min(tmax, dim=-1).values

    log_light = torch.zeros(B, dtype=dtype)
    out_rgb = torch.zeros((B, 3), dtype=dtype)
    good = torch.arange(B)
    mask = t <= tmax
    good, origins, dirs, t, sh_mult, tmax = (good[mask], origins[mask], dirs[mask], t[mask],
                                             sh_mult[mask], tmax[mask])
    C = sh.size(1)

    def fetch(lk):
        m = lk >= 0
        s = torch.zeros((lk.size(0), 1), dtype=dtype)
        c = torch.zeros((lk.size(0), C), dtype=dtype)
        idx = lk[m].long()
        s[m] = density[idx].to(dtype)
        c[m] = sh[idx].to(dtype)
        return s, c

    while good.numel() > 0:
        pos = origins + t[:, None] * dirs
        pos = pos.clamp_min(0.0)
        pos = torch.stack([torch.clamp_max(pos[:, i], gsz[i] - 1) for i in range(3)], -1)
        l = pos.to(torch.long).clamp_min(0)
        l = torch.stack([torch.clamp_max(l[:, i], int(gsz[i]) - 2) for i in range(3)], -1)
        pos = pos - l
        lx, ly, lz = l.unbind(-1)
        vals = [fetch(links[lx + dx, ly + dy, lz + dz]) for dx, dy, dz in _CORNERS]
        wa, wb = 1.0 - pos, pos

        def tri(i):
            c00 = vals[0][i] * wa[:, 2:] + vals[1][i] * wb[:, 2:]
            c01 = vals[2][i] * wa[:, 2:] + vals[3][i] * wb[:, 2:]
            c10 = vals[4][i] * wa[:, 2:] + vals[5][i] * wb[:, 2:]
            c11 = vals[6][i] * wa[:, 2:] + vals[7][i] * wb[:, 2:]
            c0 = c00 * wa[:, 1:2] + c01 * wb[:, 1:2]
            c1 = c10 * wa[:, 1:2] + c11 * wb[:, 1:2]
            return c0 * wa[:, :1] + c1 * wb[:, :1]

        sigma = tri(0)
        rgb = tri(1)
        log_att = -step_size * torch.relu(sigma[..., 0]) * delta_scale[good]
        weight = torch.exp(log_light[good]) * (1.0 - torch.exp(log_att))
        rgb = torch.clamp_min((sh_mult.unsqueeze(-2) * rgb.reshape(-1, 3, basis_dim)).sum(-1)
                              + 0.5, 0.0)
        out_rgb = out_rgb.index_add(0, good, weight[:, None] * rgb)
        log_light = log_light.index_add(0, good, log_att)
        t = t + step_size
        mask = t <= tmax
        good, origins, dirs, t, sh_mult, tmax = (good[mask], origins[mask], dirs[mask], t[mask],
                                                 sh_mult[mask], tmax[mask])
    if background_brightness:
        out_rgb = out_rgb + torch.exp(log_light).unsqueeze(-1) * background_brightness
    return out_rgb


# ---------------------------------------------------------------------------
# Sampling, TV, RMSprop, accel_dist_prop
# ---------------------------------------------------------------------------
def sample_trilerp(links: np.ndarray, density: torch.Tensor, sh: torch.Tensor, pts_grid: np.ndarray):
    """Trilinear samples at grid coordinates (svox2.py:604-653): clamp to
    [0, size-1], lower corner min(trunc, size-2), empty corners are 0.
    Returns (density (P,1), sh (P,C)) in the tensors' dtype (differentiable)."""
    size = np.array(links.shape)
    p = np.clip(pts_grid.astype(np.float64), 0.0, size - 1.0)
    l = np.minimum(p.astype(np.int64), size - 2)
    frac = torch.from_numpy(p - l).to(density.dtype)
    lk = torch.from_numpy(corner_gather(links, l).astype(np.int64))
    valid = (lk >= 0) & (lk < density.shape[0])
    lk0 = torch.where(valid, lk, torch.zeros_like(lk))
    w = corner_weights(frac) * valid.to(density.dtype)
    d = (w * density.reshape(-1)[lk0]).sum(1, keepdim=True)
    s = (w[..., None] * sh[lk0]).sum(1)
    return d, s


def tv_grad_sparse(links: np.ndarray, data: np.ndarray, rand_cells: np.ndarray, start_dim: int,
                   end_dim: int, scale: float, ignore_edge: bool, ignore_last_z: bool):
    """Gradient and touched-row mask of the sparse TV term
    (loss_kernel.cu:178-246 with the launcher's scale / n_cells, :597),
    including its edge rules, in float64 over all listed cells at once: a
    neighbour past the grid's end is link 0; a missing neighbour value is 0,
    or v000 with ignore_edge, which also skips cells whose link is 0;
    ignore_last_z skips cells at z = size2 - 2; a term is added (and its row
    marked) where its link is >= 0 and its scaled difference is nonzero."""
    data = np.asarray(data, np.float64)
    R, D = data.shape
    grad = np.zeros_like(data)
    mask = np.zeros(R, bool)
    s0, s1, s2 = links.shape
    cells = np.asarray(rand_cells, np.int64)
    n = cells.size
    if n == 0:
        return grad, mask
    scale_c = scale / n
    z = cells % s2
    y = (cells // s2) % s1
    x = cells // (s1 * s2)
    flat = links.reshape(-1).astype(np.int64)
    l000 = flat[cells]
    keep = np.ones(n, bool)
    if ignore_edge:
        keep &= l000 != 0
    if ignore_last_z:
        keep &= z != s2 - 2
    x, y, z, l000 = x[keep], y[keep], z[keep], l000[keep]
    base = (x * s1 + y) * s2 + z
    l001 = np.where(z + 1 < s2, flat[np.minimum(base + 1, flat.size - 1)], 0)
    l010 = np.where(y + 1 < s1, flat[np.minimum(base + s2, flat.size - 1)], 0)
    l100 = np.where(x + 1 < s0, flat[np.minimum(base + s1 * s2, flat.size - 1)], 0)
    ch = np.arange(start_dim, end_dim)

    def vals(lk, null):
        v = data[np.maximum(lk, 0)][:, ch]
        return np.where((lk >= 0)[:, None], v, null)

    v000 = vals(l000, 0.0)
    nul = v000 if ignore_edge else 0.0
    v001, v010, v100 = vals(l001, nul), vals(l010, nul), vals(l100, nul)
    dx, dy, dz = v100 - v000, v010 - v000, v001 - v000
    idelta = scale_c / np.sqrt(1e-9 + dx * dx + dy * dy + dz * dz)
    dx, dy, dz = dx * (s0 / 256.0), dy * (s1 / 256.0), dz * (s2 / 256.0)
    for lk, val in ((l000, -(dx + dy + dz)), (l001, dz), (l010, dy), (l100, dx)):
        act = (lk >= 0)[:, None] & (val != 0.0)
        rows = np.broadcast_to(lk[:, None], act.shape)[act]
        cols = np.broadcast_to(ch[None, :], act.shape)[act]
        np.add.at(grad, (rows, cols), (val * idelta)[act])
        mask[rows] = True
    return grad, mask


def _svox1_terms(grid: Grid, dirs, smp: dict, density: torch.Tensor, sh: torch.Tensor):
    """Per-sample FP64 density (plus 1e9 for an opaque last sample) and
    colour + 0.5 before clamping (S, 3), and the SH basis per sample."""
    B = grid.basis_dim
    link = torch.from_numpy(smp["link"].astype(np.int64))
    ray = torch.from_numpy(smp["ray"].astype(np.int64))
    basis = eval_sh_bases(B, torch.from_numpy(np.asarray(dirs, np.float32)).to(torch.float64))[ray]
    sigma = density[link] + torch.from_numpy((smp["flag"] & 8 != 0) * 1e9)
    col = (sh[link].view(-1, 3, B) * basis[:, None, :]).sum(-1) + 0.5
    return sigma, col, basis


def _segments(ray: np.ndarray, n: int):
    """(ray (S,) tensor, position within its ray, padded (n, L) shape) for
    per-ray prefix sums over samples grouped by ray."""
    r = torch.from_numpy(ray.astype(np.int64))
    S = r.numel()
    first = torch.ones(S, dtype=torch.bool)
    first[1:] = r[1:] != r[:-1]
    start = torch.nonzero(first).flatten()
    seg = torch.cumsum(first.long(), 0) - 1
    pos = torch.arange(S) - start[seg]
    return r, r, pos, (n, int(pos.max()) + 1 if S else 1)


def _ray_prefix_excl(v: torch.Tensor, r: torch.Tensor, pos: torch.Tensor, shape) -> torch.Tensor:
    """Exclusive prefix sum of v (S,) within each ray, by shifting the
    inclusive sums (no subtraction, which would lose the small terms next to
    an opaque sample's 1e9)."""
    pad = torch.zeros((shape[0], shape[1] + 1), dtype=v.dtype).index_put((r, pos + 1), v)
    return torch.cumsum(pad, 1)[r, pos]


def svox1_fp64(grid: Grid, origins, dirs, opt: Opts, density: torch.Tensor, sh: torch.Tensor,
               samples: Optional[dict] = None) -> torch.Tensor:
    """FP64 colours (n, 3) of the svox1 backend (trace_ray,
    render_svox1_kernel.cu:50-160) over the FP32 walk's samples
    (ref_svox1_samples), differentiable in density and sh: counted samples
    composite with weight T (1 - exp(-dt sigma)), colour max(sh . basis +
    0.5, 0); a ray whose transmittance before a counted sample is <=
    stop_thresh stops after that sample and its colour is divided by
    1 - that transmittance (no background); other rays add T_end bg; rays that
    miss the box get bg."""
    n = np.asarray(origins).shape[0]
    smp = march().svox1_samples(grid, origins, dirs, opt) if samples is None else samples
    bg = float(opt.background_brightness)
    out = torch.full((n, 3), bg, dtype=torch.float64) + 0 * (density.sum() + sh.sum())
    if smp["ray"].size == 0:
        return out
    sigma, col, _ = _svox1_terms(grid, dirs, smp, density, sh)
    counted = torch.from_numpy((smp["flag"] & 1) != 0)
    dt = torch.from_numpy(smp["dt"].astype(np.float64))
    pcnt = torch.where(counted, dt * sigma, torch.zeros_like(sigma))
    r, _, pos, shape = _segments(smp["ray"], n)
    t_before = torch.exp(-_ray_prefix_excl(pcnt, r, pos, shape))
    w = t_before * (1.0 - torch.exp(-pcnt)) * counted.to(torch.float64)
    rgb = torch.zeros((n, 3), dtype=torch.float64).index_add(0, r, w[:, None] * torch.clamp_min(col, 0.0))
    depth = torch.zeros(n, dtype=torch.float64).index_add(0, r, pcnt)
    hit = torch.zeros(n, dtype=torch.bool)
    hit[r] = True
    stop = torch.from_numpy((smp["flag"] & 2) != 0)
    stopped = torch.zeros(n, dtype=torch.bool)
    stopped[r[stop]] = True
    scale = torch.ones(n, dtype=torch.float64)
    scale = scale.index_put((r[stop],), 1.0 / (1.0 - t_before[stop]))
    lit = torch.where(stopped[:, None], rgb * scale[:, None],
                      rgb + (torch.exp(-depth) * bg)[:, None])
    return torch.where(hit[:, None], lit, out)


def svox1_grads_fp64(grid: Grid, origins, dirs, opt: Opts, grad_out, color_cache,
                     samples: Optional[dict] = None):
    """(grad_density (R,), grad_sh (R, C), mask (R,)) of the svox1 backward
    (trace_ray_backward, render_svox1_kernel.cu:162-275) restated in FP64
    over the FP32 walk's samples: over the counted samples up to the one
    after which the transmittance is <= stop_thresh, SH coefficient i of
    channel k gets weight g_k basis_i where colour_k > 0; the density gets
    dt (T_after sum_k [colour_k > 0] colour_k g_k - accum_after), accum
    starting at color_cache . g and losing each sample's weight times that
    sum; every counted sample marks its row.  Rays that miss the box get
    nothing."""
    n = np.asarray(origins).shape[0]
    R, C = grid.sh.shape
    B = grid.basis_dim
    gd = np.zeros(R)
    gs = np.zeros((R, C))
    mask = np.zeros(R, bool)
    smp = march().svox1_samples(grid, origins, dirs, opt) if samples is None else samples
    if smp["ray"].size == 0:
        return gd, gs, mask
    density = torch.from_numpy(np.asarray(grid.density, np.float64).reshape(-1))
    sh = torch.from_numpy(np.asarray(grid.sh, np.float64))
    sigma, col, basis = (x.numpy() for x in _svox1_terms(grid, dirs, smp, density, sh))
    g = np.asarray(grad_out, np.float64)
    cc = np.asarray(color_cache, np.float64)
    ray, link, flag = smp["ray"], smp["link"].astype(np.int64), smp["flag"]
    dt = smp["dt"].astype(np.float64)
    bounds = np.concatenate([[0], np.cumsum(np.bincount(ray, minlength=n))])
    for r in range(n):
        accum = float(cc[r] @ g[r])
        log_t = 0.0
        for i in range(bounds[r], bounds[r + 1]):
            if not flag[i] & 1:
                continue
            lt = -dt[i] * sigma[i]
            w = math.exp(log_t) * (1.0 - math.exp(lt))
            pos = col[i] > 0.0
            total = float((col[i] * g[r] * pos).sum())
            for k in range(3):
                if pos[k]:
                    gs[link[i], k * B:(k + 1) * B] += basis[i] * (w * g[r][k])
            log_t += lt
            accum -= w * total
            mask[link[i]] = True
            gd[link[i]] += dt[i] * (total * math.exp(log_t) - accum)
            if flag[i] & 4:
                break
    return gd, gs, mask


@dataclass
class Background:
    """Host copy of an MSI background: links (2 reso, reso) int32 (-1 =
    empty) into rows of data (rows, nlayers, 4) float64, channels (r, g, b,
    sigma) as the GPU stores them (float32 values)."""
    links: np.ndarray
    data: np.ndarray

    @property
    def reso(self):
        return self.links.shape[1]

    @property
    def nlayers(self):
        return self.data.shape[1]


def background_fp64(grid: Grid, bgm: Background, origins, dirs, opt: Opts,
                    data: torch.Tensor, log_t: torch.Tensor, sparsity_energy: bool = False):
    """FP64 colour the MSI background adds to each ray (n, 3) behind the
    grid's log transmittance ``log_t`` (n,), restating render_background_forward
    (render_lerp_kernel_cuvol.cu:386-476; kernel :775-796), differentiable in
    ``data`` (rows, nlayers, 4) and ``log_t``:

    * the ray in the frame where the grid spans [-1, 1] (ray_find_bounds_bg,
      render_util.cuh:630-657): o_s = (o_grid + 0.5) 2 / size - 1, d_s =
      normalised d_grid 2 / size, ws = the world length per unit of d_s;
    * first radius max(|o_s x d_s| + 1e-3, 1); n_steps = int(nlayers /
      step_size) + 2 (in float32); sample i at radius n_steps / (n_steps - i -
      0.5), on the sphere's far intersection (ConcentricSpheresIntersector,
      :546-576), skipped below the first radius or without intersection;
    * pixel (2 reso (0.5 + lon / 2 pi), reso (0.5 - lat / pi)), lat = asin(u_y),
      lon = atan2(u_x, u_z) of the unit position u; layer clamp((1 - 1 / r)
      nlayers - 0.5, 0, nlayers - 1); trilinear over columns (x, y), (x, y + 1),
      (x + 1, y), (x + 1, y + 1) wrapping at the image edges (trilerp_bg_one,
      :157-189), empty columns reading 0;
    * a sample with sigma > 0 has optical depth (1 / r_prev - 1 / r) ws sigma,
      r_prev the radius of the previous placed sample; weight T (1 - e^-depth),
      colour max(C0 rgb + 0.5, 0); the ray stops after the sample where T falls
      below stop_thresh; the remaining T times background_brightness is added;
    * rays with log_t < -25 get nothing.
    With ``sparsity_energy`` also sum over contributing samples of
    log(1 + 2 sigma^2), whose gradient is the Cauchy sparsity term the fused
    backward adds (:584-588)."""
    o = np.asarray(origins, np.float64)
    d = np.asarray(dirs, np.float64)
    n = o.shape[0]
    size = np.asarray(grid.size, np.float64)
    og = o * grid.scaling + grid.offset
    dg = d * grid.scaling
    ws = 1.0 / np.linalg.norm(dg, axis=1)
    dg = dg * ws[:, None]
    sph = 2.0 / size
    os_ = (og + 0.5) * sph - 1.0
    ds = dg * sph
    inorm = 1.0 / np.linalg.norm(ds, axis=1)
    ws = ws * inorm
    ds = ds * inorm[:, None]
    q2a = 2.0 * (ds * ds).sum(1)
    qb = 2.0 * (os_ * ds).sum(1)
    f = qb * qb - 2.0 * q2a * (os_ * os_).sum(1)
    inner = np.maximum(np.linalg.norm(np.cross(os_, ds), axis=1) + 1e-3, 1.0)
    reso, nl = bgm.reso, bgm.nlayers
    links = np.asarray(bgm.links, np.int64)
    rows = data.shape[0]
    n_steps = int(np.float32(nl) / np.float32(opt.step_size)) + 2
    active = (log_t.detach().numpy() >= -25.0)
    stopped = np.zeros(n, bool)
    invr_last = 1.0 / inner
    lt = log_t
    out = torch.zeros((n, 3), dtype=torch.float64)
    energy = torch.zeros((), dtype=torch.float64)
    ws_t = torch.from_numpy(ws)
    for i in range(n_steps):
        rad = n_steps / (n_steps - i - 0.5)
        det = f + 2.0 * q2a * rad * rad
        valid = active & ~stopped & (rad >= inner) & (det >= 0.0)
        t = (-qb + np.sqrt(np.maximum(det, 0.0))) / q2a
        p = os_ + t[:, None] * ds
        invr = 1.0 / np.linalg.norm(p, axis=1)
        u = p * invr[:, None]
        lat = np.arcsin(np.clip(u[:, 1], -1.0, 1.0))
        lon = np.arctan2(u[:, 0], u[:, 2])
        x = 2 * reso * (0.5 + lon * 0.5 / np.pi)
        y = reso * (0.5 - lat / np.pi)
        z = np.clip((1.0 - invr) * nl - 0.5, 0.0, nl - 1.0)
        l0 = np.clip(np.trunc(x).astype(np.int64), 0, 2 * reso - 1)
        l1 = np.clip(np.trunc(y).astype(np.int64), 0, reso - 1)
        l2 = np.clip(np.trunc(z).astype(np.int64), 0, nl - 2)
        fx, fy, fz = (torch.from_numpy(v) for v in (x - l0, y - l1, z - l2))
        nx = np.where(l0 < 2 * reso - 1, l0 + 1, 0)
        ny = np.where(l1 < reso - 1, l1 + 1, 0)

        def column(cx, cy):
            lk = links[cx, cy]
            ok = (lk >= 0) & (lk < rows)
            r = torch.from_numpy(np.where(ok, lk, 0))
            lz = torch.from_numpy(l2)
            v0, v1 = data[r, lz], data[r, lz + 1]
            v = v0 + fz[:, None] * (v1 - v0)
            return v * torch.from_numpy(ok)[:, None].to(torch.float64)

        c0 = column(l0, l1) + fy[:, None] * (column(l0, ny) - column(l0, l1))
        c1 = column(nx, l1) + fy[:, None] * (column(nx, ny) - column(nx, l1))
        v = c0 + fx[:, None] * (c1 - c0)
        sigma = v[:, 3]
        contrib = torch.from_numpy(valid) & (sigma > 0).detach()
        step = torch.from_numpy(invr_last - invr) * ws_t
        pcnt = torch.where(contrib, step * sigma, torch.zeros_like(sigma))
        w = torch.exp(lt) * (1.0 - torch.exp(-pcnt))
        col = torch.clamp_min(v[:, :3] * SH_C0 + 0.5, 0.0)
        out = out + (w * contrib.to(torch.float64))[:, None] * col
        if sparsity_energy:
            energy = energy + torch.where(contrib, torch.log(1.0 + 2.0 * sigma * sigma),
                                          torch.zeros_like(sigma)).sum()
        lt = lt - pcnt
        stopped |= (contrib & (torch.exp(lt) < opt.stop_thresh)).numpy()
        invr_last = np.where(valid, invr, invr_last)
    out = out + (torch.exp(lt) * float(opt.background_brightness)
                 * torch.from_numpy(active).to(torch.float64))[:, None]
    return (out, energy) if sparsity_energy else out


def render_bg_fp64(grid: Grid, bgm: Background, origins, dirs, opt: Opts,
                   density: torch.Tensor, sh: torch.Tensor, data: torch.Tensor, clip=None,
                   sparsity_energy: bool = False):
    """FP64 colours (n, 3) of a grid with an MSI background: the grid's
    compositing with background brightness 0 (render_lerp_kernel_cuvol.cu:43,
    111-113) plus ``background_fp64`` behind its final log transmittance;
    differentiable in density, sh and the background data."""
    import dataclasses
    fg, log_t = render_fp64(grid, origins, dirs, dataclasses.replace(opt, background_brightness=0.0),
                            density, sh, clip=clip, return_log_t=True)
    bg = background_fp64(grid, bgm, origins, dirs, opt, data, log_t, sparsity_energy)
    if sparsity_energy:
        return fg + bg[0], bg[1]
    return fg + bg


def msi_tv_grad_sparse(links: np.ndarray, msi: np.ndarray, rand_cells: np.ndarray,
                       scale: float, scale_last: float):
    """(grad (rows, nlayers, C), mask (rows, nlayers)) of msi_tv_grad_sparse
    (loss_kernel.cu:248-333,606-641) in FP64: for each listed flat index
    (x reso + y) nlayers + z and channel c, with v00 = the cell's value (0 for
    an empty column), its +x and +y neighbours (wrapping) and layer z + 1
    (missing: v00 for colours, 0 for the last channel): dx, dy, dz = the
    differences, idelta = (scale or scale_last for the last channel) / n_cells
    / sqrt(1e-9 + dx^2 + dy^2 + dz^2) of the unscaled differences, then
    dx, dy, dz scaled by 2 reso / 256, reso / 256, nlayers / 256, and
    -(dx + dy + dz), dz, dy, dx times idelta added to the cell, its layer
    z + 1, its +y and its +x neighbours (nonempty columns, nonzero terms,
    each marking its layer)."""
    links = np.asarray(links, np.int64)
    msi = np.asarray(msi, np.float64)
    rows, nl, C = msi.shape
    reso = links.shape[1]
    grad = np.zeros_like(msi)
    mask = np.zeros((rows, nl), bool)
    cells = np.asarray(rand_cells, np.int64)
    n = cells.size
    for idx in cells:
        z = idx % nl
        tmp = idx // nl
        y = tmp % reso
        x = tmp // reso
        nx = 0 if x == 2 * reso - 1 else x + 1
        ny = 0 if y == reso - 1 else y + 1
        l00, l01, l10 = links[x, y], links[x, ny], links[nx, y]
        for c in range(C):
            last = c == C - 1
            v00 = msi[l00, z, c] if l00 >= 0 else 0.0
            if l00 >= 0 and z + 1 < nl:
                vz = msi[l00, z + 1, c]
            else:
                vz = 0.0 if last else v00
            v01 = msi[l01, z, c] if l01 >= 0 else 0.0
            v10 = msi[l10, z, c] if l10 >= 0 else 0.0
            dx, dy, dz = v10 - v00, v01 - v00, vz - v00
            idelta = (scale_last if last else scale) / n / math.sqrt(1e-9 + dx * dx + dy * dy
                                                                     + dz * dz)
            dx *= 2 * reso / 256.0
            dy *= reso / 256.0
            dz *= nl / 256.0
            for lk, zz, val in ((l00, z, -(dx + dy + dz)), (l00, z + 1, dz), (l01, z, dy),
                                (l10, z, dx)):
                if lk < 0 or zz >= nl or val == 0.0:
                    continue
                grad[lk, zz, c] += val * idelta
                mask[lk, zz] = True
    return grad, mask


def lumisphere_tv_grad_sparse(links: np.ndarray, data: np.ndarray, rand_cells: np.ndarray,
                              basis_dim: int, sv, svu, scale: float, dir_factor: float):
    """Gradient and touched-row mask of the lumisphere TV term
    (loss_kernel.cu:336-446 with the launcher's scale / n_cells, :671), in
    float64 over all listed cells at once.  The flat index splits by the
    extents size - 1 (:360-365); a cell whose link is 0 is skipped (:378); an
    index whose x part is >= size0 - 1 is skipped too (the reference reads
    past the links array there); a missing v000 reads 0 and a missing
    neighbour v000 (:385-392); only the first 32 coefficients of a row take
    part, one per lane of the reference's warp (:351), and a channel's colour
    sums the coefficients it has among them (:402-411); a term is added (and
    its row marked) where its link is >= 0 and its value is nonzero
    (:433-445)."""
    data = np.asarray(data, np.float64)
    R, D = data.shape
    grad = np.zeros_like(data)
    mask = np.zeros(R, bool)
    s0, s1, s2 = links.shape
    cells = np.asarray(rand_cells, np.int64)
    n = cells.size
    if n == 0 or min(s0, s1, s2) < 2:
        return grad, mask
    scale_c = scale / n
    z = cells % (s2 - 1)
    xy = cells // (s2 - 1)
    y, x = xy % (s1 - 1), xy // (s1 - 1)
    keep = (cells >= 0) & (x < s0 - 1)
    base = ((x * s1 + y) * s2 + z)[keep]
    flat = links.reshape(-1).astype(np.int64)
    l000 = flat[base]
    base, l000 = base[l000 != 0], l000[l000 != 0]
    l001, l010, l100 = flat[base + 1], flat[base + s2], flat[base + s1 * s2]
    dim = min(D, 32)
    ent = np.arange(dim)
    ch = ent // basis_dim
    sv_e = np.asarray(sv, np.float64)[:basis_dim][ent % basis_dim]
    svu_e = np.asarray(svu, np.float64)[:basis_dim][ent % basis_dim]
    onehot = (ch[:, None] == np.arange(ch.max() + 1)[None, :]).astype(np.float64)

    def vals(lk, null):
        v = data[np.maximum(lk, 0)][:, :dim]
        return np.where((lk >= 0)[:, None], v, null)

    def colour(v, w):  # per-channel sums, broadcast back to the coefficients
        return ((v * w) @ onehot)[:, ch]

    v000 = vals(l000, 0.0)
    v001, v010, v100 = vals(l001, v000), vals(l010, v000), vals(l100, v000)
    a000 = colour(v000, sv_e)
    f0, f1, f2 = s0 / 256.0, s1 / 256.0, s2 / 256.0
    dx = (colour(v100, sv_e) - a000) * f0
    dy = (colour(v010, sv_e) - a000) * f1
    dz = (colour(v001, sv_e) - a000) * f2
    du = (colour(v000, svu_e) - a000) * dir_factor
    idelta = scale_c / np.sqrt(1e-9 + dx * dx + dy * dy + dz * dz + du * du)
    dx, dy, dz, du = dx * f0, dy * f1, dz * f2, du * dir_factor
    sm = -dx * sv_e - dy * sv_e - dz * sv_e + du * (svu_e - sv_e)
    for lk, val in ((l000, sm), (l001, dz * sv_e), (l010, dy * sv_e), (l100, dx * sv_e)):
        act = (lk >= 0)[:, None] & (val != 0.0)
        rows = np.broadcast_to(lk[:, None], act.shape)[act]
        cols = np.broadcast_to(ent[None, :], act.shape)[act]
        np.add.at(grad, (rows, cols), (val * idelta)[act])
        mask[rows] = True
    return grad, mask


def rmsprop_step(data, rms, grad, rows_sel, beta, lr, eps, minval, lr_last):
    """Closed form per element (optim_kernel.cu:15-25) on the selected rows;
    returns new (data, rms, grad) as float64 arrays."""
    data, rms, grad = (np.array(a, np.float64) for a in (data, rms, grad))
    D = data.shape[1]
    lrs = np.full(D, lr, np.float64)
    lrs[-1] = lr_last
    for r in rows_sel:
        g = grad[r]
        g2 = g * g
        rms[r] = np.where(rms[r] == 0.0, g2, beta * rms[r] + (1 - beta) * g2)
        data[r] = np.maximum(data[r] - lrs * g / (np.sqrt(rms[r]) + eps), minval)
        grad[r] = 0.0
    return data, rms, grad


def sgd_step(data, grad, rows_sel, lr, lr_last):
    """Closed form per element (optim_kernel.cu:97-103, lr_last on the last
    column as :111-115): data - lr * g on the selected rows, g zeroed there;
    returns new (data, grad) as float64 arrays."""
    data, grad = (np.array(a, np.float64) for a in (data, grad))
    lrs = np.full(data.shape[1], lr, np.float64)
    lrs[-1] = lr_last
    rows_sel = np.asarray(rows_sel, np.int64)
    data[rows_sel] -= lrs * grad[rows_sel]
    grad[rows_sel] = 0.0
    return data, grad


def dilate(mask: np.ndarray) -> np.ndarray:
    """One 3x3x3 dilation of a boolean volume, neighbours clipped to the grid
    (dilate_kernel, misc_kernel.cu:20-48): the OR of the 27 shifted copies of
    the volume padded with False."""
    m = np.asarray(mask, bool)
    pad = np.pad(m, 1, constant_values=False)
    out = np.zeros_like(m)
    X, Y, Z = m.shape
    for dx in range(3):
        for dy in range(3):
            for dz in range(3):
                out |= pad[dx:dx + X, dy:dy + Y, dz:dz + Z]
    return out


def mask_to_links(mask: np.ndarray, z_order: bool = False):
    """(links, kept cells) of a boolean volume (svox2.py:1397-1412): the k-th
    kept cell in raster order, or in Morton order with z_order, gets link k;
    the others -1; kept[k] is its flat raster index."""
    m = np.asarray(mask, bool)
    flat = m.reshape(-1)
    links = np.full(flat.size, -1, np.int32)
    if z_order:
        code = gen_morton(m.shape[0]).reshape(-1)  # Morton code of each raster cell
        order = np.argsort(code)                   # raster cells in Morton order
        kept = order[flat[order]]
    else:
        kept = np.nonzero(flat)[0]
    links[kept] = np.arange(kept.size, dtype=np.int32)
    return links.reshape(m.shape), kept.astype(np.int64)


def resample_axes(old_size, new_size) -> np.ndarray:
    """The new cells' coordinates along each axis in old grid coordinates,
    as svox2.py:1275-1294 computes them (float32 linspace per axis),
    concatenated x | y | z."""
    axes = []
    for a in range(3):
        f = 0.5 * old_size[a] / new_size[a]
        axes.append(torch.linspace(f - 0.5, old_size[a] - f - 0.5, new_size[a],
                                   dtype=torch.float32).numpy())
    return np.concatenate(axes)


def resample_points(axes: np.ndarray, new_size, cells: np.ndarray) -> np.ndarray:
    """Grid-coordinate points (n, 3) of flat new-grid cells."""
    X, Y, Z = new_size
    ax, ay, az = axes[:X], axes[X:X + Y], axes[X + Y:]
    cells = np.asarray(cells, np.int64)
    return np.stack([ax[cells // (Y * Z)], ay[(cells // Z) % Y], az[cells % Z]], 1)


def accel_dist_prop(links: np.ndarray) -> np.ndarray:
    """Empty links rewritten to -1 - k, k = consecutive empty ancestor levels of
    the occupancy pyramid (misc_kernel.cu:109-181,356-416), by explicit
    per-level boolean pyramids."""
    links = np.array(links, np.int32)
    occ = links >= 0
    levels = []
    cur = occ
    while all(s > 1 for s in cur.shape):
        nshape = tuple((s + 1) // 2 for s in cur.shape)
        nxt = np.zeros(nshape, bool)
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    sub = cur[dx::2, dy::2, dz::2]
                    nxt[:sub.shape[0], :sub.shape[1], :sub.shape[2]] |= sub
        levels.append(nxt)
        cur = nxt
    out = links.copy()
    X, Y, Z = np.nonzero(~occ)
    res = np.full(X.shape, -1, np.int32)
    done = np.zeros(X.shape, bool)
    x, y, z = X.copy(), Y.copy(), Z.copy()
    for lev in levels:
        x, y, z = x >> 1, y >> 1, z >> 1
        hit = lev[x, y, z] & ~done
        done |= hit
        res = np.where(done, res, res - 1)
    out[X, Y, Z] = res
    return out
