"""Seeded synthetic grids, cameras and ray batches for the benchmark and the
tests (no datasets are available offline).

Grids are sums of isotropic Gaussian "blobs" over the cube [-1, 1]^3 sampled at
voxel centres x_i = -1 + (2i + 1) / N (the grid2world map of
svox2/svox2.py:1511-1524 with radius 1, centre 0).  Sparse grids keep the cells
whose field exceeds a threshold, with rows assigned in Morton (Z-order) order
of their cells, and density = the field value.  SH coefficients are
0.1 * N(0, 1) with the DC terms 0.5 + 0.2 * N(0, 1).  All draws come from one
``torch.Generator('cpu')`` in a fixed order, in float64.

Workloads (sizes and parameters):
  dense64      64^3 dense, 8 blobs (std 0.15, amplitude 50), 4096 rays from a
               radius-1.5 sphere aimed into a radius-0.3 ball.
  sparse256    256^3, 8 blobs (std 0.1, amplitude 50) > 1: about 8% of cells.
  sparse512    512^3, 32 blobs (std U[0.0165, 0.066], amplitude U[20, 100]) > 1:
               about 3% of cells.
Camera rays come from pinhole cameras (800x800, focal 1111 px, principal point
at the image centre) on the upper (z > 0) hemisphere of radius 4 looking at the
origin, OpenCV convention; a training batch draws pixels uniformly over all
cameras' pixels (the ray sampling of opt/util/util.py:178-190).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np
import torch

from . import utils


@dataclass
class BlobSpec:
    n_blobs: int
    std_range: Tuple[float, float]
    amp_range: Tuple[float, float]
    center_range: float = 0.6
    threshold: float = 1.0


DENSE64 = BlobSpec(8, (0.15, 0.15), (50.0, 50.0))
SPARSE256 = BlobSpec(8, (0.1, 0.1), (50.0, 50.0))
SPARSE512 = BlobSpec(32, (0.0165, 0.066), (20.0, 100.0))


def draw_blobs(spec: BlobSpec, gen: torch.Generator):
    """Centres U[-c, c]^3, then stds, then amplitudes (float64)."""
    c = spec.center_range
    centers = torch.rand(spec.n_blobs, 3, generator=gen, dtype=torch.float64) * (2 * c) - c
    s0, s1 = spec.std_range
    stds = torch.rand(spec.n_blobs, generator=gen, dtype=torch.float64) * (s1 - s0) + s0
    a0, a1 = spec.amp_range
    amps = torch.rand(spec.n_blobs, generator=gen, dtype=torch.float64) * (a1 - a0) + a0
    return centers, stds, amps


def blob_field(N: int, centers, stds, amps, device="cpu") -> torch.Tensor:
    """(N, N, N) float64 field sum_k a_k exp(-|x - c_k|^2 / (2 s_k^2)) at voxel
    centres, evaluated separably (one 1-D factor per axis and blob) slab by slab."""
    dev = torch.device(device)
    xs = -1.0 + (2.0 * torch.arange(N, dtype=torch.float64, device=dev) + 1.0) / N
    c = centers.to(dev)
    s = stds.to(dev)
    a = amps.to(dev)
    inv = 1.0 / (2.0 * s * s)
    ex = torch.exp(-((xs[None, :] - c[:, 0:1]) ** 2) * inv[:, None])  # (K, N)
    ey = torch.exp(-((xs[None, :] - c[:, 1:2]) ** 2) * inv[:, None])
    ez = torch.exp(-((xs[None, :] - c[:, 2:3]) ** 2) * inv[:, None])
    out = torch.empty((N, N, N), dtype=torch.float64, device=dev)
    for x in range(N):
        w = a * ex[:, x]  # (K,)
        out[x] = (ey.T * w[None, :]) @ ez
    return out


def morton_rows(mask: torch.Tensor) -> Tuple[torch.Tensor, int]:
    """Links for occupied cells of ``mask`` (X, Y, Z), rows in Morton order of
    the cells; empty cells get -1."""
    X, Y, Z = mask.shape
    idx = torch.nonzero(mask.flatten(), as_tuple=False).flatten()
    z = idx % Z
    y = (idx // Z) % Y
    x = idx // (Y * Z)
    code = utils.morton_code_3(x, y, z)
    order = torch.argsort(code)
    links = torch.full((X * Y * Z,), -1, dtype=torch.int32, device=mask.device)
    rows = torch.empty_like(order)
    rows[order] = torch.arange(order.numel(), device=mask.device)
    links[idx] = rows.to(torch.int32)
    return links.view(X, Y, Z), int(idx.numel())


def random_sh(R: int, C: int, gen: torch.Generator, basis_dim: int) -> torch.Tensor:
    """0.1 * N(0,1) for every coefficient, DC terms 0.5 + 0.2 * N(0,1)."""
    sh = 0.1 * torch.randn(R, C, generator=gen, dtype=torch.float64)
    dc = 0.5 + 0.2 * torch.randn(R, 3, generator=gen, dtype=torch.float64)
    sh[:, [0, basis_dim, 2 * basis_dim]] = dc
    return sh


@dataclass
class SceneData:
    """A grid as plain tensors (host or device)."""
    links: torch.Tensor       # (N, N, N) int32
    density: torch.Tensor     # (R, 1) float64
    sh: torch.Tensor          # (R, C) float64
    basis_dim: int
    occupancy: float


def make_blob_grid(N: int, spec: BlobSpec, gen: torch.Generator, dense: bool = False,
                   basis_dim: int = 9, device="cpu") -> SceneData:
    """Blob grid; ``dense`` keeps every cell (row-major rows)."""
    centers, stds, amps = draw_blobs(spec, gen)
    field = blob_field(N, centers, stds, amps, device=device)
    if dense:
        links = torch.arange(N ** 3, dtype=torch.int32, device=device).view(N, N, N)
        dens = field.reshape(-1, 1)
        R = N ** 3
    else:
        mask = field > spec.threshold
        links, R = morton_rows(mask)
        order = torch.empty(R, dtype=torch.long, device=field.device)
        valid = links.flatten() >= 0
        order[links.flatten()[valid].long()] = torch.nonzero(valid).flatten()
        dens = field.flatten()[order].reshape(-1, 1)
    sh = random_sh(R, 3 * basis_dim, gen, basis_dim).to(device)
    return SceneData(links, dens, sh, basis_dim, R / float(N ** 3))


def to_sparse_grid(scene: SceneData, device="cuda", dtype=torch.float32, accelerate=True):
    """A svox2_amd.SparseGrid holding the scene (storage ``dtype``)."""
    from .svox2 import SparseGrid
    N = scene.links.shape
    grid = SparseGrid(reso=list(N), basis_dim=scene.basis_dim, device="cpu", dtype=dtype)
    grid.links = scene.links.to(device)
    grid._set_data(scene.density.to(device=device, dtype=torch.float64),
                   scene.sh.to(device=device, dtype=torch.float64), dtype, device)
    grid.capacity = grid.sh_data.size(0)
    grid.basis_data = torch.nn.Parameter(grid.basis_data.data.to(device), requires_grad=False)
    grid.background_data = torch.nn.Parameter(grid.background_data.data.to(device),
                                              requires_grad=False)
    if accelerate and grid.links.is_cuda:
        grid.accelerate()
    return grid


# ---------------------------------------------------------------------------
# Rays and cameras
# ---------------------------------------------------------------------------
def sphere_rays(n: int, gen: torch.Generator, radius: float = 1.5, target: float = 0.3):
    """Origins uniform on a sphere, each aimed at a uniform point of a ball
    about the centre; unit float32 directions."""
    v = torch.randn(n, 3, generator=gen, dtype=torch.float64)
    o = radius * v / v.norm(dim=1, keepdim=True)
    u = torch.randn(n, 3, generator=gen, dtype=torch.float64)
    U = torch.rand(n, 1, generator=gen, dtype=torch.float64)
    p = target * u / u.norm(dim=1, keepdim=True) * U.pow(1.0 / 3.0)
    d = p - o
    d = d / d.norm(dim=1, keepdim=True)
    return o.float().contiguous(), d.float().contiguous()


def look_at_c2w(center: torch.Tensor) -> torch.Tensor:
    """OpenCV camera-to-world (3, 4) at ``center`` looking at the origin, +z up."""
    f = -center / center.norm()
    up = torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64)
    x = torch.linalg.cross(f, up)
    if x.norm() < 1e-9:
        x = torch.tensor([1.0, 0.0, 0.0], dtype=torch.float64)
    x = x / x.norm()
    y = torch.linalg.cross(f, x)
    return torch.stack([x, y, f, center], dim=1)


def hemisphere_cameras(n: int, gen: torch.Generator, radius: float = 4.0) -> torch.Tensor:
    """(n, 3, 4) float64 c2w matrices, centres uniform on the upper hemisphere."""
    cos_t = torch.rand(n, generator=gen, dtype=torch.float64)
    phi = torch.rand(n, generator=gen, dtype=torch.float64) * (2 * math.pi)
    sin_t = torch.sqrt(1 - cos_t ** 2)
    c = radius * torch.stack([sin_t * torch.cos(phi), sin_t * torch.sin(phi), cos_t], 1)
    return torch.stack([look_at_c2w(c[i]) for i in range(n)])


@dataclass
class CameraSet:
    c2w: torch.Tensor  # (n, 3, 4) float64
    width: int = 800
    height: int = 800
    focal: float = 1111.0

    def pixel_rays(self, cam: torch.Tensor, iy: torch.Tensor, ix: torch.Tensor):
        """World rays of pixels (cam, iy, ix): pixel-centre directions
        ((ix + 0.5 - cx) / f, (iy + 0.5 - cy) / f, 1) normalised and rotated
        (opt/util/dataset_base.py:36-71), float32."""
        cx, cy = self.width * 0.5, self.height * 0.5
        x = (ix.double() + 0.5 - cx) / self.focal
        y = (iy.double() + 0.5 - cy) / self.focal
        d = torch.stack([x, y, torch.ones_like(x)], -1)
        d = d / d.norm(dim=-1, keepdim=True)
        R = self.c2w[cam, :, :3]
        dirs = (R @ d[..., None])[..., 0]
        origins = self.c2w[cam, :, 3]
        return origins.float().contiguous(), dirs.float().contiguous()

    def random_batch(self, n: int, gen: torch.Generator):
        """n rays drawn uniformly over every pixel of every camera."""
        total = self.c2w.shape[0] * self.width * self.height
        idx = torch.randint(0, total, (n,), generator=gen)
        cam = idx // (self.width * self.height)
        pix = idx % (self.width * self.height)
        return self.pixel_rays(cam, pix // self.width, pix % self.width)

    def camera(self, i: int, width: Optional[int] = None, height: Optional[int] = None):
        """svox2_amd.Camera for camera i (optionally another image size with
        the same field of view)."""
        from .svox2 import Camera
        w = width or self.width
        h = height or self.height
        f = self.focal * w / self.width
        return Camera(self.c2w[i].float(), fx=f, fy=f, width=w, height=h)


def config_clip(n: int, gen: torch.Generator):
    """Per-ray (t_near, t_far): t_near ~ U[0, 0.5], t_far ~ U[2, 4]."""
    near = torch.rand(n, generator=gen, dtype=torch.float64) * 0.5
    far = torch.rand(n, generator=gen, dtype=torch.float64) * 2.0 + 2.0
    return torch.stack([near, far], 1).float().contiguous()


# ---------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------
def workload_dense64(seed: int = 0, n_rays: int = 4096, device="cpu"):
    """64^3 dense grid, sphere rays and U(0,1) target colours (one generator,
    drawn in that order)."""
    gen = torch.Generator("cpu").manual_seed(seed)
    scene = make_blob_grid(64, DENSE64, gen, dense=True, device=device)
    origins, dirs = sphere_rays(n_rays, gen)
    gt = torch.rand(n_rays, 3, generator=gen, dtype=torch.float64).float()
    return scene, origins, dirs, gt


def workload_blobs(N: int, spec: BlobSpec, seed: int = 0, n_cameras: int = 100, device="cpu"):
    """Sparse blob grid at resolution N and the camera set; returns the
    generator positioned after these draws, for ray batches."""
    gen = torch.Generator("cpu").manual_seed(seed)
    scene = make_blob_grid(N, spec, gen, device=device)
    cams = CameraSet(hemisphere_cameras(n_cameras, gen))
    return scene, cams, gen
