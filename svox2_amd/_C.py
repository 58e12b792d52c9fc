"""Python face of the gfx950 kernel library, shaped like the reference's
pybind11 module ``svox2.csrc`` (svox2/csrc/svox2.cpp:74-172).

The spec classes (``SparseGridSpec``, ``RaysSpec``, ``CameraSpec``,
``RenderOptions``, ``GridOutputGrads``) carry the same fields as the reference's
C++ structs (svox2/csrc/include/data_spec.hpp:18-135), and the functions keep
the reference's names, argument order and in-place semantics.  Underneath, each
call validates its tensors the way ``CHECK_INPUT`` does (svox2/csrc/include/
util.hpp:3-12: device tensor, contiguous, expected dtype) and raises
``RuntimeError`` on violation, then calls the C ABI declared in
``include/svox2_amd.h`` through ctypes with raw device pointers and PyTorch's
current HIP stream.  There is no fallback: if the library or a GPU is missing,
calls raise.
"""
from __future__ import annotations

import ctypes
import os
from collections import OrderedDict
from typing import Optional, Tuple

import torch

__all__ = [
    "SparseGridSpec", "RaysSpec", "CameraSpec", "RenderOptions", "GridOutputGrads",
    "volume_render_cuvol", "volume_render_cuvol_image", "volume_render_cuvol_backward",
    "volume_render_cuvol_fused", "volume_render_cuvol_stats", "volume_render_expected_term",
    "volume_render_sigma_thresh", "tv_grad_sparse", "tv_grad_sparse_run",
    "rmsprop_step", "sgd_step", "rmsprop_step_grouped", "dilate", "grid_weight_render",
    "resample_density", "resample_rows", "threshold_mask", "mask_to_links", "sample_grid", "sample_grid_backward", "accel_dist_prop",
    "library_path", "load_library", "EXPORTED_SYMBOLS",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libsvox2_amd.so")

STORAGE_F32 = 0
STORAGE_BF16 = 1
BASIS_TYPE_SH = 1
INDEXER_ALL, INDEXER_MASK, INDEXER_INDEX = 0, 1, 2

# Every function the C header declares (checked against include/svox2_amd.h by
# the tests).
EXPORTED_SYMBOLS = (
    "svox2_volume_render_cuvol", "svox2_volume_render_cuvol_ordered",
    "svox2_volume_render_cuvol_image",
    "svox2_volume_render_cuvol_backward", "svox2_volume_render_cuvol_fused",
    "svox2_volume_render_cuvol_fused_noalloc",
    "svox2_ray_order_workspace_bytes", "svox2_volume_render_cuvol_fused_workspace_bytes",
    "svox2_background_workspace_bytes", "svox2_cell_map_bytes",
    "svox2_build_cell_map", "svox2_volume_render_cuvol_tiles",
    "svox2_volume_render_cuvol_stats", "svox2_volume_render_expected_term",
    "svox2_volume_render_sigma_thresh", "svox2_volume_render_nvol",
    "svox2_volume_render_nvol_backward", "svox2_volume_render_nvol_fused",
    "svox2_volume_render_svox1", "svox2_volume_render_svox1_backward",
    "svox2_volume_render_svox1_fused",
    "svox2_sample_grid", "svox2_sample_grid_backward",
    "svox2_tv_grad_sparse", "svox2_tv_grad_sparse_run", "svox2_lumisphere_tv_grad_sparse",
    "svox2_msi_tv_grad_sparse", "svox2_rmsprop_step", "svox2_sgd_step",
    "svox2_rmsprop_step_grouped", "svox2_accel_dist_prop_workspace_bytes",
    "svox2_accel_dist_prop", "svox2_resample_density", "svox2_resample_rows",
    "svox2_grid_weight_render", "svox2_threshold_mask", "svox2_dilate",
    "svox2_mask_compact_workspace_bytes", "svox2_mask_compact_count",
    "svox2_mask_compact_assign", "svox2_last_error", "svox2_version",
)


# ---------------------------------------------------------------------------
# ctypes mirrors of the C structs
# ---------------------------------------------------------------------------
class _Grid(ctypes.Structure):
    _fields_ = [
        ("links", ctypes.c_void_p),
        ("density_data", ctypes.c_void_p),
        ("sh_data", ctypes.c_void_p),
        ("rows", ctypes.c_int64),
        ("size", ctypes.c_int32 * 3),
        ("basis_dim", ctypes.c_int32),
        ("sh_data_dim", ctypes.c_int32),
        ("basis_type", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("offset", ctypes.c_float * 3),
        ("scaling", ctypes.c_float * 3),
        ("density_stride", ctypes.c_int64),
        ("sh_stride", ctypes.c_int64),
        ("cell_map", ctypes.c_void_p),
        ("background_links", ctypes.c_void_p),
        ("background_data", ctypes.c_void_p),
        ("background_rows", ctypes.c_int64),
        ("background_reso", ctypes.c_int32),
        ("background_nlayers", ctypes.c_int32),
    ]


class _Opt(ctypes.Structure):
    _fields_ = [
        ("background_brightness", ctypes.c_float),
        ("step_size", ctypes.c_float),
        ("sigma_thresh", ctypes.c_float),
        ("stop_thresh", ctypes.c_float),
        ("near_clip", ctypes.c_float),
        ("use_spheric_clip", ctypes.c_int32),
        ("last_sample_opaque", ctypes.c_int32),
    ]


class _Cam(ctypes.Structure):
    _fields_ = [
        ("c2w", ctypes.c_float * 12),
        ("fx", ctypes.c_float), ("fy", ctypes.c_float),
        ("cx", ctypes.c_float), ("cy", ctypes.c_float),
        ("width", ctypes.c_int32), ("height", ctypes.c_int32),
        ("ndc_coeffx", ctypes.c_float), ("ndc_coeffy", ctypes.c_float),
    ]


class _Tile(ctypes.Structure):
    _fields_ = [
        ("grid", ctypes.c_int32), ("camera", ctypes.c_int32),
        ("pixel0", ctypes.c_int32), ("n_pixels", ctypes.c_int32),
        ("rgb_out", ctypes.c_void_p),
    ]


class _Grads(ctypes.Structure):
    _fields_ = [
        ("grad_density_out", ctypes.c_void_p),
        ("grad_sh_out", ctypes.c_void_p),
        ("mask_out", ctypes.c_void_p),
        ("grad_density_stride", ctypes.c_int64),
        ("grad_sh_stride", ctypes.c_int64),
        ("grad_background_out", ctypes.c_void_p),
        ("mask_background_out", ctypes.c_void_p),
    ]


_lib: Optional[ctypes.CDLL] = None


def load_library(path: Optional[str] = None) -> ctypes.CDLL:
    """Loads (once) and returns the kernel library; raises if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or library_path
    if not os.path.exists(p):
        raise RuntimeError(
            f"svox2_amd kernel library not found at {p}; build it with "
            "`python -m svox2_amd.build` (hipcc, gfx950)")
    # torch is imported first so that the library binds to the HIP runtime
    # torch already loaded (same soname), giving one runtime per process.
    lib = ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
    vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
    G, O, Cm, Gr = (ctypes.POINTER(_Grid), ctypes.POINTER(_Opt), ctypes.POINTER(_Cam),
                    ctypes.POINTER(_Grads))
    sig = {
        "svox2_volume_render_cuvol": [G, vp, vp, vp, i64, O, vp, vp, vp, vp],
        "svox2_volume_render_cuvol_ordered": [G, vp, vp, vp, i64, O, vp, vp, vp, vp,
                                              ctypes.c_size_t, vp],
        "svox2_volume_render_cuvol_image": [G, Cm, O, vp, vp, vp],
        "svox2_volume_render_cuvol_backward": [G, vp, vp, vp, i64, O, vp, vp, Gr, vp, vp,
                                               ctypes.c_size_t, vp],
        "svox2_volume_render_cuvol_fused": [G, vp, vp, vp, i64, O, vp, f32, f32, vp, vp, Gr,
                                            vp, ctypes.c_size_t, vp],
        "svox2_volume_render_cuvol_fused_noalloc": [G, vp, vp, vp, i64, O, vp, f32, f32, vp, vp,
                                                    Gr, vp],
        "svox2_volume_render_cuvol_stats": [G, vp, vp, vp, i64, O, vp, vp],
        "svox2_volume_render_expected_term": [G, vp, vp, vp, i64, O, vp, vp],
        "svox2_volume_render_sigma_thresh": [G, vp, vp, vp, i64, O, f32, vp, vp],
        "svox2_volume_render_nvol": [G, vp, vp, vp, i64, O, vp, vp],
        "svox2_volume_render_nvol_backward": [G, vp, vp, vp, i64, O, vp, vp, Gr, vp],
        "svox2_volume_render_nvol_fused": [G, vp, vp, vp, i64, O, vp, f32, f32, vp, Gr, vp],
        "svox2_volume_render_svox1": [G, vp, vp, vp, i64, O, vp, vp],
        "svox2_volume_render_svox1_backward": [G, vp, vp, vp, i64, O, vp, vp, Gr, vp],
        "svox2_volume_render_svox1_fused": [G, vp, vp, vp, i64, O, vp, f32, f32, vp, Gr, vp],
        "svox2_sample_grid": [G, vp, i64, i32, vp, vp, vp],
        "svox2_sample_grid_backward": [G, vp, i64, vp, vp, vp, vp, i64, i64, i32, vp],
        "svox2_tv_grad_sparse": [vp, ctypes.POINTER(i32), vp, i32, i64, i32, i64, i64, vp, i64,
                                 vp, i32, i32, f32, i32, f32, i32, i32, f32, f32, vp, vp],
        "svox2_tv_grad_sparse_run": [vp, ctypes.POINTER(i32), vp, i32, i64, i32, i64, i64, i64,
                                     i64, vp, i32, i32, f32, i32, i32, vp, vp],
        "svox2_lumisphere_tv_grad_sparse": [G, vp, i64, vp, vp, f32, f32, f32, f32, Gr, vp],
        "svox2_msi_tv_grad_sparse": [vp, i32, vp, i64, i32, i32, vp, i64, vp, f32, f32, vp, vp],
        "svox2_rmsprop_step": [vp, i32, vp, vp, i64, i32, i64, i64, i64, i32, vp, i64, f32, f32,
                               f32, f32, f32, vp],
        "svox2_sgd_step": [vp, i32, vp, i64, i32, i64, i64, i32, vp, i64, f32, f32, vp],
        "svox2_rmsprop_step_grouped": [vp, i32, vp, vp, i64, i32, i64, i64, i64, i32, vp, vp,
                                       f32, f32, f32, f32, f32, f32, vp],
        "svox2_accel_dist_prop": [vp, ctypes.POINTER(i32), vp, ctypes.c_size_t, vp],
        "svox2_build_cell_map": [vp, ctypes.POINTER(i32), i64, vp, ctypes.c_size_t, vp],
        "svox2_resample_density": [G, vp, ctypes.POINTER(i32), vp, vp],
        "svox2_resample_rows": [G, vp, ctypes.POINTER(i32), vp, i64, vp, vp, i64, vp, i64, i32,
                                vp],
        "svox2_grid_weight_render": [vp, ctypes.POINTER(i32), Cm, f32, f32, i32,
                                     ctypes.POINTER(f32), ctypes.POINTER(f32), vp, vp],
        "svox2_threshold_mask": [vp, ctypes.POINTER(i32), f32, i32, vp, vp],
        "svox2_dilate": [vp, ctypes.POINTER(i32), vp, vp],
        "svox2_mask_compact_count": [vp, ctypes.POINTER(i32), i32, vp, ctypes.c_size_t, vp],
        "svox2_mask_compact_assign": [vp, ctypes.POINTER(i32), i32, vp, vp, vp, vp],
        "svox2_volume_render_cuvol_tiles": [vp, i32, i32, i32, vp, i32, O, vp, vp, i32, i32,
                                            vp],
    }
    for name, argtypes in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int
    lib.svox2_cell_map_bytes.argtypes = [ctypes.POINTER(i32)]
    lib.svox2_cell_map_bytes.restype = ctypes.c_size_t
    lib.svox2_volume_render_cuvol_fused_workspace_bytes.argtypes = [i64]
    lib.svox2_volume_render_cuvol_fused_workspace_bytes.restype = ctypes.c_size_t
    lib.svox2_ray_order_workspace_bytes.argtypes = [i64]
    lib.svox2_ray_order_workspace_bytes.restype = ctypes.c_size_t
    lib.svox2_background_workspace_bytes.argtypes = [i64]
    lib.svox2_background_workspace_bytes.restype = ctypes.c_size_t
    lib.svox2_mask_compact_workspace_bytes.argtypes = [ctypes.POINTER(i32)]
    lib.svox2_mask_compact_workspace_bytes.restype = ctypes.c_size_t
    lib.svox2_accel_dist_prop_workspace_bytes.argtypes = [ctypes.POINTER(i32)]
    lib.svox2_accel_dist_prop_workspace_bytes.restype = ctypes.c_size_t
    lib.svox2_last_error.restype = ctypes.c_char_p
    lib.svox2_version.restype = ctypes.c_char_p
    if path is None:
        _lib = lib
    return lib


def _call(name: str, *args) -> None:
    lib = load_library()
    st = getattr(lib, name)(*args)
    if st != 0:
        msg = lib.svox2_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed (status {st}): {msg}")


# ---------------------------------------------------------------------------
# Spec classes (same fields as the reference's C++ structs)
# ---------------------------------------------------------------------------
class SparseGridSpec:
    """SparseGridSpec (data_spec.hpp:18-52)."""

    def __init__(self):
        self.density_data: Optional[torch.Tensor] = None
        self.sh_data: Optional[torch.Tensor] = None
        self.links: Optional[torch.Tensor] = None
        self._offset: Optional[torch.Tensor] = None
        self._scaling: Optional[torch.Tensor] = None
        self.background_links: Optional[torch.Tensor] = None
        self.background_data: Optional[torch.Tensor] = None
        self.basis_dim: int = 9
        self.basis_type: int = BASIS_TYPE_SH
        self.basis_data: Optional[torch.Tensor] = None
        # Cell map of the links (svox2_build_cell_map).  None: built on first
        # use and cached per links tensor (rebuilt when the links change);
        # False: the march reads links only.
        self.cell_map = None


class RaysSpec:
    """RaysSpec (data_spec.hpp:104-113), plus an optional per-ray (t_near, t_far)
    world-distance clip ``clip`` of shape (n, 2)."""

    def __init__(self):
        self.origins: Optional[torch.Tensor] = None
        self.dirs: Optional[torch.Tensor] = None
        self.clip: Optional[torch.Tensor] = None


class CameraSpec:
    """CameraSpec (data_spec.hpp:84-102)."""

    def __init__(self):
        self.c2w: Optional[torch.Tensor] = None
        self.fx = self.fy = self.cx = self.cy = 0.0
        self.width = self.height = 0
        self.ndc_coeffx = self.ndc_coeffy = -1.0


class RenderOptions:
    """RenderOptions (data_spec.hpp:115-135)."""

    def __init__(self):
        self.background_brightness = 1.0
        self.step_size = 0.5
        self.sigma_thresh = 1e-10
        self.stop_thresh = 1e-7
        self.near_clip = 0.0
        self.use_spheric_clip = False
        self.last_sample_opaque = False


class GridOutputGrads:
    """GridOutputGrads (data_spec.hpp:54-82)."""

    def __init__(self):
        self.grad_density_out: Optional[torch.Tensor] = None
        self.grad_sh_out: Optional[torch.Tensor] = None
        self.grad_basis_out: Optional[torch.Tensor] = None
        self.grad_background_out: Optional[torch.Tensor] = None
        self.mask_out: Optional[torch.Tensor] = None
        self.mask_background_out: Optional[torch.Tensor] = None


# ---------------------------------------------------------------------------
# Checks (CHECK_INPUT / TORCH_CHECK equivalents)
# ---------------------------------------------------------------------------
def _check(t: Optional[torch.Tensor], name: str, dtypes=None, ndim: Optional[int] = None):
    if t is None:
        raise RuntimeError(f"{name} must be defined")
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} must be a tensor")
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if dtypes is not None and t.dtype not in dtypes:
        raise RuntimeError(f"{name} must have dtype in {dtypes}, got {t.dtype}")
    if ndim is not None and t.dim() != ndim:
        raise RuntimeError(f"{name} must have {ndim} dims, got {t.dim()}")
    return t


def _check_rows(t: Optional[torch.Tensor], name: str, dtypes=None, cols: Optional[int] = None):
    """A 2-d device tensor whose rows are contiguous but may be strided apart
    (the interleaved row layout): stride(1) == 1 and stride(0) >= size(1)."""
    if t is None:
        raise RuntimeError(f"{name} must be defined")
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} must be a tensor")
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if t.dim() != 2:
        raise RuntimeError(f"{name} must have 2 dims, got {t.dim()}")
    if dtypes is not None and t.dtype not in dtypes:
        raise RuntimeError(f"{name} must have dtype in {dtypes}, got {t.dtype}")
    if cols is not None and t.size(1) != cols:
        raise RuntimeError(f"{name} must have {cols} columns, got {t.size(1)}")
    if t.size(0) > 1 and not ((t.size(1) == 1 or t.stride(1) == 1) and t.stride(0) >= t.size(1)):
        raise RuntimeError(f"{name} must be contiguous or row-strided (stride(1) == 1)")
    return t


def _row_stride(t: torch.Tensor) -> int:
    return int(t.stride(0)) if t.size(0) > 1 else int(t.size(1))


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(t: torch.Tensor):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


class _on_device:
    """Makes the tensor's device current for the duration of a call
    (the reference's DEVICE_GUARD, include/cuda_util.cuh:10-11)."""

    def __init__(self, t: torch.Tensor):
        self.idx = t.device.index
        self.prev = None

    def __enter__(self):
        cur = torch.cuda.current_device()
        if self.idx is not None and cur != self.idx:
            self.prev = cur
            torch.cuda.set_device(self.idx)

    def __exit__(self, *exc):
        if self.prev is not None:
            torch.cuda.set_device(self.prev)


_DATA_DTYPES = (torch.float32, torch.bfloat16)


def _storage_of(t: torch.Tensor) -> int:
    return STORAGE_BF16 if t.dtype == torch.bfloat16 else STORAGE_F32


def _vec3(v) -> Tuple[float, float, float]:
    if isinstance(v, torch.Tensor):
        if v.is_cuda:
            raise RuntimeError("_offset/_scaling must be CPU tensors")
        v = v.detach().double().flatten().tolist()
    v = [float(x) for x in v]
    if len(v) != 3:
        raise RuntimeError("_offset/_scaling must have 3 entries")
    return v[0], v[1], v[2]


def _grid_struct(g: SparseGridSpec, need_sh: bool = True, need_map: bool = True) -> _Grid:
    links = _check(g.links, "links", (torch.int32,), 3)
    dens = _check_rows(g.density_data, "density_data", _DATA_DTYPES)
    sh = g.sh_data
    if need_sh or sh is not None:
        sh = _check_rows(sh, "sh_data", _DATA_DTYPES)
        if sh.dtype != dens.dtype:
            raise RuntimeError("density_data and sh_data must share a storage dtype")
        if sh.size(0) != dens.size(0):
            raise RuntimeError("density_data and sh_data row counts differ")
    if dens.size(1) != 1:
        raise RuntimeError("density_data must be (rows, 1)")
    if g.basis_type != BASIS_TYPE_SH:
        raise RuntimeError(f"basis_type {g.basis_type} not supported (SH only)")
    s = _Grid()
    s.links = links.data_ptr()
    s.density_data = dens.data_ptr()
    s.sh_data = sh.data_ptr() if sh is not None else None
    s.rows = dens.size(0)
    for i in range(3):
        s.size[i] = links.size(i)
    s.basis_dim = int(g.basis_dim)
    s.sh_data_dim = sh.size(1) if sh is not None else 3 * int(g.basis_dim)
    s.basis_type = int(g.basis_type)
    s.storage = _storage_of(dens)
    off, sc = _vec3(g._offset), _vec3(g._scaling)
    for i in range(3):
        s.offset[i] = off[i]
        s.scaling[i] = sc[i]
    s.density_stride = _row_stride(dens)
    s.sh_stride = _row_stride(sh) if sh is not None else 0
    cmap = g.cell_map if need_map else False
    if cmap is None:
        cmap = _cached_cell_map(links, dens.size(0))
    if cmap is not False:
        need = load_library().svox2_cell_map_bytes(s.size)
        _check(cmap, "cell_map", (torch.int16,), 1)
        if cmap.numel() * 2 < need or cmap.device != links.device:
            raise RuntimeError(f"cell_map must hold {need} bytes on the links' device")
        s.cell_map = cmap.data_ptr()
    bl, bd = g.background_links, g.background_data
    if bl is not None and bl.numel() > 0:
        # MSI background (data_spec.hpp:36-43): links (2 reso, reso) int32,
        # data (rows, nlayers, 4) float32.
        bl = _check(bl, "background_links", (torch.int32,), 2)
        if bd is None:
            raise RuntimeError("background_links given without background_data")
        bd = _check(bd, "background_data", (torch.float32,), 3)
        if bl.size(0) != 2 * bl.size(1):
            raise RuntimeError("background_links must be (2 reso, reso)")
        if bd.size(2) != 4 or bd.size(1) < 2:
            raise RuntimeError("background_data must be (rows, nlayers >= 2, 4)")
        if bl.device != links.device or bd.device != links.device:
            raise RuntimeError("background tensors must be on the links' device")
        s.background_links = bl.data_ptr()
        s.background_data = bd.data_ptr()
        s.background_rows = bd.size(0)
        s.background_reso = bl.size(1)
        s.background_nlayers = bd.size(1)
    return s


# Cell maps by links tensor: (device, address, shape, version, rows, accel
# epoch) -> map.  The accel epoch counts in-place accel_dist_prop calls, which
# write through a raw pointer and so do not bump the tensor version.
_CELL_MAPS: "OrderedDict[tuple, torch.Tensor]" = OrderedDict()
_CELL_MAP_CACHE = 4
_ACCEL_EPOCH: dict = {}


def _cached_cell_map(links: torch.Tensor, rows: int) -> torch.Tensor:
    key = (links.device.index, links.data_ptr(), tuple(links.shape), links._version, int(rows),
           _ACCEL_EPOCH.get((links.device.index, links.data_ptr()), 0))
    cmap = _CELL_MAPS.get(key)
    if cmap is None:
        cmap = build_cell_map(links, rows)
        # Drop stale maps of the same links tensor, then the oldest ones.
        for k in [k for k in _CELL_MAPS if k[:3] == key[:3]]:
            del _CELL_MAPS[k]
        _CELL_MAPS[key] = cmap
        while len(_CELL_MAPS) > _CELL_MAP_CACHE:
            _CELL_MAPS.popitem(last=False)
    else:
        _CELL_MAPS.move_to_end(key)
    return cmap


def build_cell_map(links: torch.Tensor, rows: int) -> torch.Tensor:
    """svox2_build_cell_map: the cell map (int16 words) of ``links`` for a grid
    of ``rows`` data rows, on the links' device and current stream."""
    _check(links, "links", (torch.int32,), 3)
    lib = load_library()
    size = (ctypes.c_int32 * 3)(*links.shape)
    nbytes = int(lib.svox2_cell_map_bytes(size))
    cmap = torch.empty(max(nbytes // 2, 1), dtype=torch.int16, device=links.device)
    with _on_device(links):
        _call("svox2_build_cell_map", _ptr(links), size, int(rows), _ptr(cmap),
              ctypes.c_size_t(nbytes), _stream(links))
    return cmap


def _opt_struct(o: RenderOptions) -> _Opt:
    s = _Opt()
    s.background_brightness = float(o.background_brightness)
    s.step_size = float(o.step_size)
    s.sigma_thresh = float(o.sigma_thresh)
    s.stop_thresh = float(o.stop_thresh)
    s.near_clip = float(o.near_clip)
    s.use_spheric_clip = int(bool(o.use_spheric_clip))
    s.last_sample_opaque = int(bool(o.last_sample_opaque))
    return s


def _rays(r: RaysSpec):
    o = _check(r.origins, "origins", (torch.float32,), 2)
    d = _check(r.dirs, "dirs", (torch.float32,), 2)
    if o.size(1) != 3 or d.shape != o.shape:
        raise RuntimeError("origins/dirs must both be (n, 3)")
    clip = getattr(r, "clip", None)
    if clip is not None:
        _check(clip, "clip", (torch.float32,), 2)
        if clip.shape != (o.size(0), 2):
            raise RuntimeError("clip must be (n, 2)")
    return o, d, clip


def _grads_struct(gr: GridOutputGrads, rows: int, sh_dim: int, grid: Optional[_Grid] = None
                  ) -> _Grads:
    s = _Grads()
    if gr.grad_density_out is not None:
        t = _check_rows(gr.grad_density_out, "grad_density_out", (torch.float32,), 1)
        if t.size(0) != rows:
            raise RuntimeError("grad_density_out must have one value per row")
        s.grad_density_out = t.data_ptr()
        s.grad_density_stride = _row_stride(t)
    if gr.grad_sh_out is not None:
        t = _check_rows(gr.grad_sh_out, "grad_sh_out", (torch.float32,), sh_dim)
        if t.size(0) != rows:
            raise RuntimeError("grad_sh_out must be (rows, sh_data_dim)")
        s.grad_sh_out = t.data_ptr()
        s.grad_sh_stride = _row_stride(t)
    if gr.mask_out is not None and gr.mask_out.numel() > 0:
        t = _check(gr.mask_out, "mask_out", (torch.bool, torch.uint8))
        if t.numel() != rows:
            raise RuntimeError("mask_out must have one entry per row")
        s.mask_out = t.data_ptr()
    if gr.grad_basis_out is not None:
        raise RuntimeError("basis gradients are not supported (SH basis only)")
    # Background gradients, for the entry points that composite a background
    # (the others ignore them, as the reference's do).
    if grid is not None and grid.background_nlayers > 0:
        shape = (grid.background_rows, grid.background_nlayers)
        if gr.grad_background_out is not None:
            t = _check(gr.grad_background_out, "grad_background_out", (torch.float32,), 3)
            if tuple(t.shape) != shape + (4,):
                raise RuntimeError("grad_background_out must be (rows, nlayers, 4)")
            s.grad_background_out = t.data_ptr()
        if gr.mask_background_out is not None and gr.mask_background_out.numel() > 0:
            t = _check(gr.mask_background_out, "mask_background_out", (torch.bool, torch.uint8))
            if t.numel() != shape[0] * shape[1]:
                raise RuntimeError("mask_background_out must have rows * nlayers entries")
            s.mask_background_out = t.data_ptr()
    return s


# ---------------------------------------------------------------------------
# Entry points
# ---------------------------------------------------------------------------
def volume_render_cuvol(grid: SparseGridSpec, rays: RaysSpec, opt: RenderOptions,
                        log_transmit_out: Optional[torch.Tensor] = None,
                        ray_cost_out: Optional[torch.Tensor] = None,
                        ordered: bool = False) -> torch.Tensor:
    """volume_render_cuvol (render_lerp_kernel_cuvol.cu:918-958): (n, 3) colours.

    ray_cost_out: optional (n,) int32 tensor receiving each ray's march chunk
    count, for ``volume_render_cuvol_backward(..., ray_cost=...)``.
    ordered: large batches run in descending order of segment length, with a
    workspace from torch's caching allocator (same results).
    """
    g = _grid_struct(grid)
    o, d, clip = _rays(rays)
    out = torch.empty_like(o)
    for t, name in ((log_transmit_out, "log_transmit_out"), (ray_cost_out, "ray_cost_out")):
        if t is not None:
            _check(t, name, (torch.float32,) if t is log_transmit_out else (torch.int32,))
            if t.numel() != o.size(0):
                raise RuntimeError(f"{name} must have one value per ray")
    if g.background_nlayers > 0 and log_transmit_out is None:
        log_transmit_out = torch.empty(o.size(0), dtype=torch.float32, device=o.device)
    nbytes = int(load_library().svox2_ray_order_workspace_bytes(o.size(0))) if ordered else 0
    ws = torch.empty(nbytes, dtype=torch.uint8, device=o.device) if nbytes else None
    with _on_device(o):
        _call("svox2_volume_render_cuvol_ordered", ctypes.byref(g), _ptr(o), _ptr(d),
              _ptr(clip), o.size(0), ctypes.byref(_opt_struct(opt)), _ptr(out),
              _ptr(log_transmit_out), _ptr(ray_cost_out), _ptr(ws), nbytes, _stream(o))
    return out


def volume_render_cuvol_stats(grid: SparseGridSpec, rays: RaysSpec,
                              opt: RenderOptions) -> torch.Tensor:
    """Per-ray (iterations, evaluated, contributing, hash) int32 counters of the
    forward march (no reference counterpart; used for measurement and bit-exact
    index tests)."""
    g = _grid_struct(grid)
    o, d, clip = _rays(rays)
    out = torch.empty((o.size(0), 4), dtype=torch.int32, device=o.device)
    with _on_device(o):
        _call("svox2_volume_render_cuvol_stats", ctypes.byref(g), _ptr(o), _ptr(d), _ptr(clip),
              o.size(0), ctypes.byref(_opt_struct(opt)), _ptr(out), _stream(o))
    return out


def volume_render_expected_term(grid: SparseGridSpec, rays: RaysSpec,
                                opt: RenderOptions) -> torch.Tensor:
    """volume_render_expected_term (render_lerp_kernel_cuvol.cu:1154-1174):
    (n,) expected termination distance along each ray."""
    g = _grid_struct(grid, need_sh=False)
    o, d, clip = _rays(rays)
    out = torch.empty(o.size(0), dtype=torch.float32, device=o.device)
    with _on_device(o):
        _call("svox2_volume_render_expected_term", ctypes.byref(g), _ptr(o), _ptr(d), _ptr(clip),
              o.size(0), ctypes.byref(_opt_struct(opt)), _ptr(out), _stream(o))
    return out


def volume_render_sigma_thresh(grid: SparseGridSpec, rays: RaysSpec, opt: RenderOptions,
                               sigma_thresh: float) -> torch.Tensor:
    """volume_render_sigma_thresh (render_lerp_kernel_cuvol.cu:1176-1197): (n,)
    distance of the first sample with density above sigma_thresh (0 if none)."""
    g = _grid_struct(grid, need_sh=False)
    o, d, clip = _rays(rays)
    out = torch.empty(o.size(0), dtype=torch.float32, device=o.device)
    with _on_device(o):
        _call("svox2_volume_render_sigma_thresh", ctypes.byref(g), _ptr(o), _ptr(d), _ptr(clip),
              o.size(0), ctypes.byref(_opt_struct(opt)), float(sigma_thresh), _ptr(out),
              _stream(o))
    return out


def _cam_struct(cam: CameraSpec) -> _Cam:
    c2w = cam.c2w
    if not isinstance(c2w, torch.Tensor) or c2w.dim() != 2 or c2w.size(1) != 4 \
            or c2w.size(0) < 3 or not c2w.is_floating_point():
        raise RuntimeError("c2w must be a floating (3 or 4, 4) tensor")
    vals = c2w[:3].detach().double().cpu().flatten().tolist()
    s = _Cam()
    for i in range(12):
        s.c2w[i] = vals[i]
    s.fx, s.fy, s.cx, s.cy = float(cam.fx), float(cam.fy), float(cam.cx), float(cam.cy)
    s.width, s.height = int(cam.width), int(cam.height)
    s.ndc_coeffx, s.ndc_coeffy = float(cam.ndc_coeffx), float(cam.ndc_coeffy)
    return s


def volume_render_cuvol_image(grid: SparseGridSpec, cam: CameraSpec, opt: RenderOptions,
                              log_transmit_out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """volume_render_cuvol_image (render_lerp_kernel_cuvol.cu:960-1008): (H, W, 3)."""
    g = _grid_struct(grid)
    dev = grid.sh_data.device
    out = torch.empty((int(cam.height), int(cam.width), 3), dtype=torch.float32, device=dev)
    if log_transmit_out is not None:
        _check(log_transmit_out, "log_transmit_out", (torch.float32,))
        if log_transmit_out.numel() != int(cam.height) * int(cam.width):
            raise RuntimeError("log_transmit_out must have one value per pixel")
    elif g.background_nlayers > 0:
        log_transmit_out = torch.empty(int(cam.height) * int(cam.width), dtype=torch.float32,
                                       device=dev)
    with _on_device(grid.sh_data):
        _call("svox2_volume_render_cuvol_image", ctypes.byref(g), ctypes.byref(_cam_struct(cam)),
              ctypes.byref(_opt_struct(opt)), _ptr(out), _ptr(log_transmit_out),
              _stream(grid.sh_data))
    return out


def volume_render_cuvol_backward(grid: SparseGridSpec, rays: RaysSpec, opt: RenderOptions,
                                 grad_out: torch.Tensor, color_cache: torch.Tensor,
                                 grads: GridOutputGrads,
                                 ray_cost: Optional[torch.Tensor] = None) -> None:
    """volume_render_cuvol_backward (render_lerp_kernel_cuvol.cu:1010-1068).

    ray_cost: the forward's ``ray_cost_out``; with it, large batches run in
    descending order of cost (same results).
    """
    g = _grid_struct(grid)
    o, d, clip = _rays(rays)
    _check(grad_out, "grad_out", (torch.float32,), 2)
    _check(color_cache, "color_cache", (torch.float32,), 2)
    if grad_out.shape != o.shape or color_cache.shape != o.shape:
        raise RuntimeError("grad_out / color_cache must be (n, 3)")
    gs = _grads_struct(grads, g.rows, g.sh_data_dim, g)
    ws, nbytes = None, 0
    if ray_cost is not None:
        _check(ray_cost, "ray_cost", (torch.int32,))
        if ray_cost.numel() != o.size(0):
            raise RuntimeError("ray_cost must have one value per ray")
        nbytes = int(load_library().svox2_ray_order_workspace_bytes(o.size(0)))
    if g.background_nlayers > 0:
        nbytes = int(load_library().svox2_background_workspace_bytes(o.size(0)))
    if nbytes:
        ws = torch.empty(nbytes, dtype=torch.uint8, device=o.device)
    with _on_device(o):
        _call("svox2_volume_render_cuvol_backward", ctypes.byref(g), _ptr(o), _ptr(d),
              _ptr(clip), o.size(0), ctypes.byref(_opt_struct(opt)), _ptr(grad_out),
              _ptr(color_cache), ctypes.byref(gs), _ptr(ray_cost), _ptr(ws), nbytes,
              _stream(o))


def volume_render_cuvol_fused(grid: SparseGridSpec, rays: RaysSpec, opt: RenderOptions,
                              rgb_gt: torch.Tensor, beta_loss: float, sparsity_loss: float,
                              rgb_out: torch.Tensor, grads: GridOutputGrads,
                              ordered: bool = True, record: bool = True) -> None:
    """volume_render_cuvol_fused (render_lerp_kernel_cuvol.cu:1070-1152).

    ordered: run the backward over the rays in descending order of the work
    the forward measured (same results, shorter tail); False keeps input order.
    record (with ordered): the forward records every contributing sample and
    the backward replays the records instead of re-marching (same results).
    """
    g = _grid_struct(grid)
    o, d, clip = _rays(rays)
    _check(rgb_gt, "rgb_gt", (torch.float32,), 2)
    _check(rgb_out, "rgb_out", (torch.float32,), 2)
    if rgb_gt.shape != o.shape or rgb_out.shape != o.shape:
        raise RuntimeError("rgb_gt / rgb_out must be (n, 3)")
    gs = _grads_struct(grads, g.rows, g.sh_data_dim, g)
    log_t = (torch.empty(o.size(0), dtype=torch.float32, device=o.device)
             if beta_loss > 0 else None)
    # Workspace for the ray ordering and, for small batches, the record
    # stream (from torch's caching allocator).
    lib = load_library()
    if ordered and record:
        nbytes = lib.svox2_volume_render_cuvol_fused_workspace_bytes(o.size(0))
    elif ordered:
        nbytes = lib.svox2_ray_order_workspace_bytes(o.size(0))
    else:
        nbytes = 0
    if g.background_nlayers > 0:  # the background's per-ray buffers
        nbytes = max(nbytes, lib.svox2_background_workspace_bytes(o.size(0)))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=o.device) if nbytes else None
    with _on_device(o):
        _call("svox2_volume_render_cuvol_fused", ctypes.byref(g), _ptr(o), _ptr(d), _ptr(clip),
              o.size(0), ctypes.byref(_opt_struct(opt)), _ptr(rgb_gt), float(beta_loss),
              float(sparsity_loss), _ptr(rgb_out), _ptr(log_t), ctypes.byref(gs), _ptr(ws),
              nbytes, _stream(o))


def _no_clip(rays: RaysSpec, what: str):
    if getattr(rays, "clip", None) is not None:
        raise RuntimeError(f"{what}: the svox1 backend takes no ray clips")


def volume_render_svox1(grid: SparseGridSpec, rays: RaysSpec, opt: RenderOptions) -> torch.Tensor:
    """volume_render_svox1 (render_svox1_kernel.cu:334-349): (n, 3) colours."""
    _no_clip(rays, "volume_render_svox1")
    g = _grid_struct(grid)
    o, d, _ = _rays(rays)
    out = torch.empty(o.size(0), 3, dtype=torch.float32, device=o.device)
    with _on_device(o):
        _call("svox2_volume_render_svox1", ctypes.byref(g), _ptr(o), _ptr(d), None, o.size(0),
              ctypes.byref(_opt_struct(opt)), _ptr(out), _stream(o))
    return out


def volume_render_svox1_backward(grid: SparseGridSpec, rays: RaysSpec, opt: RenderOptions,
                                 grad_out: torch.Tensor, color_cache: torch.Tensor,
                                 grads: GridOutputGrads) -> None:
    """volume_render_svox1_backward (render_svox1_kernel.cu:351-378)."""
    _no_clip(rays, "volume_render_svox1_backward")
    g = _grid_struct(grid)
    o, d, _ = _rays(rays)
    _check(grad_out, "grad_out", (torch.float32,), 2)
    _check(color_cache, "color_cache", (torch.float32,), 2)
    if grad_out.shape != o.shape or color_cache.shape != o.shape:
        raise RuntimeError("grad_out / color_cache must be (n, 3)")
    gs = _grads_struct(grads, g.rows, g.sh_data_dim)
    with _on_device(o):
        _call("svox2_volume_render_svox1_backward", ctypes.byref(g), _ptr(o), _ptr(d), None,
              o.size(0), ctypes.byref(_opt_struct(opt)), _ptr(grad_out), _ptr(color_cache),
              ctypes.byref(gs), _stream(o))


def volume_render_svox1_fused(grid: SparseGridSpec, rays: RaysSpec, opt: RenderOptions,
                              rgb_gt: torch.Tensor, beta_loss: float, sparsity_loss: float,
                              rgb_out: torch.Tensor, grads: GridOutputGrads) -> None:
    """volume_render_svox1_fused (render_svox1_kernel.cu:380-419); beta_loss and
    sparsity_loss are ignored, as in the reference."""
    _no_clip(rays, "volume_render_svox1_fused")
    g = _grid_struct(grid)
    o, d, _ = _rays(rays)
    _check(rgb_gt, "rgb_gt", (torch.float32,), 2)
    _check(rgb_out, "rgb_out", (torch.float32,), 2)
    if rgb_gt.shape != o.shape or rgb_out.shape != o.shape:
        raise RuntimeError("rgb_gt / rgb_out must be (n, 3)")
    gs = _grads_struct(grads, g.rows, g.sh_data_dim)
    with _on_device(o):
        _call("svox2_volume_render_svox1_fused", ctypes.byref(g), _ptr(o), _ptr(d), None,
              o.size(0), ctypes.byref(_opt_struct(opt)), _ptr(rgb_gt), float(beta_loss),
              float(sparsity_loss), _ptr(rgb_out), ctypes.byref(gs), _stream(o))


def volume_render_nvol(grid: SparseGridSpec, rays: RaysSpec, opt: RenderOptions) -> torch.Tensor:
    """volume_render_nvol (render_lerp_kernel_nvol.cu:411-427): (n, 3) colours."""
    g = _grid_struct(grid)
    o, d, clip = _rays(rays)
    out = torch.empty(o.size(0), 3, dtype=torch.float32, device=o.device)
    with _on_device(o):
        _call("svox2_volume_render_nvol", ctypes.byref(g), _ptr(o), _ptr(d), _ptr(clip),
              o.size(0), ctypes.byref(_opt_struct(opt)), _ptr(out), _stream(o))
    return out


def volume_render_nvol_backward(grid: SparseGridSpec, rays: RaysSpec, opt: RenderOptions,
                                grad_out: torch.Tensor, color_cache: torch.Tensor,
                                grads: GridOutputGrads) -> None:
    """volume_render_nvol_backward (render_lerp_kernel_nvol.cu:431-458)."""
    g = _grid_struct(grid)
    o, d, clip = _rays(rays)
    _check(grad_out, "grad_out", (torch.float32,), 2)
    if grad_out.shape != o.shape:
        raise RuntimeError("grad_out must be (n, 3)")
    gs = _grads_struct(grads, g.rows, g.sh_data_dim)
    with _on_device(o):
        _call("svox2_volume_render_nvol_backward", ctypes.byref(g), _ptr(o), _ptr(d),
              _ptr(clip), o.size(0), ctypes.byref(_opt_struct(opt)), _ptr(grad_out),
              _ptr(color_cache), ctypes.byref(gs), _stream(o))


def volume_render_nvol_fused(grid: SparseGridSpec, rays: RaysSpec, opt: RenderOptions,
                             rgb_gt: torch.Tensor, beta_loss: float, sparsity_loss: float,
                             rgb_out: torch.Tensor, grads: GridOutputGrads) -> None:
    """volume_render_nvol_fused (render_lerp_kernel_nvol.cu:461-502)."""
    g = _grid_struct(grid)
    o, d, clip = _rays(rays)
    _check(rgb_gt, "rgb_gt", (torch.float32,), 2)
    _check(rgb_out, "rgb_out", (torch.float32,), 2)
    if rgb_gt.shape != o.shape or rgb_out.shape != o.shape:
        raise RuntimeError("rgb_gt / rgb_out must be (n, 3)")
    gs = _grads_struct(grads, g.rows, g.sh_data_dim)
    with _on_device(o):
        _call("svox2_volume_render_nvol_fused", ctypes.byref(g), _ptr(o), _ptr(d), _ptr(clip),
              o.size(0), ctypes.byref(_opt_struct(opt)), _ptr(rgb_gt), float(beta_loss),
              float(sparsity_loss), _ptr(rgb_out), ctypes.byref(gs), _stream(o))


def sample_grid(grid: SparseGridSpec, points: torch.Tensor,
                want_colors: bool) -> Tuple[torch.Tensor, torch.Tensor]:
    """sample_grid (svox2_kernel.cu:181-217): (density (n,1), sh (n,C) or (0,C))."""
    g = _grid_struct(grid, need_sh=bool(want_colors), need_map=False)
    _check(points, "points", (torch.float32,), 2)
    if points.size(1) != 3:
        raise RuntimeError("points must be (n, 3)")
    n = points.size(0)
    dens = torch.empty((n, 1), dtype=torch.float32, device=points.device)
    sh = torch.empty((n if want_colors else 0, g.sh_data_dim), dtype=torch.float32,
                     device=points.device)
    with _on_device(points):
        _call("svox2_sample_grid", ctypes.byref(g), _ptr(points), n, int(bool(want_colors)),
              _ptr(dens), _ptr(sh) if want_colors else None, _stream(points))
    return dens, sh


def sample_grid_backward(grid: SparseGridSpec, points: torch.Tensor,
                         grad_out_density: torch.Tensor, grad_out_sh: torch.Tensor,
                         grad_density_out: torch.Tensor, grad_sh_out: torch.Tensor,
                         want_colors: bool) -> None:
    """sample_grid_backward (svox2_kernel.cu:219-267)."""
    g = _grid_struct(grid, need_sh=False, need_map=False)
    _check(points, "points", (torch.float32,), 2)
    n = points.size(0)
    _check(grad_out_density, "grad_out_density", (torch.float32,), 2)
    _check_rows(grad_density_out, "grad_density_out", (torch.float32,), 1)
    if grad_out_density.shape != (n, 1) or grad_density_out.size(0) != g.rows:
        raise RuntimeError("density gradient shapes do not match")
    gss = 0
    if want_colors:
        _check(grad_out_sh, "grad_out_sh", (torch.float32,), 2)
        _check_rows(grad_sh_out, "grad_sh_out", (torch.float32,), g.sh_data_dim)
        if grad_out_sh.shape != (n, g.sh_data_dim) or grad_sh_out.size(0) != g.rows:
            raise RuntimeError("sh gradient shapes do not match")
        gss = _row_stride(grad_sh_out)
    with _on_device(points):
        _call("svox2_sample_grid_backward", ctypes.byref(g), _ptr(points), n,
              _ptr(grad_out_density), _ptr(grad_out_sh) if want_colors else None,
              _ptr(grad_density_out), _ptr(grad_sh_out) if want_colors else None,
              _row_stride(grad_density_out), gss, int(bool(want_colors)), _stream(points))


def tv_grad_sparse(links: torch.Tensor, data: torch.Tensor, rand_cells: torch.Tensor,
                   mask_out: torch.Tensor, start_dim: int, end_dim: int, scale: float,
                   use_logalpha: bool, logalpha_delta: float, ignore_edge: bool,
                   ignore_last_z: bool, ndc_coeffx: float, ndc_coeffy: float,
                   grad_data: torch.Tensor) -> None:
    """tv_grad_sparse (loss_kernel.cu:558-604)."""
    _check(links, "links", (torch.int32,), 3)
    _check_rows(data, "data", _DATA_DTYPES)
    _check_rows(grad_data, "grad_data", (torch.float32,))
    _check(rand_cells, "rand_cells", (torch.int32,))
    if grad_data.shape != data.shape:
        raise RuntimeError("grad_data must match data's shape")
    mask_ptr = None
    if mask_out is not None and mask_out.dim() > 0 and mask_out.numel() > 0:
        _check(mask_out, "mask_out", (torch.bool, torch.uint8))
        if mask_out.numel() != data.size(0):
            raise RuntimeError("mask_out must have one entry per row")
        mask_ptr = _ptr(mask_out)
    size = (ctypes.c_int32 * 3)(*links.shape)
    with _on_device(data):
        _call("svox2_tv_grad_sparse", _ptr(links), size, _ptr(data), _storage_of(data),
              data.size(0), data.size(1), _row_stride(data), _row_stride(grad_data),
              _ptr(rand_cells), rand_cells.numel(), mask_ptr,
              int(start_dim), int(end_dim), float(scale), int(bool(use_logalpha)),
              float(logalpha_delta), int(bool(ignore_edge)), int(bool(ignore_last_z)),
              float(ndc_coeffx), float(ndc_coeffy), _ptr(grad_data), _stream(data))


def tv_grad_sparse_run(links: torch.Tensor, data: torch.Tensor, first_cell: int, n_cells: int,
                       mask_out: Optional[torch.Tensor], start_dim: int, end_dim: int,
                       scale: float, ignore_edge: bool, ignore_last_z: bool,
                       grad_data: torch.Tensor) -> None:
    """tv_grad_sparse over the cells first_cell, first_cell + 1, ... (n_cells
    of them, wrapping past the last cell), generated in the kernel: the
    contiguous run of the sparse regularisers without a cell tensor.  No
    reference counterpart."""
    _check(links, "links", (torch.int32,), 3)
    _check_rows(data, "data", _DATA_DTYPES)
    _check_rows(grad_data, "grad_data", (torch.float32,))
    if grad_data.shape != data.shape:
        raise RuntimeError("grad_data must match data's shape")
    mask_ptr = None
    if mask_out is not None and mask_out.dim() > 0 and mask_out.numel() > 0:
        _check(mask_out, "mask_out", (torch.bool, torch.uint8))
        if mask_out.numel() != data.size(0):
            raise RuntimeError("mask_out must have one entry per row")
        mask_ptr = _ptr(mask_out)
    size = (ctypes.c_int32 * 3)(*links.shape)
    with _on_device(data):
        _call("svox2_tv_grad_sparse_run", _ptr(links), size, _ptr(data), _storage_of(data),
              data.size(0), data.size(1), _row_stride(data), _row_stride(grad_data),
              int(first_cell), int(n_cells), mask_ptr, int(start_dim), int(end_dim),
              float(scale), int(bool(ignore_edge)), int(bool(ignore_last_z)), _ptr(grad_data),
              _stream(data))


def lumisphere_tv_grad_sparse(grid: SparseGridSpec, rand_cells: torch.Tensor,
                              basis_fn: torch.Tensor, basis_fn_u: torch.Tensor, scale: float,
                              ndc_coeffx: float, ndc_coeffy: float, dir_factor: float,
                              grads: GridOutputGrads) -> None:
    """lumisphere_tv_grad_sparse (loss_kernel.cu:643-678): the lumisphere TV
    gradient of grid.sh_data at the listed cells into grads.grad_sh_out (and
    grads.mask_out).  basis_fn / basis_fn_u hold at least basis_dim float32
    values on the grid's device (the kernel reads the first basis_dim)."""
    g = _grid_struct(grid, need_sh=True, need_map=False)
    _check(rand_cells, "rand_cells", (torch.int32,))
    for name, t in (("basis_fn", basis_fn), ("basis_fn_u", basis_fn_u)):
        _check(t, name, (torch.float32,))
        if t.numel() < g.basis_dim:
            raise RuntimeError(f"{name} must hold basis_dim = {g.basis_dim} values")
        if t.device != grid.sh_data.device:
            raise RuntimeError(f"{name} must be on the grid's device")
    if basis_fn.dim() != 1:
        raise RuntimeError("basis_fn must have 1 dim")
    if grads.grad_sh_out is None:
        raise RuntimeError("grads.grad_sh_out must be defined")
    gr = _grads_struct(grads, g.rows, g.sh_data_dim)
    with _on_device(grid.sh_data):
        _call("svox2_lumisphere_tv_grad_sparse", ctypes.byref(g), _ptr(rand_cells),
              rand_cells.numel(), _ptr(basis_fn), _ptr(basis_fn_u), float(scale),
              float(ndc_coeffx), float(ndc_coeffy), float(dir_factor), ctypes.byref(gr),
              _stream(grid.sh_data))


def msi_tv_grad_sparse(links: torch.Tensor, msi: torch.Tensor, rand_cells: torch.Tensor,
                       mask_out: torch.Tensor, scale: float, scale_last: float,
                       grad_msi: torch.Tensor) -> None:
    """msi_tv_grad_sparse (loss_kernel.cu:606-641): total variation of an MSI
    background, links (2 reso, reso) int32, msi and grad_msi (rows, nlayers,
    channels) float32, rand_cells int32 flat indices over (2 reso, reso,
    nlayers); mask_out (rows, nlayers) bool, or empty for none."""
    links = _check(links, "links", (torch.int32,), 2)
    msi = _check(msi, "msi", (torch.float32,), 3)
    grad_msi = _check(grad_msi, "grad_msi", (torch.float32,), 3)
    rand_cells = _check(rand_cells, "rand_cells", (torch.int32,), 1)
    if links.size(0) != 2 * links.size(1):
        raise RuntimeError("links must be (2 reso, reso)")
    if grad_msi.shape != msi.shape:
        raise RuntimeError("grad_msi must have msi's shape")
    mask = None
    if mask_out is not None and mask_out.numel() > 0:
        mask = _check(mask_out, "mask_out", (torch.bool, torch.uint8))
        if mask.numel() != msi.size(0) * msi.size(1):
            raise RuntimeError("mask_out must be (rows, nlayers)")
    with _on_device(msi):
        _call("svox2_msi_tv_grad_sparse", _ptr(links), links.size(1), _ptr(msi), msi.size(0),
              msi.size(1), msi.size(2), _ptr(rand_cells), rand_cells.numel(), _ptr(mask),
              float(scale), float(scale_last), _ptr(grad_msi), _stream(msi))


def _indexer_args(indexer: torch.Tensor, rows: int):
    """(kind, pointer, length) of the reference's indexer: a 0-dim tensor
    selects every row, a bool/uint8 (rows,) tensor is a row mask, an int64
    tensor lists rows; None for an empty 1-d indexer (nothing to update)."""
    if indexer.dim() == 0:
        return INDEXER_ALL, None, 0
    if indexer.numel() == 0:
        return None
    if indexer.dtype in (torch.bool, torch.uint8):
        _check(indexer, "indexer", (torch.bool, torch.uint8), 1)
        if indexer.numel() != rows:
            raise RuntimeError("a mask indexer must have one entry per row")
        return INDEXER_MASK, _ptr(indexer), indexer.numel()
    _check(indexer, "indexer", (torch.int64,), 1)
    return INDEXER_INDEX, _ptr(indexer), indexer.numel()


def rmsprop_step(data: torch.Tensor, rms: torch.Tensor, grad: torch.Tensor,
                 indexer: torch.Tensor, beta: float, lr: float, epsilon: float, minval: float,
                 lr_last: float) -> None:
    """rmsprop_step (optim_kernel.cu:154-218).  ``indexer``: 0-dim tensor = all
    rows, bool/uint8 (rows,) mask, or int64 row indices; an empty 1-d indexer
    skips the update."""
    _check_rows(data, "data", _DATA_DTYPES)
    _check_rows(rms, "rms", (torch.float32,))
    _check_rows(grad, "grad", (torch.float32,))
    if rms.shape != data.shape or grad.shape != data.shape:
        raise RuntimeError("rms / grad must match data's shape")
    rows, dim = data.shape
    sel = _indexer_args(indexer, rows)
    if sel is None:
        return
    kind, iptr, n = sel
    with _on_device(data):
        _call("svox2_rmsprop_step", _ptr(data), _storage_of(data), _ptr(rms), _ptr(grad), rows,
              dim, _row_stride(data), _row_stride(rms), _row_stride(grad), kind, iptr, n,
              float(beta), float(lr), float(epsilon), float(minval), float(lr_last),
              _stream(data))


def sgd_step(data: torch.Tensor, grad: torch.Tensor, indexer: torch.Tensor, lr: float,
             lr_last: float) -> None:
    """sgd_step (optim_kernel.cu:218-266): data -= lr * grad over the indexed
    rows (lr_last on the last column), grad zeroed there."""
    _check_rows(data, "data", _DATA_DTYPES)
    _check_rows(grad, "grad", (torch.float32,))
    if grad.shape != data.shape:
        raise RuntimeError("grad must match data's shape")
    rows, dim = data.shape
    sel = _indexer_args(indexer, rows)
    if sel is None:
        return
    kind, iptr, n = sel
    with _on_device(data):
        _call("svox2_sgd_step", _ptr(data), _storage_of(data), _ptr(grad), rows, dim,
              _row_stride(data), _row_stride(grad), kind, iptr, n, float(lr), float(lr_last),
              _stream(data))


def rmsprop_step_grouped(data: torch.Tensor, rms: torch.Tensor, grad: torch.Tensor, split: int,
                         mask_a: Optional[torch.Tensor], mask_b: Optional[torch.Tensor],
                         beta_a: float, lr_a: float, beta_b: float, lr_b: float,
                         epsilon: float, minval: float = -1e9) -> None:
    """svox2_rmsprop_step_grouped: RMSprop of columns [0, split) with
    (mask_a, beta_a, lr_a) and [split, ncols) with (mask_b, beta_b, lr_b) in
    one pass; masks None = every row.  No reference counterpart."""
    _check_rows(data, "data", _DATA_DTYPES)
    _check_rows(rms, "rms", (torch.float32,))
    _check_rows(grad, "grad", (torch.float32,))
    if rms.shape != data.shape or grad.shape != data.shape:
        raise RuntimeError("rms / grad must match data's shape")
    rows, ncols = data.shape
    if (mask_a is None) != (mask_b is None):
        raise RuntimeError("give both group masks or neither")
    for m, name in ((mask_a, "mask_a"), (mask_b, "mask_b")):
        if m is not None:
            _check(m, name, (torch.bool, torch.uint8), 1)
            if m.numel() != rows:
                raise RuntimeError(f"{name} must have one entry per row")
    with _on_device(data):
        _call("svox2_rmsprop_step_grouped", _ptr(data), _storage_of(data), _ptr(rms), _ptr(grad),
              rows, ncols, _row_stride(data), _row_stride(rms), _row_stride(grad), int(split),
              _ptr(mask_a), _ptr(mask_b), float(beta_a), float(lr_a), float(beta_b),
              float(lr_b), float(epsilon), float(minval), _stream(data))


def accel_dist_prop(links: torch.Tensor) -> None:
    """accel_dist_prop (misc_kernel.cu:356-416), in place on ``links``."""
    _check(links, "links", (torch.int32,), 3)
    lib = load_library()
    size = (ctypes.c_int32 * 3)(*links.shape)
    nbytes = lib.svox2_accel_dist_prop_workspace_bytes(size)
    ws = torch.empty(max(int(nbytes), 4), dtype=torch.uint8, device=links.device)
    with _on_device(links):
        _call("svox2_accel_dist_prop", _ptr(links), size, _ptr(ws), ctypes.c_size_t(nbytes),
              _stream(links))
    key = (links.device.index, links.data_ptr())
    _ACCEL_EPOCH[key] = _ACCEL_EPOCH.get(key, 0) + 1


# ---------------------------------------------------------------------------
# Resampling pieces
# ---------------------------------------------------------------------------
def _size3(t: torch.Tensor):
    return (ctypes.c_int32 * 3)(*t.shape)


def dilate(grid: torch.Tensor) -> torch.Tensor:
    """dilate (misc_kernel.cu:339-354): a new bool/uint8 (X, Y, Z) mask, set
    where any cell of the 3x3x3 neighbourhood (clipped to the grid) is set."""
    _check(grid, "grid", (torch.bool, torch.uint8), 3)
    out = torch.empty_like(grid)
    with _on_device(grid):
        _call("svox2_dilate", _ptr(grid), _size3(grid), _ptr(out), _stream(grid))
    return out


def grid_weight_render(data: torch.Tensor, cam: CameraSpec, step_size: float,
                       stop_thresh: float, last_sample_opaque: bool, offset, scaling,
                       grid_weight_out: torch.Tensor) -> None:
    """grid_weight_render (misc_kernel.cu:418-445): raises grid_weight_out
    (float32, data's shape, non-negative) to the per-voxel maximum weight of
    the camera's rays through the dense density volume ``data``.  offset and
    scaling (3 values each, tensors on any device or sequences) map world to
    volume coordinates."""
    _check(data, "data", (torch.float32,), 3)
    _check(grid_weight_out, "grid_weight_out", (torch.float32,), 3)
    if grid_weight_out.shape != data.shape:
        raise RuntimeError("grid_weight_out must match data's shape")

    def vec(v):
        v = v.detach().double().cpu().flatten().tolist() if isinstance(v, torch.Tensor) else v
        if len(v) != 3:
            raise RuntimeError("offset / scaling must have 3 entries")
        return (ctypes.c_float * 3)(*[float(x) for x in v])

    with _on_device(data):
        _call("svox2_grid_weight_render", _ptr(data), _size3(data),
              ctypes.byref(_cam_struct(cam)), float(step_size), float(stop_thresh),
              int(bool(last_sample_opaque)), vec(offset), vec(scaling), _ptr(grid_weight_out),
              _stream(data))


def resample_density(grid: SparseGridSpec, axes: torch.Tensor, reso) -> torch.Tensor:
    """svox2_resample_density: (reso) float32 volume of the grid's density at
    the new cells' points (axes = [x coords | y coords | z coords], grid
    coordinates of the old grid, float32 on the grid's device)."""
    g = _grid_struct(grid, need_sh=False, need_map=False)
    _check(axes, "axes", (torch.float32,), 1)
    reso = [int(r) for r in reso]
    if axes.numel() != sum(reso):
        raise RuntimeError("axes must hold reso[0] + reso[1] + reso[2] values")
    out = torch.empty(reso, dtype=torch.float32, device=axes.device)
    with _on_device(axes):
        _call("svox2_resample_density", ctypes.byref(g), _ptr(axes), (ctypes.c_int32 * 3)(*reso),
              _ptr(out), _stream(axes))
    return out


def resample_rows(grid: SparseGridSpec, axes: torch.Tensor, reso, kept_cells: torch.Tensor,
                  density_volume: torch.Tensor, sh_out: torch.Tensor,
                  density_out: torch.Tensor) -> None:
    """svox2_resample_rows: fills the new grid's rows (sh_out (n, C),
    density_out (n, 1), float32 or bf16, row-strided) from the kept cells."""
    g = _grid_struct(grid, need_sh=True, need_map=False)
    _check(axes, "axes", (torch.float32,), 1)
    _check(kept_cells, "kept_cells", (torch.int32,), 1)
    _check(density_volume, "density_volume", (torch.float32,))
    _check_rows(sh_out, "sh_out", _DATA_DTYPES, g.sh_data_dim)
    _check_rows(density_out, "density_out", (sh_out.dtype,), 1)
    n = kept_cells.numel()
    if sh_out.size(0) != n or density_out.size(0) != n:
        raise RuntimeError("sh_out / density_out must have one row per kept cell")
    with _on_device(axes):
        _call("svox2_resample_rows", ctypes.byref(g), _ptr(axes),
              (ctypes.c_int32 * 3)(*[int(r) for r in reso]), _ptr(kept_cells), n,
              _ptr(density_volume), _ptr(sh_out), _row_stride(sh_out), _ptr(density_out),
              _row_stride(density_out), _storage_of(sh_out), _stream(axes))


def threshold_mask(values: torch.Tensor, thresh: float, keep_last_z: bool = False) -> torch.Tensor:
    """svox2_threshold_mask: bool volume values >= thresh (last z slice set
    when keep_last_z)."""
    _check(values, "values", (torch.float32,), 3)
    out = torch.empty(values.shape, dtype=torch.bool, device=values.device)
    with _on_device(values):
        _call("svox2_threshold_mask", _ptr(values), _size3(values), float(thresh),
              int(bool(keep_last_z)), _ptr(out), _stream(values))
    return out


def mask_to_links(mask: torch.Tensor, z_order: bool = False):
    """(links (X, Y, Z) int32, kept cells (n,) int32): the k-th set cell in
    raster (or Morton, z_order) order gets link k, the others -1; kept[k] is
    its flat cell index.  One host synchronisation reads the count."""
    _check(mask, "mask", (torch.bool, torch.uint8), 3)
    lib = load_library()
    size = _size3(mask)
    nbytes = int(lib.svox2_mask_compact_workspace_bytes(size))
    ws = torch.empty(nbytes // 4, dtype=torch.int32, device=mask.device)
    with _on_device(mask):
        _call("svox2_mask_compact_count", _ptr(mask), size, int(bool(z_order)), _ptr(ws),
              ctypes.c_size_t(nbytes), _stream(mask))
        n = int(ws[-1].item())
        links = torch.empty(mask.shape, dtype=torch.int32, device=mask.device)
        kept = torch.empty(n, dtype=torch.int32, device=mask.device)
        _call("svox2_mask_compact_assign", _ptr(mask), size, int(bool(z_order)), _ptr(ws),
              _ptr(links), _ptr(kept) if n > 0 else None, _stream(mask))
    return links, kept


# ---------------------------------------------------------------------------
# Multi-request tile rendering (the serving path; no reference counterpart)
# ---------------------------------------------------------------------------
def pack_structs(items) -> torch.Tensor:
    """Bytes of a list of ctypes structures as a CPU uint8 tensor."""
    if not items:
        return torch.empty(0, dtype=torch.uint8)
    blob = b"".join(bytes(memoryview(x)) for x in items)
    return torch.frombuffer(bytearray(blob), dtype=torch.uint8)


def grid_struct(grid: SparseGridSpec) -> _Grid:
    """Validated C descriptor of a grid (its cell map built and cached)."""
    return _grid_struct(grid)


def camera_struct(cam: CameraSpec) -> _Cam:
    return _cam_struct(cam)


def tile_struct(grid: int, camera: int, pixel0: int, n_pixels: int,
                rgb_out: torch.Tensor) -> _Tile:
    t = _Tile()
    t.grid, t.camera, t.pixel0, t.n_pixels = int(grid), int(camera), int(pixel0), int(n_pixels)
    t.rgb_out = rgb_out.data_ptr()
    return t


def volume_render_cuvol_tiles(grids: torch.Tensor, n_grids: int, storage: int, basis_dim: int,
                              cameras: torch.Tensor, n_cameras: int, opt: RenderOptions,
                              tiles: torch.Tensor, n_tiles: torch.Tensor, max_tiles: int,
                              tile_pixels: int) -> None:
    """svox2_volume_render_cuvol_tiles: grids / cameras / tiles are device
    uint8 tensors holding packed svox2_grid_t / svox2_camera_t / svox2_tile_t
    arrays, n_tiles a device int32 tensor of one element."""
    for t, name in ((grids, "grids"), (cameras, "cameras"), (tiles, "tiles")):
        _check(t, name, (torch.uint8,), 1)
    _check(n_tiles, "n_tiles", (torch.int32,))
    if grids.numel() < n_grids * ctypes.sizeof(_Grid) or \
            cameras.numel() < n_cameras * ctypes.sizeof(_Cam) or \
            tiles.numel() < max_tiles * ctypes.sizeof(_Tile):
        raise RuntimeError("grid / camera / tile tables are smaller than their counts")
    with _on_device(tiles):
        _call("svox2_volume_render_cuvol_tiles", _ptr(grids), int(n_grids), int(storage),
              int(basis_dim), _ptr(cameras), int(n_cameras), ctypes.byref(_opt_struct(opt)),
              _ptr(tiles), _ptr(n_tiles), int(max_tiles), int(tile_pixels), _stream(tiles))
