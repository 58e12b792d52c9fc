"""Host-side helpers of the svox2 interface: integer roots, Morton (Z-order)
codes for the initial row layout, the real SH basis, and the NDC ray warp.

Counterparts: svox2/utils.py:9-27 (inthroot, isqrt, is_pow2), :45-79 (Morton),
:82-162 (SH constants, eval_sh_bases), :576-596 (convert_to_ndc).
"""
from __future__ import annotations

from typing import Optional

import torch

MAX_SH_BASIS = 10

SH_C0 = 0.28209479177387814
SH_C1 = 0.4886025119029199
SH_C2 = (1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
         -1.0925484305920792, 0.5462742152960396)


def inthroot(x: int, n: int) -> Optional[int]:
    """Exact integer n-th root of x, or None."""
    if x <= 0:
        return None
    r = int(round(x ** (1.0 / n)))
    for c in (r - 1, r, r + 1):
        if c > 0 and c ** n == x:
            return c
    return None


def isqrt(x: int) -> Optional[int]:
    return inthroot(x, 2)


def is_pow2(x: int) -> bool:
    return x > 0 and (x & (x - 1)) == 0


def _spread3(v: torch.Tensor) -> torch.Tensor:
    """Inserts two zero bits between the low 10 bits of v (int64)."""
    v = v & 0x3FF
    v = (v | (v << 16)) & 0x030000FF
    v = (v | (v << 8)) & 0x0300F00F
    v = (v | (v << 4)) & 0x030C30C3
    v = (v | (v << 2)) & 0x09249249
    return v


def _compact3(v: torch.Tensor) -> torch.Tensor:
    v = v & 0x09249249
    v = (v | (v >> 2)) & 0x030C30C3
    v = (v | (v >> 4)) & 0x0300F00F
    v = (v | (v >> 8)) & 0x030000FF
    v = (v | (v >> 16)) & 0x000003FF
    return v


def morton_code_3(x: torch.Tensor, y: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
    """Z-order code with x in the most significant position of each bit triple:
    (spread(x) << 2) | (spread(y) << 1) | spread(z)."""
    return (_spread3(x) << 2) | (_spread3(y) << 1) | _spread3(z)


def inv_morton_code_3(code: torch.Tensor):
    return _compact3(code >> 2), _compact3(code >> 1), _compact3(code)


def gen_morton(D: int, device="cpu", dtype=torch.long) -> torch.Tensor:
    """(D, D, D) tensor of Morton codes of every cell (D a power of two <= 1024)."""
    if not is_pow2(D) or D > 1024:
        raise ValueError("Morton code requires a power-of-2 resolution <= 1024")
    a = torch.arange(D, device=device, dtype=torch.long)
    X, Y, Z = torch.meshgrid(a, a, a, indexing="ij")
    return morton_code_3(X, Y, Z).to(dtype)


def eval_sh_bases(basis_dim: int, dirs: torch.Tensor) -> torch.Tensor:
    """Real SH basis (degree <= 2, basis_dim in {1, 4, 9}) at unit dirs (..., 3)."""
    if basis_dim not in (1, 4, 9):
        raise ValueError("basis_dim must be 1, 4 or 9")
    out = torch.empty((*dirs.shape[:-1], basis_dim), dtype=dirs.dtype, device=dirs.device)
    out[..., 0] = SH_C0
    if basis_dim > 1:
        x, y, z = dirs.unbind(-1)
        out[..., 1] = -SH_C1 * y
        out[..., 2] = SH_C1 * z
        out[..., 3] = -SH_C1 * x
        if basis_dim > 4:
            out[..., 4] = SH_C2[0] * (x * y)
            out[..., 5] = SH_C2[1] * (y * z)
            out[..., 6] = SH_C2[2] * (2.0 * z * z - x * x - y * y)
            out[..., 7] = SH_C2[3] * (x * z)
            out[..., 8] = SH_C2[4] * (x * x - y * y)
    return out


def convert_to_ndc(origins: torch.Tensor, directions: torch.Tensor, ndc_coeffs,
                   near: float = 1.0):
    """Warps rays into normalised device coordinates of a forward-facing scene:
    origins are moved to the z = near plane, then projected."""
    t = (near - origins[..., 2]) / directions[..., 2]
    origins = origins + t[..., None] * directions
    dx, dy, dz = directions.unbind(-1)
    ox, oy, oz = origins.unbind(-1)
    o = torch.stack([ndc_coeffs[0] * (ox / oz), ndc_coeffs[1] * (oy / oz), 1 - 2 * near / oz], -1)
    d = torch.stack([ndc_coeffs[0] * (dx / dz - ox / oz), ndc_coeffs[1] * (dy / dz - oy / oz),
                     2 * near / oz], -1)
    return o, d
