"""A tissue region: a signed-distance field on a regular grid, negative inside
(geometry/region.py:18-78 of the reference).

The SDF and its node gradient live on the GPU, as float32 (nx, ny, nz) and (nx, ny, nz, 3)
tensors; the gradient is computed there once, at construction, by ``wbc_region_gradient``
(numpy.gradient's differences in float64, rounded to float32). The contact pass samples both
per cell (``cells.mechanics``); ``sample_at`` samples the SDF alone. The host-side queries
(``sample``, ``contains``, ``volume``) work on a host copy of the SDF in float64, as the
reference's do.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from warpbiocell_amd import _lib
from warpbiocell_amd.fields.scalar_field import GridGeometry
from warpbiocell_amd.geometry.shapes import evaluate_on_grid


class RegionStruct(ctypes.Structure):
    """``wbc_region_t``."""

    _fields_ = [("sdf", ctypes.c_void_p), ("gradient", ctypes.c_void_p), ("nx", ctypes.c_int32),
                ("ny", ctypes.c_int32), ("nz", ctypes.c_int32), ("origin", ctypes.c_float * 3),
                ("inv_dx", ctypes.c_float)]


def node_gradient(values, dx: float, device=None) -> torch.Tensor:
    """Central differences inside, one-sided at the faces, of a float32 (nx, ny, nz) grid with
    spacing ``dx``: a float32 (nx, ny, nz, 3) tensor on the GPU."""
    dev = _lib.require_gpu(device)
    sdf = torch.as_tensor(np.ascontiguousarray(values, dtype=np.float32)).to(dev) if not torch.is_tensor(values) \
        else values.to(device=dev, dtype=torch.float32).contiguous()
    if sdf.dim() != 3 or min(sdf.shape) < 2:
        raise ValueError("node_gradient needs a 3-D grid with at least 2 nodes per axis")
    out = torch.empty((*sdf.shape, 3), dtype=torch.float32, device=dev)
    _lib.call("wbc_region_gradient", sdf.data_ptr(), *sdf.shape, float(dx), out.data_ptr(), _lib.stream_of(dev))
    return out


class TissueRegion:
    def __init__(self, geometry: GridGeometry, sdf: np.ndarray, device=None, name: str = "region"):
        sdf = np.ascontiguousarray(sdf, dtype=np.float32)
        if sdf.shape != tuple(geometry.shape):
            raise ValueError(f"sdf shape {sdf.shape} does not match the grid {geometry.shape}")
        self.geometry = geometry
        self.name = name
        self.device = _lib.require_gpu(device)
        self._sdf_np = sdf
        self.sdf = torch.from_numpy(sdf).to(self.device)
        self.gradient = node_gradient(self.sdf, geometry.dx, self.device)
        self._struct = RegionStruct(self.sdf.data_ptr(), self.gradient.data_ptr(), *sdf.shape,
                                    (ctypes.c_float * 3)(*geometry.origin), 1.0 / geometry.dx)

    @classmethod
    def from_shape(cls, shape, geometry: GridGeometry, device=None) -> TissueRegion:
        return cls(geometry, evaluate_on_grid(shape, geometry), device=device, name=type(shape).__name__.lower())

    @property
    def struct_ptr(self) -> int:
        """Address of the ``wbc_region_t`` describing this region (valid while the region lives)."""
        return ctypes.addressof(self._struct)

    # ---- host-side queries ----------------------------------------------------------------------

    def sdf_numpy(self) -> np.ndarray:
        return self._sdf_np.copy()

    def inside_mask(self) -> np.ndarray:
        """Nodes strictly inside the tissue."""
        return self._sdf_np < 0.0

    def volume(self) -> float:
        """Tissue volume [um^3] counted in whole voxels."""
        return float(self.inside_mask().sum()) * self.geometry.voxel_volume

    def sample(self, positions: np.ndarray) -> np.ndarray:
        """Trilinear SDF at arbitrary points [um] in float64 on the host; points outside the grid
        are clamped to it (geometry/region.py:53-67)."""
        geom = self.geometry
        shape = np.array(geom.shape)
        u = (np.asarray(positions, dtype=np.float64) - np.asarray(geom.origin)) / geom.dx
        u = np.clip(u, 0.0, shape - 1.0)
        base = np.minimum(np.floor(u).astype(np.int64), shape - 2)
        fx, fy, fz = (u - base).T
        i, j, k = base.T
        v = self._sdf_np.astype(np.float64)
        c00 = v[i, j, k] * (1 - fx) + v[i + 1, j, k] * fx
        c10 = v[i, j + 1, k] * (1 - fx) + v[i + 1, j + 1, k] * fx
        c01 = v[i, j, k + 1] * (1 - fx) + v[i + 1, j, k + 1] * fx
        c11 = v[i, j + 1, k + 1] * (1 - fx) + v[i + 1, j + 1, k + 1] * fx
        return (c00 * (1 - fy) + c10 * fy) * (1 - fz) + (c01 * (1 - fy) + c11 * fy) * fz

    def contains(self, positions: np.ndarray, margin: float = 0.0) -> np.ndarray:
        """True where a point lies at least ``margin`` [um] inside the surface."""
        return self.sample(positions) < -margin

    # ---- device-side queries --------------------------------------------------------------------

    def sample_at(self, position: torch.Tensor, count: int, out: torch.Tensor) -> None:
        """``out[:count]`` = the trilinear SDF at ``position[:count]`` (float32, on the GPU)."""
        if count < 0 or count > position.shape[0] or count > out.shape[0]:
            raise ValueError("count exceeds the position or output arrays")
        if position.dtype != torch.float32 or out.dtype != torch.float32:
            raise TypeError("sample_at takes float32 tensors")
        if not (position.is_contiguous() and out.is_contiguous()):
            raise ValueError("sample_at takes contiguous tensors")
        if position.device != self.device or out.device != self.device:
            raise RuntimeError("sample_at: tensors must live on the region's device")
        if count > 0:
            _lib.call("wbc_sample_sdf", self.struct_ptr, position.data_ptr(), count, out.data_ptr(),
                      _lib.stream_of(self.device))
