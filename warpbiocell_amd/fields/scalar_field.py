"""Regular 3-D scalar grid with per-axis boundary conditions (fields/scalar_field.py:1-150).

Node ``(i, j, k)`` sits at ``origin + dx * (i, j, k)``; values are a ``(nx, ny, nz)`` float32
tensor, k contiguous. Dirichlet faces hold ``boundary_value`` and are never written by the
solvers; Neumann faces mirror (``f[-1] = f[1]``). ``fixed`` pins further nodes to the boundary
value; the solvers receive no mask at all while nothing is pinned.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from enum import IntEnum

import numpy as np
import torch

from warpbiocell_amd import _lib


class Boundary(IntEnum):
    DIRICHLET = 0
    NEUMANN = 1


@dataclass(frozen=True)
class GridGeometry:
    origin: tuple[float, float, float]
    dx: float
    shape: tuple[int, int, int]

    def __post_init__(self):
        if self.dx <= 0.0:
            raise ValueError("dx must be positive")
        if len(self.shape) != 3 or min(self.shape) < 3:
            raise ValueError("shape needs at least 3 nodes per axis")

    @classmethod
    def centered_cube(cls, side: float, dx: float) -> GridGeometry:
        """Cube of edge ``side`` centred on the origin, rounded up to whole voxels."""
        n = int(np.ceil(side / dx)) + 1
        half = 0.5 * (n - 1) * dx
        return cls(origin=(-half, -half, -half), dx=dx, shape=(n, n, n))

    @property
    def node_count(self) -> int:
        return int(np.prod(self.shape))

    @property
    def voxel_volume(self) -> float:
        return self.dx ** 3

    @property
    def upper(self) -> tuple[float, float, float]:
        return tuple(self.origin[a] + self.dx * (self.shape[a] - 1) for a in range(3))

    def node_coordinates(self):
        return tuple(self.origin[a] + self.dx * np.arange(self.shape[a]) for a in range(3))

    def contains(self, positions: np.ndarray) -> np.ndarray:
        p = np.asarray(positions)
        return np.all((p >= np.asarray(self.origin)) & (p <= np.asarray(self.upper)), axis=1)


@dataclass
class ScalarField:
    geometry: GridGeometry
    boundary: tuple[Boundary, Boundary, Boundary]
    boundary_value: float
    device: torch.device
    values: torch.Tensor  # (nx, ny, nz) float32
    fixed: torch.Tensor  # (nx, ny, nz) int32, 1 where pinned
    has_pinned: bool = False
    _scratch: dict = field(default_factory=dict, repr=False)

    @classmethod
    def create(cls, geometry: GridGeometry, boundary_value: float,
               boundary=(Boundary.DIRICHLET,) * 3, initial_value: float | None = None, device=None) -> ScalarField:
        dev = _lib.require_gpu(device)
        fill = boundary_value if initial_value is None else initial_value
        values = torch.full(geometry.shape, float(fill), dtype=torch.float32, device=dev)
        fixed = torch.zeros(geometry.shape, dtype=torch.int32, device=dev)
        f = cls(geometry=geometry, boundary=tuple(Boundary(b) for b in boundary), boundary_value=float(boundary_value),
                device=dev, values=values, fixed=fixed)
        f.apply_dirichlet()
        return f

    def set_fixed(self, mask: np.ndarray | None) -> None:
        if mask is None:
            self.fixed.zero_()
            self.has_pinned = False
            return
        mask = np.asarray(mask, dtype=bool)
        if mask.shape != tuple(self.geometry.shape):
            raise ValueError(f"mask shape {mask.shape} does not match the grid {self.geometry.shape}")
        self.fixed.copy_(torch.from_numpy(mask.astype(np.int32)))
        self.has_pinned = bool(mask.any())
        self.apply_dirichlet()

    def fixed_mask(self) -> np.ndarray:
        return self.fixed.cpu().numpy().astype(bool)

    @property
    def fixed_or_none(self) -> torch.Tensor | None:
        return self.fixed if self.has_pinned else None

    def fixed_zero_like(self) -> torch.Tensor:
        z = self._scratch.get("zero")
        if z is None:
            z = torch.zeros(self.geometry.shape, dtype=torch.float32, device=self.device)
            self._scratch["zero"] = z
        return z

    def solver_workspace(self, max_sweeps: int, check_every: int):
        """Device workspace and pinned host residual slots of the solve loop (allocated once)."""
        key = ("solve", max_sweeps, check_every)
        ws = self._scratch.get(key)
        if ws is None:
            nbytes = _lib.query("wbc_solve_workspace_bytes", int(max_sweeps), int(check_every), *self.geometry.shape)
            chunks = -(-int(max_sweeps) // int(check_every))
            ws = (torch.empty(nbytes, dtype=torch.uint8, device=self.device),
                  torch.zeros(chunks, dtype=torch.float32, pin_memory=True))
            self._scratch[key] = ws
        return ws

    @property
    def boundary_flags(self) -> tuple[int, int, int]:
        return tuple(int(b) for b in self.boundary)

    def numpy(self) -> np.ndarray:
        return self.values.cpu().numpy().copy()

    def assign(self, array) -> None:
        array = np.ascontiguousarray(array, dtype=np.float32)
        if array.shape != tuple(self.geometry.shape):
            raise ValueError(f"expected shape {self.geometry.shape}, got {array.shape}")
        self.values.copy_(torch.from_numpy(array))
        self.apply_dirichlet()

    def apply_dirichlet(self) -> None:
        pinned = torch.from_numpy(~self.interior_mask()).to(self.device)
        self.values.masked_fill_(pinned, self.boundary_value)

    def interior_mask(self) -> np.ndarray:
        mask = np.ones(self.geometry.shape, dtype=bool)
        for axis, b in enumerate(self.boundary):
            if b == Boundary.DIRICHLET:
                idx = [slice(None)] * 3
                idx[axis] = 0
                mask[tuple(idx)] = False
                idx[axis] = -1
                mask[tuple(idx)] = False
        if self.has_pinned:
            mask &= ~self.fixed_mask()
        return mask
