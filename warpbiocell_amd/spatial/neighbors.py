"""Neighbour search over the active cells (spatial/neighbors.py:16-37 of the reference).

``dim`` is the number of hash buckets per axis (the reference's ``wp.HashGrid(dim, dim, dim)``):
points anywhere in space hash into ``dim**3`` buckets of edge ``cell_size``. The device
workspace (bucket counts and starts, plus a bucket-ordered copy of positions and radii) is
allocated on first use for the population's capacity and reused by every build.
"""

from __future__ import annotations

import torch

from warpbiocell_amd import _lib


class NeighborGrid:
    def __init__(self, cell_size: float, dim: int = 128, device=None):
        if cell_size <= 0.0:
            raise ValueError("cell_size must be positive")
        if dim < 3:
            raise ValueError("dim must be at least 3")
        self.cell_size = float(cell_size)
        self.dim = int(dim)
        self.device = _lib.require_gpu(device)
        self._capacity = 0
        self._workspace: torch.Tensor | None = None
        self._built_count = 0

    @property
    def id(self) -> int:
        """Handle of the grid's device workspace (the kernels' view of the grid)."""
        return 0 if self._workspace is None else self._workspace.data_ptr()

    @property
    def workspace(self) -> torch.Tensor:
        if self._workspace is None:
            raise RuntimeError("the grid has not been reserved or built yet")
        return self._workspace

    @property
    def capacity(self) -> int:
        return self._capacity

    def reserve(self, num_points: int) -> None:
        if num_points <= self._capacity and self._workspace is not None:
            return
        nbytes = _lib.query("wbc_hashgrid_workspace_bytes", self.dim, int(num_points))
        if nbytes == 0:
            raise ValueError(f"unsupported grid: dim={self.dim}, points={num_points}")
        self._workspace = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        self._capacity = int(num_points)
        _lib.call("wbc_hashgrid_init", self._workspace.data_ptr(), nbytes, self.dim, self._capacity,
                  _lib.stream_of(self.device))

    def build(self, population) -> None:
        """Index the active prefix ``[0, count)``; query results are global slot indices."""
        if population.device != self.device:
            raise RuntimeError("population and NeighborGrid must live on the same device")
        self.reserve(population.capacity)
        _lib.call("wbc_hashgrid_build", population.position.data_ptr(), population.radius.data_ptr(),
                  population.count, self.cell_size, self.dim, self._capacity, self._workspace.data_ptr(),
                  self._workspace.numel(), _lib.stream_of(self.device))
        self._built_count = population.count

    def point_ids(self) -> torch.Tensor:
        """Slot ids of the last build in bucket order (``hash_grid_point_id`` order)."""
        return self._point_ids_view().clone()

    def _point_ids_view(self) -> torch.Tensor:
        """The same ids as a view of the workspace (valid until the next build)."""
        import ctypes

        pid, starts = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.call("wbc_hashgrid_point_ids", self.workspace.data_ptr(), self.workspace.numel(), self.dim,
                  self._capacity, ctypes.byref(pid), ctypes.byref(starts))
        offset = pid.value - self.workspace.data_ptr()
        return self.workspace[offset: offset + 4 * self._built_count].view(torch.int32)

    def bucket_starts(self) -> torch.Tensor:
        """The dim**3 + 1 bucket starts of the last build."""
        import ctypes

        pid, starts = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.call("wbc_hashgrid_point_ids", self.workspace.data_ptr(), self.workspace.numel(), self.dim,
                  self._capacity, ctypes.byref(pid), ctypes.byref(starts))
        offset = starts.value - self.workspace.data_ptr()
        nb = self.dim ** 3 + 1
        return self.workspace[offset: offset + 4 * nb].view(torch.int32).clone()
