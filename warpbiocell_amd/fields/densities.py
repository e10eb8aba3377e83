"""Per-state cell number densities on the grid (fields/densities.py:14-56 of the reference).

The deposit accumulates int64 fixed-point trilinear weights (2^24 per cell), so the result
does not depend on the order in which the atomics land. ``density`` (per state,
``(4, nx, ny, nz)`` float32) is produced on demand from the accumulator: a species that owns
its densities asks for its weighted sum directly (``deposit_combined``) and never writes the
four per-state layers unless something reads them. A bit set on the device records the layers
(states) deposits have written; the others stay zero and are neither zeroed nor read, so a
population in one state touches a quarter of the accumulator.
"""

from __future__ import annotations

import numpy as np
import torch

from warpbiocell_amd import _lib
from warpbiocell_amd.cells.model import NUM_STATES
from warpbiocell_amd.fields.scalar_field import GridGeometry


class CellDensities:
    def __init__(self, geometry: GridGeometry, device=None):
        if NUM_STATES != 4:
            raise RuntimeError("combine_densities assumes four cell states")
        self.geometry = geometry
        self.device = _lib.require_gpu(device)
        shape = (NUM_STATES, *geometry.shape)
        self._accumulator = torch.zeros(shape, dtype=torch.int64, device=self.device)
        # Bit set of the accumulator layers (states) a deposit has written; the others stay zero
        # and are neither zeroed nor read.
        self._layers = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._density = torch.zeros(shape, dtype=torch.float32, device=self.device)
        self._density_stale = False

    @property
    def accumulator(self) -> torch.Tensor:
        return self._accumulator

    def _scalars(self):
        g = self.geometry
        return g.origin, 1.0 / g.dx, 1.0 / g.voxel_volume

    def _accumulate(self, population) -> None:
        if population.device != self.device:
            raise RuntimeError("population and CellDensities must live on the same device")
        stream = _lib.stream_of(self.device)
        nx, ny, nz = self.geometry.shape
        origin, inv_dx, _ = self._scalars()
        # A permutation of the live slots (deposit_order) gives the same accumulator, faster.
        order = population.deposit_order() if population.count > 0 else None
        _lib.call("wbc_deposit_accumulate_i64", population.position.data_ptr(), population.cell_state.data_ptr(),
                  _lib.ptr(order), population.count, origin[0], origin[1], origin[2], inv_dx, nx, ny, nz,
                  self._accumulator.data_ptr(), self._layers.data_ptr(), stream)

    def deposit(self, population) -> None:
        """Zero the accumulator, deposit every active cell in its state's layer, convert."""
        self._accumulate(population)
        self._finalize(None, None, per_state=True)

    def deposit_combined(self, population, weights, out: torch.Tensor) -> None:
        """Deposit and write ``sum_k weights[k] * density[k]`` into ``out`` in one grid pass; the
        per-state layers are materialised later only if ``density`` is read."""
        self._accumulate(population)
        self._finalize(weights, out, per_state=False)

    def _finalize(self, weights, out, per_state: bool) -> None:
        w = [0.0] * 4 if weights is None else [float(x) for x in weights]
        if len(w) != NUM_STATES:
            raise ValueError(f"need {NUM_STATES} weights, got {len(w)}")
        _, _, inv_vol = self._scalars()
        _lib.call("wbc_density_finalize", self._accumulator.data_ptr(), self.geometry.node_count, inv_vol, *w,
                  _lib.ptr(out), self._density.data_ptr() if per_state else None, self._layers.data_ptr(),
                  _lib.stream_of(self.device))
        self._density_stale = not per_state

    @property
    def density(self) -> torch.Tensor:
        if self._density_stale:
            self._finalize(None, None, per_state=True)
        return self._density

    def combine(self, weights, out: torch.Tensor) -> None:
        """``out = sum_k weights[k] * density[k]`` (weights indexed by CellState)."""
        w = [float(x) for x in weights]
        if len(w) != NUM_STATES:
            raise ValueError(f"need {NUM_STATES} weights, got {len(w)}")
        if self._density_stale:
            _, _, inv_vol = self._scalars()
            _lib.call("wbc_density_finalize", self._accumulator.data_ptr(), self.geometry.node_count, inv_vol, *w,
                      out.data_ptr(), None, self._layers.data_ptr(), _lib.stream_of(self.device))
            return
        _lib.call("wbc_combine_densities", self._density.data_ptr(), self.geometry.node_count, *w, out.data_ptr(),
                  _lib.stream_of(self.device))

    def numpy(self) -> np.ndarray:
        return self.density.cpu().numpy().copy()

    def total_numpy(self) -> np.ndarray:
        return self.density.cpu().numpy().sum(axis=0)
