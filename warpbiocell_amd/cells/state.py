"""Structure-of-arrays cell storage with preallocated capacity (cells/state.py:31-159).

Slots ``[0, count)`` hold cells (alive or dead), ``[count, capacity)`` are free; a slot is a
cell's id for life. Arrays are torch tensors on a ROCm device: ``position`` and ``velocity`` are
``(capacity, 3)`` float32 rows (the byte layout of a Warp ``vec3`` array), ``rng_state`` holds
each slot's uint32 RNG state in an int32 tensor.

Every slot's RNG stream is seeded at creation, ``rand_init(seed, slot)``, so results never
depend on the capacity (cells/state.py:108-110).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from warpbiocell_amd import _lib
from warpbiocell_amd.cells.model import CellState, CellType


@dataclass
class CellPopulation:
    capacity: int
    count: int
    seed: int
    device: torch.device
    position: torch.Tensor  # (capacity, 3) float32, um
    radius: torch.Tensor  # float32, um
    cell_state: torch.Tensor  # int32, CellState
    cell_type: torch.Tensor  # int32, CellType
    age: torch.Tensor  # float32, h
    oxygen_local: torch.Tensor  # float32, mmHg at the cell centre
    glucose_local: torch.Tensor  # float32, mM at the cell centre
    rng_state: torch.Tensor  # int32 storage of uint32 stream states
    velocity: torch.Tensor  # (capacity, 3) float32, um/h
    neighbor_count: torch.Tensor  # int32
    divide_flag: torch.Tensor  # int32
    division_offset: torch.Tensor  # int32, inclusive prefix sum of divide_flag
    fate_flags: torch.Tensor  # int32
    extra_local: dict = field(default_factory=dict)
    _scan_workspace: torch.Tensor | None = None
    _deposit_order: torch.Tensor | None = None  # int32 (capacity,): a permutation of [0, count)
    _deposit_order_count: int = -1

    @classmethod
    def from_numpy(cls, positions: np.ndarray, radii: np.ndarray, capacity: int | None = None, device=None,
                   seed: int = 0, states: np.ndarray | None = None) -> CellPopulation:
        positions = np.ascontiguousarray(positions, dtype=np.float32)
        radii = np.ascontiguousarray(radii, dtype=np.float32)
        if positions.ndim != 2 or positions.shape[1] != 3:
            raise ValueError(f"positions must have shape (n, 3), got {positions.shape}")
        count = positions.shape[0]
        if radii.shape != (count,):
            raise ValueError(f"radii must have shape ({count},), got {radii.shape}")
        capacity = count if capacity is None else int(capacity)
        if capacity < count:
            raise ValueError(f"capacity {capacity} smaller than initial cell count {count}")
        if states is None:
            states = np.full(count, int(CellState.PROLIFERATIVE), dtype=np.int32)
        states = np.ascontiguousarray(states, dtype=np.int32)
        if states.shape != (count,):
            raise ValueError(f"states must have shape ({count},), got {states.shape}")
        dev = _lib.require_gpu(device)

        def padded(values, dtype, fill=0):
            buf = np.full((capacity,) + values.shape[1:], fill, dtype=dtype)
            buf[:count] = values
            return torch.from_numpy(buf).to(dev)

        def zeros(*shape, dtype=torch.float32):
            return torch.zeros(shape, dtype=dtype, device=dev)

        pop = cls(
            capacity=capacity,
            count=count,
            seed=int(seed),
            device=dev,
            position=padded(positions, np.float32),
            radius=padded(radii, np.float32),
            cell_state=padded(states, np.int32, fill=int(CellState.DEAD)),
            cell_type=torch.full((capacity,), int(CellType.TUMOR), dtype=torch.int32, device=dev),
            age=zeros(capacity),
            oxygen_local=zeros(capacity),
            glucose_local=zeros(capacity),
            rng_state=zeros(capacity, dtype=torch.int32),
            velocity=zeros(capacity, 3),
            neighbor_count=zeros(capacity, dtype=torch.int32),
            divide_flag=zeros(capacity, dtype=torch.int32),
            division_offset=zeros(capacity, dtype=torch.int32),
            fate_flags=zeros(capacity, dtype=torch.int32),
        )
        _lib.call("wbc_init_rng", pop.rng_state.data_ptr(), capacity, int(seed), _lib.stream_of(dev))
        return pop

    @property
    def stream(self) -> int:
        return _lib.stream_of(self.device)

    def record_deposit_order(self, order: torch.Tensor) -> None:
        """Keep a copy of ``order`` (int32, a permutation of the live slots [0, count), such as the
        hash grid's bucket order after a contact relaxation) as the order the deposit visits cells
        in. Any permutation gives the same int64 accumulator; a spatially coherent one halves the
        deposit kernel's time."""
        n = self.count
        if order.dtype != torch.int32 or order.numel() != n or order.device != self.device:
            raise ValueError(f"the order must be {n} int32 values on {self.device}")
        if self._deposit_order is None:
            self._deposit_order = torch.empty(self.capacity, dtype=torch.int32, device=self.device)
        self._deposit_order[:n].copy_(order)
        self._deposit_order_count = n

    def forget_deposit_order(self) -> None:
        """Drop the recorded order (its cells were moved or replaced by another path)."""
        self._deposit_order_count = -1

    def deposit_order(self) -> torch.Tensor | None:
        """The recorded order while it covers exactly the live slots (no cells added since), else None."""
        if self.count > 0 and self._deposit_order_count == self.count:
            return self._deposit_order
        return None

    def scan_workspace(self) -> torch.Tensor:
        """Tile sums of the division-slot scan, sized for the capacity (allocated once)."""
        if self._scan_workspace is None:
            nbytes = _lib.query("wbc_scan_workspace_bytes", max(self.capacity, 1))
            self._scan_workspace = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        return self._scan_workspace

    # Views over the active prefix (no copies).
    @property
    def active_position(self) -> torch.Tensor:
        return self.position[: self.count]

    @property
    def active_radius(self) -> torch.Tensor:
        return self.radius[: self.count]

    def _host(self, t: torch.Tensor) -> np.ndarray:
        return t[: self.count].cpu().numpy().copy()

    def positions_numpy(self) -> np.ndarray:
        return self._host(self.position)

    def radii_numpy(self) -> np.ndarray:
        return self._host(self.radius)

    def states_numpy(self) -> np.ndarray:
        return self._host(self.cell_state)

    def ages_numpy(self) -> np.ndarray:
        return self._host(self.age)

    def oxygen_numpy(self) -> np.ndarray:
        return self._host(self.oxygen_local)

    def glucose_numpy(self) -> np.ndarray:
        return self._host(self.glucose_local)

    def rng_numpy(self) -> np.ndarray:
        return self.rng_state.cpu().numpy().view(np.uint32).copy()

    def local_field(self, name: str) -> torch.Tensor:
        if name == "oxygen":
            return self.oxygen_local
        if name == "glucose":
            return self.glucose_local
        if name not in self.extra_local:
            self.extra_local[name] = torch.zeros(self.capacity, dtype=torch.float32, device=self.device)
        return self.extra_local[name]

    def local_numpy(self, name: str) -> np.ndarray:
        return self._host(self.local_field(name))

    def neighbor_counts_numpy(self) -> np.ndarray:
        return self._host(self.neighbor_count)

    def fate_flags_numpy(self) -> np.ndarray:
        return self._host(self.fate_flags)
