"""Checkpoints in the reference's ``.npz`` layout (io/checkpoints.py:16-49), plus a restore.

``save_checkpoint`` writes the same keys, shapes and dtypes as the reference: every persistent
cell array of the live cells, ``count``, ``capacity``, ``seed``, ``time_h``, extra per-cell fields
as ``<name>_local``, the field, the consuming density, ``grid_origin`` and ``grid_dx``, the
tissue SDF as ``tissue_sdf`` when a region is given, so the
reference's loaders and post-processing (``load_checkpoint``, io/export.py) read it unchanged.
One key is added: ``neighbor_count``, the crowding of the last mechanics substep, which the next
step's lifecycle decisions read (simulation/simulator.py:16-18); with it a run restored by
``restore_population`` / ``restore_field`` continues bit for bit (the reference has no restart).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from .. import _lib
from ..cells.state import CellPopulation

CELL_ARRAYS = ("position", "radius", "cell_state", "cell_type", "age", "oxygen_local", "glucose_local", "fate_flags",
               "rng_state")
_DTYPES = {"position": np.float32, "radius": np.float32, "cell_state": np.int32, "cell_type": np.int32,
           "age": np.float32, "oxygen_local": np.float32, "glucose_local": np.float32, "fate_flags": np.int32,
           "rng_state": np.uint32}


def _host(t: torch.Tensor, n: int, dtype) -> np.ndarray:
    a = t[:n].cpu().numpy()
    return a.view(dtype) if a.dtype != dtype and a.dtype.itemsize == np.dtype(dtype).itemsize else a.astype(dtype)


def save_checkpoint(path: str | Path, population: CellPopulation, oxygen, time_h: float, region=None,
                    network=None) -> Path:
    """Write ``path`` (compressed ``.npz``) in the reference's layout; returns the path."""
    if network is not None:
        raise NotImplementedError("gene networks are not part of this package")
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    n = population.count
    arrays = {name: _host(getattr(population, name), n, _DTYPES[name]).copy() for name in CELL_ARRAYS}
    arrays["count"] = np.array(n)
    arrays["capacity"] = np.array(population.capacity)
    arrays["seed"] = np.array(population.seed)
    arrays["time_h"] = np.array(time_h)
    for name, values in population.extra_local.items():
        arrays[f"{name}_local"] = values[:n].cpu().numpy().copy()
    if oxygen is not None:
        fields = getattr(oxygen, "species_fields", None) or {"oxygen": oxygen}
        for name, species in fields.items():
            arrays["field" if name == "oxygen" else f"field_{name}"] = species.numpy()
        arrays["density"] = fields["oxygen"].density.cpu().numpy()
        arrays["grid_origin"] = np.array(oxygen.field.geometry.origin)
        arrays["grid_dx"] = np.array(oxygen.field.geometry.dx)
    if region is not None:
        arrays["tissue_sdf"] = region.sdf_numpy()
        arrays["grid_origin"] = np.array(region.geometry.origin)
        arrays["grid_dx"] = np.array(region.geometry.dx)
    arrays["neighbor_count"] = _host(population.neighbor_count, n, np.int32).copy()
    np.savez_compressed(path, **arrays)
    return path


def load_checkpoint(path: str | Path) -> dict[str, np.ndarray]:
    """Every array of a checkpoint, by key."""
    with np.load(Path(path)) as data:
        return {key: data[key] for key in data.files}


def restore_population(data: dict[str, np.ndarray], device=None) -> CellPopulation:
    """A population in the saved state: the saved cell arrays in slots [0, count), every other
    slot as a fresh population of the same capacity and seed leaves it (its RNG stream
    ``rand_init(seed, slot)`` included), and the saved crowding when the checkpoint has it."""
    n, capacity, seed = int(data["count"]), int(data["capacity"]), int(data["seed"])
    pop = CellPopulation.from_numpy(data["position"], data["radius"], capacity=capacity, device=device, seed=seed,
                                    states=data["cell_state"])
    for name in ("cell_type", "age", "oxygen_local", "glucose_local", "fate_flags", "rng_state"):
        target = getattr(pop, name)
        values = np.ascontiguousarray(data[name])
        if name == "rng_state":
            values = values.view(np.int32)
        target[:n] = torch.from_numpy(values).to(device=pop.device, dtype=target.dtype)
    if "neighbor_count" in data:
        pop.neighbor_count[:n] = torch.from_numpy(np.ascontiguousarray(data["neighbor_count"], np.int32)).to(pop.device)
    for key, values in data.items():
        if key.endswith("_local") and key not in ("oxygen_local", "glucose_local"):
            pop.local_field(key[: -len("_local")])[:n] = torch.from_numpy(np.ascontiguousarray(values)).to(pop.device)
    torch.cuda.synchronize(pop.device)
    return pop


def restore_field(data: dict[str, np.ndarray], oxygen) -> None:
    """Put the saved field(s) (the next solves' warm starts) into ``oxygen`` (an OxygenField or a
    MetabolicFields: every species by its ``field_<name>`` key); the grid must match."""
    geom = oxygen.field.geometry
    if tuple(data["field"].shape) != tuple(geom.shape) or not np.allclose(data["grid_origin"], geom.origin) \
            or float(data["grid_dx"]) != float(geom.dx):
        raise ValueError("the checkpoint's grid does not match the field's")
    fields = getattr(oxygen, "species_fields", None) or {"oxygen": oxygen}
    for name, species in fields.items():
        species.field.assign(data["field" if name == "oxygen" else f"field_{name}"])
    torch.cuda.synchronize(oxygen.device)


def restore_region(data: dict[str, np.ndarray], device=None, name: str = "region"):
    """The TissueRegion a checkpoint saved (``tissue_sdf`` on ``grid_origin`` / ``grid_dx``), or
    None when it holds none."""
    if "tissue_sdf" not in data:
        return None
    from ..fields.scalar_field import GridGeometry
    from ..geometry.region import TissueRegion

    sdf = np.asarray(data["tissue_sdf"], dtype=np.float32)
    geom = GridGeometry(origin=tuple(float(o) for o in data["grid_origin"]), dx=float(data["grid_dx"]),
                        shape=tuple(int(d) for d in sdf.shape))
    return TissueRegion(geom, sdf, device=device, name=name)
