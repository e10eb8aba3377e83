"""Population-level metrics (reference: metrics/population.py:13-135), computed on the device.

The reference copies the whole position, radius, state and oxygen arrays and the field to the host
and reduces them with numpy. Here the HIP library reduces them where they live
(``wbc_cell_moments``, ``wbc_centre_distances``, ``wbc_select_rank``, ``wbc_grid_moments``,
``wbc_radial_profile``) and a few numbers come back. Same names, arguments and dictionary keys.

Numbers: counts and minima are exact; sums are taken in double (in a fixed order, so they repeat
bit for bit), where the reference's numpy means accumulate in float32, so means agree to float32
rounding. The 99th-percentile centre distance applies numpy's float32 linear interpolation to the
two order statistics the device selects; the centre is the double-precision mean rounded to
float32 (numpy sums the float32 rows in float32), so distances agree to float32 rounding.
"""

from __future__ import annotations

import ctypes
from ctypes import c_double, c_float, c_int32

import numpy as np
import torch

from .. import _lib
from ..cells.model import CellState
from ..cells.state import CellPopulation


class _CellMoments(ctypes.Structure):
    _fields_ = [("sum_x", c_double), ("sum_y", c_double), ("sum_z", c_double), ("sum_radius", c_double),
                ("sum_oxygen_alive", c_double), ("min_oxygen_alive", c_float), ("count", c_int32),
                ("alive", c_int32), ("states", c_int32 * 4)]


class _GridMoments(ctypes.Structure):
    _fields_ = [("sum", c_double), ("min", c_float), ("count", c_int32)]


def _workspace(population: CellPopulation, n_bins: int = 0) -> torch.Tensor:
    size = _lib.query("wbc_metrics_workspace_bytes", population.count, n_bins)
    if size == 0:
        raise ValueError("invalid metrics size")
    return torch.empty(size, dtype=torch.uint8, device=population.device)


def _read(struct_type, device_bytes: torch.Tensor):
    return struct_type.from_buffer_copy(device_bytes.cpu().numpy().tobytes())


def _cell_moments(population: CellPopulation, ws: torch.Tensor) -> _CellMoments:
    out = torch.empty(ctypes.sizeof(_CellMoments), dtype=torch.uint8, device=population.device)
    _lib.call("wbc_cell_moments", population.position.data_ptr(), population.radius.data_ptr(),
              population.cell_state.data_ptr(), population.oxygen_local.data_ptr(), population.count, ws.data_ptr(),
              ws.numel(), out.data_ptr(), population.stream)
    return _read(_CellMoments, out)


def _centroid(m: _CellMoments) -> np.ndarray:
    return np.array([m.sum_x / m.count, m.sum_y / m.count, m.sum_z / m.count], dtype=np.float32)


def _centre_distances(population: CellPopulation, m: _CellMoments, ws: torch.Tensor):
    """(r on the device, max r): |x - centre| per cell, as numpy.linalg.norm computes it in float32."""
    n = population.count
    centre = torch.from_numpy(_centroid(m)).to(population.device)
    r = torch.empty(n, dtype=torch.float32, device=population.device)
    r_max = torch.empty(1, dtype=torch.float32, device=population.device)
    _lib.call("wbc_centre_distances", population.position.data_ptr(), n, centre.data_ptr(), r.data_ptr(),
              r_max.data_ptr(), ws.data_ptr(), ws.numel(), population.stream)
    return r, float(r_max.item())


def _select(values: torch.Tensor, n: int, rank: int, ws: torch.Tensor, stream: int) -> np.float32:
    out = torch.empty(1, dtype=torch.float32, device=values.device)
    _lib.call("wbc_select_rank", values.data_ptr(), n, rank, out.data_ptr(), ws.data_ptr(), ws.numel(), stream)
    return np.float32(out.item())


def _percentile(select, n: int, percent: float) -> float:
    """numpy.percentile(a, percent) (method 'linear') of n float32 values, given select(k) = the
    k-th smallest, with numpy 2.x's arithmetic for a float32 array: q = percent / float32(100), the
    virtual index (n - 1) * q, gamma and the two-sided interpolation, all in float32."""
    f32 = np.float32
    q = np.true_divide(percent, f32(100))
    virtual = (n - 1) * q
    if virtual >= n - 1:
        return float(f32(select(n - 1)))
    if virtual < 0:
        return float(f32(select(0)))
    lo = np.floor(virtual)
    gamma = f32(virtual - lo)
    a, b = f32(select(int(lo))), f32(select(int(lo) + 1))
    diff = f32(b - a)
    if gamma >= 0.5:
        return float(f32(b - diff * f32(1 - gamma)))
    return float(f32(a + diff * gamma))


def state_counts(population: CellPopulation) -> dict[CellState, int]:
    """Number of cells in each state (exact)."""
    m = _cell_moments(population, _workspace(population))
    return {state: int(m.states[int(state)]) for state in CellState}


def population_summary(population: CellPopulation, oxygen=None) -> dict[str, float]:
    """Counts plus geometric and oxygen summaries; the keys of the reference's population_summary."""
    ws = _workspace(population)
    m = _cell_moments(population, ws)
    n = population.count
    spheroid_radius = 0.0
    if n > 0:
        r, _ = _centre_distances(population, m, ws)
        spheroid_radius = _percentile(lambda k: _select(r, n, k, ws, population.stream), n, 99)
    dead = int(m.states[int(CellState.DEAD)])
    summary = {
        "cells": n,
        "living_cells": n - dead,
        "dead_cells": dead,
        "proliferative_cells": int(m.states[int(CellState.PROLIFERATIVE)]),
        "quiescent_cells": int(m.states[int(CellState.QUIESCENT)]),
        "hypoxic_cells": int(m.states[int(CellState.HYPOXIC)]),
        "spheroid_radius": spheroid_radius,
        "mean_radius": m.sum_radius / n if n else 0.0,
    }
    if oxygen is not None:
        summary.update(_oxygen_summary(population, oxygen, m, ws))
        summary.update(species_summary(population, oxygen))
    return summary


def _oxygen_summary(population: CellPopulation, oxygen, m: _CellMoments, ws: torch.Tensor) -> dict[str, float]:
    field = oxygen.field
    nx, ny, nz = field.geometry.shape
    flags = tuple(field.boundary_flags)
    fixed = field.fixed_or_none
    out = torch.empty(ctypes.sizeof(_GridMoments), dtype=torch.uint8, device=population.device)
    _lib.call("wbc_grid_moments", field.values.data_ptr(), fixed.data_ptr() if fixed is not None else None, *flags,
              nx, ny, nz, ws.data_ptr(), ws.numel(), out.data_ptr(), population.stream)
    g = _read(_GridMoments, out)
    nan = float("nan")
    return {
        "oxygen_cells_mean": m.sum_oxygen_alive / m.alive if m.alive else nan,
        "oxygen_cells_min": float(m.min_oxygen_alive) if m.alive else nan,
        "oxygen_grid_mean": g.sum / g.count if g.count else nan,
        "oxygen_grid_min": float(g.min) if g.count else nan,
        "field_sweeps": oxygen.last_report.sweeps if oxygen.last_report else 0,
    }


def oxygen_summary(population: CellPopulation, oxygen) -> dict[str, float]:
    """Oxygen at living cells and over the updated grid nodes [mmHg]."""
    ws = _workspace(population)
    return _oxygen_summary(population, oxygen, _cell_moments(population, ws), ws)


def species_summary(population: CellPopulation, fields) -> dict[str, float]:
    """Mean/min/max at living cells of every species beyond oxygen (empty for a plain OxygenField,
    the only field this package solves)."""
    names = [n for n in getattr(fields, "names", []) if n != "oxygen"]
    if names:
        raise NotImplementedError("species beyond oxygen are not solved by this package")
    return {}


def radial_profile(population: CellPopulation, n_bins: int | None = None,
                   shell_width: float | None = None) -> dict[str, np.ndarray]:
    """Mean oxygen (and glucose, and lactate when the population carries it) and state fractions
    in concentric shells around the population centroid: ``n_bins`` equal shells up to the
    outermost cell, or shells of fixed ``shell_width`` [um]."""
    n = population.count
    probe = _workspace(population)
    m = _cell_moments(population, probe)
    if n > 0:
        r, r_max = _centre_distances(population, m, probe)
    else:
        r, r_max = torch.empty(1, dtype=torch.float32, device=population.device), 0.0
    if shell_width is not None:
        n_bins = max(1, int(np.ceil((r_max + 1e-6) / shell_width)))
        edges = shell_width * np.arange(n_bins + 1)
    else:
        n_bins = n_bins or 10
        edges = np.linspace(0.0, r_max + 1e-6, n_bins + 1)
    ws = _workspace(population, n_bins)
    edges_dev = torch.from_numpy(np.ascontiguousarray(edges, dtype=np.float64)).to(population.device)
    counts = torch.empty(n_bins * 5, dtype=torch.int32, device=population.device)
    sums = torch.empty(n_bins * 3, dtype=torch.float64, device=population.device)
    lactate = population.extra_local.get("lactate")
    _lib.call("wbc_radial_profile", r.data_ptr(), population.cell_state.data_ptr(), population.oxygen_local.data_ptr(),
              population.glucose_local.data_ptr(), lactate.data_ptr() if lactate is not None else None, n,
              edges_dev.data_ptr(), n_bins, counts.data_ptr(), sums.data_ptr(), ws.data_ptr(), ws.numel(),
              population.stream)
    c = counts.cpu().numpy().reshape(n_bins, 5).astype(np.int64)
    s = sums.cpu().numpy().reshape(n_bins, 3)
    cells = c[:, 0]
    with np.errstate(invalid="ignore", divide="ignore"):
        def mean(k):
            return np.where(cells > 0, s[:, k] / np.maximum(cells, 1), np.nan)

        out = {"r_outer": edges[1:], "cells": cells, "oxygen_mean": mean(0), "glucose_mean": mean(1)}
        if lactate is not None:
            out["lactate_mean"] = mean(2)
        for state in CellState:
            out[f"fraction_{state.name.lower()}"] = np.where(cells > 0, c[:, 1 + int(state)] / np.maximum(cells, 1),
                                                              np.nan)
    return out


def shell_radii(profile: dict[str, np.ndarray]) -> dict[str, float]:
    """Characteristic radii [um] from a radial profile: the outer edge of the outermost shell
    whose fraction of dead / (hypoxic or dead) / non-proliferative cells is at least one half;
    0 when no shell qualifies (host arithmetic on the profile's few numbers)."""

    def outermost(fraction):
        hit = np.flatnonzero(np.nan_to_num(fraction) >= 0.5)
        return float(profile["r_outer"][hit[-1]]) if hit.size else 0.0

    dead = profile["fraction_dead"]
    return {
        "necrotic_radius": outermost(dead),
        "hypoxic_radius": outermost(profile["fraction_hypoxic"] + dead),
        "non_proliferative_radius": outermost(1.0 - profile["fraction_proliferative"]),
    }
