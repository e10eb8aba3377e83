"""GPU parity of RNG streams, the slot-allocation scan, lifecycle decisions, daughter placement
and the full cell step against the oracle: bit-exact, including positions."""

import math
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from warpbiocell_amd import _lib
from warpbiocell_amd import workloads as W
from warpbiocell_amd.cells.initialization import spherical_cluster
from warpbiocell_amd.cells.lifecycle import CapacityError, LifecycleParams, lifecycle_step
from warpbiocell_amd.cells.mechanics import ContactParams, make_neighbor_grid
from warpbiocell_amd.cells.state import CellPopulation
from warpbiocell_amd.fields.diffusion import SolverSettings
from warpbiocell_amd.fields.oxygen import OxygenField, OxygenParams
from warpbiocell_amd.fields.scalar_field import GridGeometry
from warpbiocell_amd.simulation.simulator import TimeStepping, cell_step

pytestmark = pytest.mark.gpu
R = 8.0


def test_rng_streams_are_bit_exact(oracle, device):
    pop = CellPopulation.from_numpy(np.zeros((10, 3), np.float32), np.ones(10), capacity=70001, device=device, seed=77)
    np.testing.assert_array_equal(pop.rng_numpy(), oracle.init_rng(70001, 77))


@pytest.mark.parametrize("n", [1, 63, 64, 65, 4095, 4096, 4097, 100_003, 1_048_577, 4_194_304, 4_194_305, 9_000_001])
def test_inclusive_scan_matches_cumsum(device, n):
    flags = (np.random.default_rng(n).random(n) < 0.37).astype(np.int32)
    inp = torch.from_numpy(flags).to(device)
    out = torch.empty_like(inp)
    ws = torch.empty(_lib.query("wbc_scan_workspace_bytes", n), dtype=torch.uint8, device=device)
    _lib.call("wbc_scan_inclusive_i32", inp.data_ptr(), out.data_ptr(), n, ws.data_ptr(), ws.numel(),
              _lib.stream_of(device))
    np.testing.assert_array_equal(out.cpu().numpy(), np.cumsum(flags))
    _lib.call("wbc_scan_exclusive_i32", inp.data_ptr(), out.data_ptr(), n, ws.data_ptr(), ws.numel(),
              _lib.stream_of(device))
    np.testing.assert_array_equal(out.cpu().numpy(), np.cumsum(flags) - flags)


def _pair(n, device, oracle, capacity, seed=0, states=None, oxygen=None, counts=None):
    pos = spherical_cluster(n, R, spacing_factor=1.0, jitter=0.1, seed=seed)
    gpu = CellPopulation.from_numpy(pos, np.full(n, R), capacity=capacity, device=device, seed=seed, states=states)
    ref = oracle.OraclePopulation.from_numpy(pos, np.full(n, R), capacity=capacity, seed=seed, states=states)
    if oxygen is not None:
        gpu.oxygen_local[:n] = torch.from_numpy(oxygen).to(device)
        ref.oxygen_local[:n] = oxygen
    if counts is not None:
        gpu.neighbor_count[:n] = torch.from_numpy(counts).to(device)
        ref.neighbor_count[:n] = counts
    return gpu, ref


def _same(gpu, ref, fields=("position", "radius", "cell_state", "cell_type", "age", "oxygen_local",
                            "glucose_local", "velocity", "neighbor_count", "divide_flag", "fate_flags")):
    assert gpu.count == ref.count
    for name in fields:
        np.testing.assert_array_equal(getattr(gpu, name).cpu().numpy(), getattr(ref, name), err_msg=name)
    np.testing.assert_array_equal(gpu.rng_numpy(), ref.rng_state, err_msg="rng_state")


def test_lifecycle_with_oxygen_thresholds_is_bit_exact(oracle, device):
    n = 20000
    rng = np.random.default_rng(3)
    states = rng.integers(0, 4, n).astype(np.int32)
    oxygen = rng.uniform(0.0, 12.0, n).astype(np.float32)
    counts = rng.integers(0, 14, n).astype(np.int32)
    gpu, ref = _pair(n, device, oracle, 3 * n, seed=3, states=states, oxygen=oxygen, counts=counts)
    params = LifecycleParams(division_rate=0.4, death_rate=0.02, anoxic_death_rate=0.5, hypoxia_threshold=8.0,
                             death_threshold=2.0, inhibition_threshold=8)
    for _ in range(3):
        created = lifecycle_step(gpu, params, 0.5)
        oracle.lifecycle_decide(ref, params, 0.5)
        assert created == oracle.commit_divisions(ref, params)
    _same(gpu, ref)
    assert gpu.count > n


def test_every_cell_divides_once_and_pairs_are_symmetric(device):
    n = 300
    pos = spherical_cluster(n, R, spacing_factor=1.0, jitter=0.1, seed=0)
    pop = CellPopulation.from_numpy(pos, np.full(n, R), capacity=3 * n, device=device)
    params = LifecycleParams(division_rate=1.0e3, death_rate=0.0, inhibition_threshold=10 ** 6)
    assert lifecycle_step(pop, params, dt=1.0) == n and pop.count == 2 * n
    x = pop.positions_numpy()
    np.testing.assert_allclose(np.linalg.norm(x[n:] - x[:n], axis=1), R, rtol=1e-5)
    np.testing.assert_allclose(0.5 * (x[n:] + x[:n]), pos, atol=1e-5)
    assert np.all(pop.cell_state.cpu().numpy()[2 * n:] == 3)


def test_capacity_overflow_raises_and_places_nothing(device):
    n = 50
    pos = spherical_cluster(n, R, spacing_factor=1.0, jitter=0.1, seed=0)
    pop = CellPopulation.from_numpy(pos, np.full(n, R), capacity=n + 10, device=device)
    with pytest.raises(CapacityError):
        lifecycle_step(pop, LifecycleParams(division_rate=1e3, inhibition_threshold=10 ** 6), dt=1.0)
    assert pop.count == n
    np.testing.assert_array_equal(pop.positions_numpy(), pos)
    assert np.all(pop.position.cpu().numpy()[n:] == 0.0)


def test_full_cell_step_with_oxygen_is_bit_exact(oracle, device):
    """simulation/simulator.py:58-81: field, lifecycle with divisions, 4 contact substeps."""
    n0 = 1500
    pos = spherical_cluster(n0, 5.0, spacing_factor=1.0, jitter=0.1, seed=4)
    radii = np.full(n0, 5.0, dtype=np.float32)
    geom = GridGeometry.centered_cube(400.0, 10.0)
    params = OxygenParams(W.D_UM2_PER_H, W.UPTAKE_MMHG_UM3_PER_H, 3.4, 40.0)
    settings = SolverSettings(omega=1.8, tolerance=1e-5, max_sweeps=2000, check_every=10)
    lifecycle = LifecycleParams(division_rate=0.5, death_rate=0.01, anoxic_death_rate=0.5, hypoxia_threshold=8.0,
                                death_threshold=2.0, inhibition_threshold=8)
    contact = ContactParams(stiffness=10.0, damping=1.0, max_radius=5.0, margin=0.5)
    stepping = TimeStepping(dt_cells=0.1, dt_mechanics=0.025)
    pop = CellPopulation.from_numpy(pos, radii, capacity=4 * n0, device=device, seed=11)
    ox = OxygenField(geom, params, settings, device=device)
    grid = make_neighbor_grid(contact, device)
    ref = oracle.OraclePopulation.from_numpy(pos, radii, capacity=4 * n0, seed=11)
    a = oracle.SolveArgs(dx=geom.dx, diffusion=params.diffusion_coefficient, uptake_max=params.uptake_max,
                         michaelis_k=params.michaelis_k, boundary_value=40.0)
    rox = oracle.OracleOxygen(shape=geom.shape, origin=geom.origin, args=a, omega=1.8, max_sweeps=2000,
                              tolerance=1e-5, check_every=10)
    for step in range(6):
        report = cell_step(pop, grid, lifecycle, contact, stepping, oxygen=ox)
        daughters, field = oracle.oracle_cell_step(ref, lifecycle, contact, stepping, oxygen=rox)
        assert report.daughters == daughters and report.field_sweeps == field[0], f"step {step}"
        _same(pop, ref, fields=("position", "cell_state", "age", "oxygen_local", "neighbor_count"))
        np.testing.assert_array_equal(ox.numpy(), rox.values)
    assert pop.count > n0


def test_result_does_not_depend_on_capacity(device):
    def run(capacity):
        pos = spherical_cluster(400, R, spacing_factor=1.0, jitter=0.1, seed=5)
        pop = CellPopulation.from_numpy(pos, np.full(400, R), capacity=capacity, device=device, seed=5)
        contact = ContactParams(stiffness=10.0, damping=1.0, max_radius=R, margin=2.0)
        grid = make_neighbor_grid(contact, device)
        lc = LifecycleParams(division_rate=0.2, death_rate=0.01, inhibition_threshold=8)
        for _ in range(12):
            cell_step(pop, grid, lc, contact, TimeStepping(dt_cells=0.25, dt_mechanics=0.025))
        return pop.count, pop.positions_numpy(), pop.states_numpy()

    a, b = run(4000), run(9000)
    assert a[0] == b[0] > 400
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])
