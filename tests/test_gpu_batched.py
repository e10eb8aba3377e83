"""Batched replica sweep vs running each replica alone with the plain step, and vs the oracle.

Each replica is built twice from the same seed: one copy advances with
``simulation.simulator.cell_step`` (hash grid of 128 buckets per axis), the other inside one
``BatchedSweep`` (64 buckets per axis; the contact result does not depend on the bucket count).
After every step the reports (solve and daughters) must match and, at the end, every per-cell
array, the cell counts and every field bit for bit. The replicas cover:

- cubic and non-cubic grids with partial tiles on every axis;
- an empty population;
- capacity gaps between replicas;
- a solve that stops at max_sweeps unconverged;
- oxygen-dependent death;
- growth: many divisions per step with deaths, and the lifecycle of the reference's sweep
  (configs/tumor_spheroid.yaml:31-38) with contact inhibition.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CELL_ARRAYS = ("position", "radius", "cell_state", "cell_type", "age", "oxygen_local", "glucose_local", "velocity",
               "neighbor_count", "divide_flag", "division_offset", "fate_flags", "rng_state")

SPECS = [
    # name, cells, capacity, grid shape, dx, boundary, seed, lifecycle kind, max_sweeps, tolerance
    ("cube21", 800, 900, (21, 21, 21), 20.0, 40.0, 0, "plain", 2000, 1e-6),
    ("hypoxic33", 3000, 3000, (33, 33, 33), 20.0, 20.0, 1, "oxygen", 2000, 1e-6),
    ("skew", 1500, 1600, (29, 37, 45), 16.0, 60.0, 2, "plain", 2000, 1e-6),
    ("capped", 200, 256, (9, 9, 9), 40.0, 100.0, 3, "plain", 35, 1e-12),
    ("empty", 0, 16, (11, 13, 15), 30.0, 150.0, 4, "growth", 2000, 1e-6),
    ("growing", 1200, 2600, (25, 27, 23), 18.0, 40.0, 5, "growth", 2000, 1e-6),
    ("sweep_lifecycle", 2500, 3600, (31, 31, 31), 20.0, 20.0, 6, "sweep", 2000, 1e-6),
]


def build(spec, device):
    from warpbiocell_amd import workloads as W
    from warpbiocell_amd.cells.lifecycle import LifecycleParams
    from warpbiocell_amd.cells.mechanics import ContactParams
    from warpbiocell_amd.cells.state import CellPopulation
    from warpbiocell_amd.fields.diffusion import SolverSettings
    from warpbiocell_amd.fields.oxygen import OxygenField, OxygenParams
    from warpbiocell_amd.fields.scalar_field import GridGeometry
    from warpbiocell_amd.simulation.batched import Replica
    from warpbiocell_amd.simulation.simulator import TimeStepping

    name, n, cap, shape, dx, ob, seed, kind, max_sweeps, tol = spec
    rng = np.random.default_rng(seed)
    radii = (5.0 * np.exp(0.1 * rng.standard_normal(n))).astype(np.float32)
    R = W.ball_radius_for_fraction(max(n, 1), 125.0 * math.exp(4.5 * 0.01), 0.6)
    pos = W.ball_uniform(n, R, rng) if n else np.zeros((0, 3), np.float32)
    origin = tuple(-0.5 * (s - 1) * dx for s in shape)
    geom = GridGeometry(origin=origin, dx=dx, shape=shape)
    params = OxygenParams(W.D_UM2_PER_H, W.UPTAKE_MMHG_UM3_PER_H, W.MICHAELIS_K, ob)
    settings = SolverSettings(omega=1.8, tolerance=tol, max_sweeps=max_sweeps, check_every=10)
    if kind == "oxygen":
        lifecycle = LifecycleParams.oxygen_dependent(division_rate=0.0, inhibition_threshold=10 ** 6)
    elif kind == "growth":  # about 14 % of cells divide per step, 0.2 % die
        lifecycle = LifecycleParams(division_rate=1.5, death_rate=0.02, inhibition_threshold=10 ** 6)
    elif kind == "sweep":
        lifecycle = LifecycleParams(**W.SWEEP_LIFECYCLE)
    else:
        lifecycle = LifecycleParams(division_rate=0.0, death_rate=0.01, inhibition_threshold=10 ** 6)
    contact = ContactParams(stiffness=W.CONTACT_RATE, damping=1.0, max_radius=float(radii.max()) if n else 5.0)
    stepping = TimeStepping(dt_cells=W.DT_CELLS, dt_mechanics=W.DT_MECHANICS)
    pop = CellPopulation.from_numpy(pos, radii, capacity=cap, device=device, seed=seed)
    ox = OxygenField(geom, params, settings, device=device)
    return Replica(pop, ox, lifecycle, contact, stepping, label=name)


def run_serial(replicas, steps):
    from warpbiocell_amd.cells.mechanics import make_neighbor_grid
    from warpbiocell_amd.simulation.simulator import cell_step

    reports = []
    grids = [make_neighbor_grid(r.contact, r.population.device) for r in replicas]
    for _ in range(steps):
        reports.append([cell_step(r.population, g, r.lifecycle, r.contact, r.stepping, oxygen=r.oxygen)
                        for r, g in zip(replicas, grids)])
    torch.cuda.synchronize()
    return reports


def assert_same_state(serial, batch):
    assert batch.counts() == [r.population.count for r in serial]
    for i, r in enumerate(serial):
        n = r.population.count
        for name in CELL_ARRAYS:
            a = getattr(r.population, name)[:n].cpu().numpy()
            b = batch.population_view(i, name)[:n].cpu().numpy()
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), f"{r.label}: {name} differs"
        fa = r.oxygen.values.cpu().numpy()
        fb = batch.field_view(i).cpu().numpy()
        assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32)), f"{r.label}: field differs"


def use_schedule(batch, name):
    """Select the schedule called `name` ('default' keeps the constructor's)."""
    if name == "default":
        return
    from warpbiocell_amd.autotune import batch_schedules

    tile, bpc, _ = next(s for s in batch_schedules() if s[2] == name)
    batch.set_schedule(tile, bpc)


@pytest.mark.parametrize("schedule", ["default", "wide/2", "wide/4", "wide/quad"])
def test_batched_steps_are_bit_exact_per_replica(device, schedule):
    from warpbiocell_amd.simulation.batched import BatchedSweep

    steps = 3
    serial = [build(s, device) for s in SPECS]
    batched = [build(s, device) for s in SPECS]
    ref_reports = run_serial(serial, steps)
    batch = BatchedSweep(batched, hash_dim=64)
    use_schedule(batch, schedule)
    for k in range(steps):
        print(f"batched step {k}", flush=True)
        batch.step(1, use_graph=True)
        got = batch.step_reports()
        for r, a, b in zip(serial, ref_reports[k], got):
            assert (a.daughters, a.field_sweeps, a.field_converged) == \
                   (b.daughters, b.field_sweeps, b.field_converged), (r.label, k, a, b)
            assert np.float32(a.field_residual) == np.float32(b.field_residual), (r.label, k, a, b)
    capped = [i for i, s in enumerate(SPECS) if s[0] == "capped"][0]
    assert ref_reports[0][capped].field_sweeps == 35 and not ref_reports[0][capped].field_converged
    assert_same_state(serial, batch)
    assert batch.total_sweeps() == [sum(rep[i].field_sweeps for rep in ref_reports) for i in range(len(SPECS))]
    # the oxygen-dependent replica lost cells to anoxia, so the lifecycle path was exercised
    hyp = [i for i, s in enumerate(SPECS) if s[0] == "hypoxic33"][0]
    states = batch.population_view(hyp, "cell_state")[:3000].cpu().numpy()
    assert (states == 3).any() and (states != 3).any()
    # the growing replicas placed daughters, so the batched slot scan and placement ran
    grow = [i for i, s in enumerate(SPECS) if s[0] == "growing"][0]
    assert sum(rep[grow].daughters for rep in ref_reports) > 200
    batch.write_back()
    assert [r.population.count for r in batched] == [r.population.count for r in serial]
    batch.close()


@pytest.mark.parametrize("schedule", ["default", "wide/2", "wide/quad"])
def test_graph_replays_match_eager_steps(device, schedule):
    from warpbiocell_amd.simulation.batched import BatchedSweep

    eager = BatchedSweep([build(s, device) for s in SPECS[:3]], hash_dim=64)
    graph = BatchedSweep([build(s, device) for s in SPECS[:3]], hash_dim=64)
    use_schedule(eager, schedule)
    use_schedule(graph, schedule)
    for k in range(3):
        print(f"replay {k}", flush=True)
        eager.step(1, use_graph=False)
        graph.step(1, use_graph=True)
        # unrelated eager work and a synchronisation between replays
        torch.ones(1 << 20, device=device).sum().item()
        assert [r.sweeps for r in eager.report()] == [r.sweeps for r in graph.report()]
        for i in range(3):
            assert torch.equal(eager.field_view(i), graph.field_view(i))
            for name in ("position", "velocity", "oxygen_local", "rng_state"):
                assert torch.equal(eager.population_view(i, name), graph.population_view(i, name))
    eager.close()
    graph.close()


@pytest.mark.parametrize("blocks_per_cu", [0, 1])
def test_every_schedule_gives_the_same_solve(device, blocks_per_cu):
    from warpbiocell_amd import _lib
    from warpbiocell_amd.simulation.batched import BatchedSweep

    batch = BatchedSweep([build(s, device) for s in SPECS], hash_dim=64)
    batch.step(1)
    start = batch.field.clone()
    results = []
    for tile in range(_lib.query("wbc_batch_schedule_count")):
        print(f"schedule tile {tile} blocks/CU {blocks_per_cu}", flush=True)
        batch.field.copy_(start)
        batch.set_schedule(tile, blocks_per_cu)
        batch.solve()
        results.append(([r.sweeps for r in batch.report()], batch.field.clone()))
    for sweeps, field in results[1:]:
        assert sweeps == results[0][0]
        assert torch.equal(field, results[0][1])
    batch.close()


def test_unsupported_replicas_raise(device):
    from warpbiocell_amd.fields.oxygen import OxygenField
    from warpbiocell_amd.fields.scalar_field import Boundary
    from warpbiocell_amd.simulation.batched import BatchedSweep

    r = build(SPECS[0], device)
    ox = r.oxygen
    r.oxygen = OxygenField(ox.geometry, ox.params, ox.settings, device=device,
                           boundary=(Boundary.DIRICHLET, Boundary.NEUMANN, Boundary.DIRICHLET))
    with pytest.raises(NotImplementedError):
        BatchedSweep([r])


def _oracle_replica(oracle, spec):
    """The oracle's population and oxygen field for `spec`, built from the same arrays as build()."""
    from warpbiocell_amd import workloads as W

    name, n, cap, shape, dx, ob, seed, kind, max_sweeps, tol = spec
    rng = np.random.default_rng(seed)
    radii = (5.0 * np.exp(0.1 * rng.standard_normal(n))).astype(np.float32)
    R = W.ball_radius_for_fraction(max(n, 1), 125.0 * math.exp(4.5 * 0.01), 0.6)
    pos = W.ball_uniform(n, R, rng) if n else np.zeros((0, 3), np.float32)
    origin = tuple(-0.5 * (s - 1) * dx for s in shape)
    pop = oracle.OraclePopulation.from_numpy(pos, radii, capacity=cap, seed=seed)
    a = oracle.SolveArgs(dx=dx, diffusion=W.D_UM2_PER_H, uptake_max=W.UPTAKE_MMHG_UM3_PER_H,
                         michaelis_k=W.MICHAELIS_K, boundary_value=ob)
    ox = oracle.OracleOxygen(shape=shape, origin=origin, args=a, omega=1.8, max_sweeps=max_sweeps, tolerance=tol,
                             check_every=10)
    return pop, ox


def test_batched_steps_equal_the_oracle_step(oracle, device):
    """Every replica after three batched steps (divisions, deaths and all) against
    oracle.oracle_cell_step run three times on the same inputs: cell arrays, counts, fields and
    reports bit for bit."""
    from warpbiocell_amd.simulation.batched import BatchedSweep

    specs = [s for s in SPECS if s[0] in ("cube21", "hypoxic33", "empty", "growing", "sweep_lifecycle")]
    replicas = [build(s, device) for s in specs]
    batch = BatchedSweep(replicas, hash_dim=64)
    steps = 3
    refs = [_oracle_replica(oracle, s) for s in specs]
    for k in range(steps):
        print(f"oracle comparison step {k}", flush=True)
        batch.step(1, use_graph=True)
        got = batch.step_reports()
        for r, (pop, ox), b in zip(replicas, refs, got):
            daughters, field = oracle.oracle_cell_step(pop, r.lifecycle, r.contact, r.stepping, oxygen=ox)
            assert (daughters, field[0], field[2]) == (b.daughters, b.field_sweeps, b.field_converged), (r.label, k)
            assert np.float32(field[1]) == np.float32(b.field_residual), (r.label, k)
    counts = batch.counts()
    for i, (r, (pop, ox)) in enumerate(zip(replicas, refs)):
        n = pop.count
        assert counts[i] == n, r.label
        for name in ("position", "radius", "cell_state", "age", "oxygen_local", "velocity", "neighbor_count",
                     "divide_flag", "division_offset", "rng_state"):
            got = batch.population_view(i, name)[:n].cpu().numpy()
            exp = getattr(pop, name)[:n]
            assert np.array_equal(got.view(np.uint8), np.ascontiguousarray(exp).view(np.uint8)), (r.label, name)
        assert np.array_equal(batch.field_view(i).cpu().numpy(), ox.values), r.label
    batch.close()


@pytest.mark.parametrize("schedule", ["default", "wide/quad"])
def test_repeated_batched_runs_give_the_serial_result_every_time(device, schedule):
    """Eight fresh batched runs of the same replicas, eager and captured alternately, each equal to
    one serial run bit for bit. The finalize of the batched deposit reads accumulators that other
    kernels built with atomics; plain loads there once returned wrong values for a few nodes in
    some runs and not others, which a single run does not show."""
    from warpbiocell_amd.simulation.batched import BatchedSweep

    steps = 3
    specs = [s for s in SPECS if s[0] in ("cube21", "hypoxic33", "growing", "sweep_lifecycle")]
    serial = [build(s, device) for s in specs]
    run_serial(serial, steps)
    for run in range(8):
        print(f"repeat {run}", flush=True)
        batch = BatchedSweep([build(s, device) for s in specs], hash_dim=64)
        use_schedule(batch, schedule)
        for _ in range(steps):
            batch.step(1, use_graph=bool(run & 1))
        torch.cuda.synchronize()
        assert_same_state(serial, batch)
        batch.close()


def test_capacity_overflow_places_nothing_for_that_replica(device):
    """A replica whose daughters would not fit: none are placed for it (its count stays), the
    others grow as serial steps do, and reading the counts raises CapacityError."""
    from warpbiocell_amd.cells.lifecycle import CapacityError
    from warpbiocell_amd.simulation.batched import BatchedSweep

    tight = ("tight", 1200, 1210, (25, 27, 23), 18.0, 40.0, 5, "growth", 2000, 1e-6)
    roomy = [s for s in SPECS if s[0] == "growing"][0]
    serial = build(roomy, device)
    ref = run_serial([serial], 1)[0][0]
    batch = BatchedSweep([build(tight, device), build(roomy, device)], hash_dim=64)
    batch.step(1, use_graph=True)
    with pytest.raises(CapacityError, match="tight"):
        batch.counts()
    with pytest.raises(CapacityError):
        batch.step_reports()
    from warpbiocell_amd import _lib

    counts, daughters, err = (ctypes.c_int32 * 2)(), (ctypes.c_int32 * 2)(), (ctypes.c_int32 * 2)()
    _lib.call("wbc_batch_cell_counts", batch._h(), batch.stream, ctypes.byref(counts), ctypes.byref(daughters),
              ctypes.byref(err))
    assert list(err) == [1, 0]
    assert list(counts) == [1200, serial.population.count] and list(daughters) == [0, ref.daughters]
    assert ref.daughters > 0
    assert batch.population_view(0, "radius")[1200:].abs().sum().item() == 0.0  # no daughter placed
    batch.close()


def test_three_replicas_at_128_buckets_scan_their_tile_sums_in_a_launch(device):
    """3 x 128^3 buckets = 1,536 scan tiles, beyond the 1,024 up to which each tile adds up the
    earlier tiles' sums itself: the bucket scan takes its tile-sum launch in every substep, and the
    replicas still equal the serial steps bit for bit."""
    from warpbiocell_amd.simulation.batched import BatchedSweep

    specs = [s for s in SPECS if s[0] in ("growing", "sweep_lifecycle", "skew")]
    serial = [build(s, device) for s in specs]
    batched = [build(s, device) for s in specs]
    ref_reports = run_serial(serial, 2)
    batch = BatchedSweep(batched, hash_dim=128)
    for k in range(2):
        batch.step(1, use_graph=True)
        assert [(a.daughters, a.field_sweeps) for a in ref_reports[k]] == \
               [(b.daughters, b.field_sweeps) for b in batch.step_reports()]
    assert_same_state(serial, batch)
    batch.close()
