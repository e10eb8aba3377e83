"""Batched replica sweep vs running each replica alone with the plain step.

Each replica is built twice from the same seed: one copy advances with
``simulation.simulator.cell_step`` (hash grid of 128 buckets per axis), the other inside one
``BatchedSweep`` (64 buckets per axis; the contact result does not depend on the bucket count).
After every step the solve reports must match and, at the end, every per-cell array and every
field bit for bit. The replicas cover:

- cubic and non-cubic grids with partial tiles on every axis;
- an empty population;
- capacity gaps between replicas;
- a solve that stops at max_sweeps unconverged;
- oxygen-dependent death.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CELL_ARRAYS = ("position", "radius", "cell_state", "age", "oxygen_local", "glucose_local", "velocity",
               "neighbor_count", "divide_flag", "division_offset", "fate_flags", "rng_state")

SPECS = [
    # name, cells, capacity, grid shape, dx, boundary, seed, lifecycle kind, max_sweeps, tolerance
    ("cube21", 800, 900, (21, 21, 21), 20.0, 40.0, 0, "plain", 2000, 1e-6),
    ("hypoxic33", 3000, 3000, (33, 33, 33), 20.0, 20.0, 1, "oxygen", 2000, 1e-6),
    ("skew", 1500, 1600, (29, 37, 45), 16.0, 60.0, 2, "plain", 2000, 1e-6),
    ("capped", 200, 256, (9, 9, 9), 40.0, 100.0, 3, "plain", 35, 1e-12),
    ("empty", 0, 16, (11, 13, 15), 30.0, 150.0, 4, "plain", 2000, 1e-6),
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


@pytest.mark.parametrize("schedule", ["default", "wide/2", "wide/4"])
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
            assert (a.field_sweeps, a.field_converged) == (b.field_sweeps, b.field_converged), (r.label, k, a, b)
            assert np.float32(a.field_residual) == np.float32(b.field_residual), (r.label, k, a, b)
    capped = [i for i, s in enumerate(SPECS) if s[0] == "capped"][0]
    assert ref_reports[0][capped].field_sweeps == 35 and not ref_reports[0][capped].field_converged
    assert_same_state(serial, batch)
    assert batch.total_sweeps() == [sum(rep[i].field_sweeps for rep in ref_reports) for i in range(len(SPECS))]
    # the oxygen-dependent replica lost cells to anoxia, so the lifecycle path was exercised
    hyp = [i for i, s in enumerate(SPECS) if s[0] == "hypoxic33"][0]
    states = batch.population_view(hyp, "cell_state")[:3000].cpu().numpy()
    assert (states == 3).any() and (states != 3).any()
    batch.write_back()
    batch.close()


@pytest.mark.parametrize("schedule", ["default", "wide/2"])
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
    from dataclasses import replace

    from warpbiocell_amd.simulation.batched import BatchedSweep

    r = build(SPECS[0], device)
    with pytest.raises(NotImplementedError):
        BatchedSweep([replace(r, lifecycle=replace(r.lifecycle, division_rate=0.1))])
