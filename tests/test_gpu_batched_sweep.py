"""The 15-replica parameter sweep at full size, batched against the plain step run replica after
replica (the reference's Sweep.run order, simulation/sweep.py:113-162).

sweep_15: 609,575 cells over 15 spheroids of 2k-182k cells, 12.5M grid nodes on 41^3 ... 161^3
grids, the lifecycle of the reference sweep's base experiment (divisions at 0.0289 /h, anoxic
death, hypoxia, contact inhibition). Three steps: reports (daughters, sweeps, residual) after each
step, then every cell array, the counts and every field, bit for bit."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CELL_ARRAYS = ("position", "radius", "cell_state", "cell_type", "age", "oxygen_local", "glucose_local", "velocity",
               "neighbor_count", "divide_flag", "division_offset", "fate_flags", "rng_state")


def test_sweep_15_batched_equals_serial_steps(device):
    from warpbiocell_amd import workloads as W
    from warpbiocell_amd.simulation import scenario
    from warpbiocell_amd.simulation.batched import BatchedSweep
    from warpbiocell_amd.simulation.simulator import cell_step

    ws = W.sweep_15()
    serial = [scenario.build(w, device) for w in ws]
    batched = [scenario.build(w, device) for w in ws]
    batch = BatchedSweep([sc.replica() for sc in batched])
    assert sum(sc.population.count for sc in batched) == 609_575
    total_daughters = 0
    for k in range(3):
        print(f"sweep_15 step {k}", flush=True)
        ref = [cell_step(sc.population, sc.grid, sc.lifecycle, sc.contact, sc.stepping, oxygen=sc.oxygen)
               for sc in serial]
        batch.step(1, use_graph=True)
        got = batch.step_reports()
        for w, a, b in zip(ws, ref, got):
            assert (a.daughters, a.field_sweeps, a.field_converged) == \
                   (b.daughters, b.field_sweeps, b.field_converged), (w.name, k, a, b)
            assert np.float32(a.field_residual) == np.float32(b.field_residual), (w.name, k)
        total_daughters += sum(a.daughters for a in ref)
    torch.cuda.synchronize()
    assert total_daughters > 0
    assert batch.counts() == [sc.population.count for sc in serial]
    for i, sc in enumerate(serial):
        n = sc.population.count
        for name in CELL_ARRAYS:
            a = getattr(sc.population, name)[:n].cpu().numpy()
            b = batch.population_view(i, name)[:n].cpu().numpy()
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (ws[i].name, name)
        assert np.array_equal(sc.oxygen.values.cpu().numpy().view(np.uint32),
                              batch.field_view(i).cpu().numpy().view(np.uint32)), ws[i].name
    batch.close()
