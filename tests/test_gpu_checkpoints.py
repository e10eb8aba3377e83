"""Checkpoints: the reference's .npz layout (io/checkpoints.py:16-49) and a restored run that
continues bit for bit, divisions included (slot allocation and per-slot RNG streams)."""

from dataclasses import replace

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def scenario(device):
    from warpbiocell_amd import workloads as W
    from warpbiocell_amd.cells.lifecycle import LifecycleParams
    from warpbiocell_amd.simulation import scenario as S

    w = W.spheroid_1k_41()
    w = replace(w, capacity=4 * w.n)
    sc = S.build(w, device)
    # divisions, deaths and crowding on, so the continuation exercises slots, streams and counts
    sc = replace(sc, lifecycle=LifecycleParams(division_rate=0.3, death_rate=0.05, inhibition_threshold=6))
    return sc


def step(sc, k):
    from warpbiocell_amd.simulation.simulator import cell_step

    for _ in range(k):
        cell_step(sc.population, sc.grid, sc.lifecycle, sc.contact, sc.stepping, oxygen=sc.oxygen)
    torch.cuda.synchronize()


def test_checkpoint_has_the_reference_layout(device, tmp_path):
    from warpbiocell_amd.io.checkpoints import CELL_ARRAYS, load_checkpoint, save_checkpoint

    sc = scenario(device)
    step(sc, 2)
    path = save_checkpoint(tmp_path / "ck" / "state_0000.20h.npz", sc.population, sc.oxygen, 0.2)
    data = load_checkpoint(path)
    n = sc.population.count
    expected = set(CELL_ARRAYS) | {"count", "capacity", "seed", "time_h", "field", "density", "grid_origin",
                                   "grid_dx", "neighbor_count"}
    assert set(data) == expected
    assert data["position"].shape == (n, 3) and data["position"].dtype == np.float32
    for name in ("radius", "age", "oxygen_local", "glucose_local"):
        assert data[name].shape == (n,) and data[name].dtype == np.float32, name
    for name in ("cell_state", "cell_type", "fate_flags", "neighbor_count"):
        assert data[name].shape == (n,) and data[name].dtype == np.int32, name
    assert data["rng_state"].dtype == np.uint32
    assert int(data["count"]) == n and int(data["capacity"]) == sc.population.capacity
    assert float(data["time_h"]) == 0.2 and float(data["grid_dx"]) == sc.geometry.dx
    assert data["field"].shape == tuple(sc.geometry.shape) and data["field"].dtype == np.float32
    np.testing.assert_array_equal(data["field"], sc.oxygen.numpy())
    np.testing.assert_array_equal(data["rng_state"], sc.population.rng_numpy()[:n].view(np.uint32))


def test_restored_run_continues_bit_for_bit(device, tmp_path):
    from warpbiocell_amd.io.checkpoints import load_checkpoint, restore_field, restore_population, save_checkpoint

    a = scenario(device)
    step(a, 3)
    path = save_checkpoint(tmp_path / "ck.npz", a.population, a.oxygen, 0.3)
    n_saved = a.population.count
    step(a, 3)
    b = scenario(device)
    data = load_checkpoint(path)
    pop = restore_population(data, device)
    assert pop.count == n_saved
    restore_field(data, b.oxygen)
    b = replace(b, population=pop)
    step(b, 3)
    assert b.population.count == a.population.count > n_saved
    n = a.population.count
    for name in ("position", "radius", "cell_state", "age", "oxygen_local", "velocity", "neighbor_count",
                 "divide_flag", "division_offset", "rng_state"):
        np.testing.assert_array_equal(getattr(b.population, name)[:n].cpu().numpy().view(np.uint8),
                                      getattr(a.population, name)[:n].cpu().numpy().view(np.uint8), err_msg=name)
    np.testing.assert_array_equal(b.oxygen.numpy().view(np.uint32), a.oxygen.numpy().view(np.uint32))


def test_mismatched_grid_is_refused(device, tmp_path):
    from warpbiocell_amd.io.checkpoints import load_checkpoint, restore_field, save_checkpoint

    sc = scenario(device)
    data = load_checkpoint(save_checkpoint(tmp_path / "ck.npz", sc.population, sc.oxygen, 0.0))
    data["grid_dx"] = np.array(sc.geometry.dx * 2)
    with pytest.raises(ValueError):
        restore_field(data, sc.oxygen)
