"""The HIP path reproduces the committed golden fixtures (tests/golden/*.npz, written from the CPU
oracle by tests/golden/make_golden.py) bit for bit, from the stored inputs: RNG streams, the int64
deposit, a steady-state solve, the hash-grid order, contact velocities and relaxation, the
tissue-wall push, and three full cell steps with oxygen, divisions and mechanics."""

from pathlib import Path

import numpy as np
import pytest
import torch

from warpbiocell_amd import _lib
from warpbiocell_amd.cells.lifecycle import LifecycleParams
from warpbiocell_amd.cells.mechanics import ContactParams, contact_substep, make_neighbor_grid, relax_contacts
from warpbiocell_amd.cells.state import CellPopulation
from warpbiocell_amd.fields import diffusion as Dn
from warpbiocell_amd.fields.densities import CellDensities
from warpbiocell_amd.fields.diffusion import SolverSettings, UptakeKinetics
from warpbiocell_amd.fields.oxygen import OxygenField, OxygenParams
from warpbiocell_amd.fields.scalar_field import GridGeometry, ScalarField
from warpbiocell_amd.geometry.region import TissueRegion
from warpbiocell_amd.simulation.simulator import TimeStepping, cell_step

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def _geom(g):
    return GridGeometry(origin=tuple(float(v) for v in g["origin"]), dx=float(g["dx"]),
                        shape=tuple(int(v) for v in g["shape"]))


def test_rng_streams(device):
    g = load("rng")
    pop = CellPopulation.from_numpy(np.zeros((1, 3), np.float32), np.ones(1), capacity=int(g["capacity"]),
                                    device=device, seed=int(g["seed"]))
    np.testing.assert_array_equal(pop.rng_numpy(), g["rng_state"])


def test_deposit_accumulator(device):
    g = load("deposit")
    pop = CellPopulation.from_numpy(g["position"], np.full(len(g["position"]), 5.0), device=device,
                                    states=g["cell_state"])
    dens = CellDensities(_geom(g), device=device)
    dens.deposit(pop)
    np.testing.assert_array_equal(dens.accumulator.cpu().numpy(), g["accumulator"])


@pytest.mark.parametrize("shape_name", ["tile60x8x8", "packed/2", "packed/quad"])
def test_steady_state_solve(device, shape_name):
    from warpbiocell_amd.autotune import config_shapes

    g = load("solve")
    geom = _geom(g)
    previous = _lib.query("wbc_sor_fused_get_config")
    _lib.call("wbc_sor_fused_set_config", config_shapes().index(shape_name))
    try:
        f = ScalarField.create(geom, 40.0, device=device)
        rep = Dn.solve_steady_state(f, torch.from_numpy(g["density"]).to(device), UptakeKinetics(7.2e6, 5.538e7, 3.4),
                                    SolverSettings(1.8, 2000, 1e-6, 10))
    finally:
        _lib.call("wbc_sor_fused_set_config", previous)
    assert (rep.sweeps, np.float32(rep.residual), int(rep.converged)) == \
           (int(g["sweeps"]), g["residual"], int(g["converged"]))
    np.testing.assert_array_equal(f.numpy(), g["field"])


def test_hash_grid_contact_and_relaxation(device):
    g = load("contact")
    q = float(g["query_radius"])
    params = ContactParams(stiffness=1.0, damping=1.0, max_radius=9.0, margin=0.5)
    assert params.query_radius == q
    pop = CellPopulation.from_numpy(g["position"], g["radius"], device=device)
    grid = make_neighbor_grid(params, device)
    contact_substep(pop, grid, params, dt=0.0)
    np.testing.assert_array_equal(grid.point_ids().cpu().numpy(), g["point_ids"])
    starts = grid.bucket_starts().cpu().numpy()
    np.testing.assert_array_equal(np.flatnonzero(np.diff(starts)), g["bucket_start_nonzero"])
    np.testing.assert_array_equal(pop.velocity.cpu().numpy()[: len(g["position"])], g["velocity"])
    np.testing.assert_array_equal(pop.neighbor_counts_numpy(), g["neighbor_count"])
    moved = CellPopulation.from_numpy(g["position"], g["radius"], device=device)
    relax_contacts(moved, make_neighbor_grid(params, device),
                   ContactParams(stiffness=10.0, damping=1.0, max_radius=9.0, margin=0.5), 0.01, 5)
    np.testing.assert_array_equal(moved.positions_numpy(), g["relaxed_position"])
    np.testing.assert_array_equal(moved.neighbor_counts_numpy(), g["relaxed_count"])


def test_node_gradient_and_wall_push(device):
    g = load("confinement")
    n = len(g["position"])
    geom = GridGeometry(origin=tuple(float(v) for v in g["origin"]), dx=float(g["dx"]), shape=g["sdf"].shape)
    region = TissueRegion(geom, g["sdf"], device=device)
    np.testing.assert_array_equal(region.gradient.cpu().numpy(), g["gradient"])
    pos = torch.from_numpy(g["position"]).to(device)
    rad = torch.from_numpy(g["radius"]).to(device)
    vel = torch.from_numpy(g["velocity"]).to(device)
    _lib.call("wbc_confinement_velocities", region.struct_ptr, float(g["wall_rate"]), pos.data_ptr(), rad.data_ptr(),
              n, vel.data_ptr(), _lib.stream_of(device))
    np.testing.assert_array_equal(vel.cpu().numpy(), g["pushed"])


def test_three_full_cell_steps(device):
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", GOLDEN / "make_golden.py")
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)  # the case's parameters (the generator is plain Python on numpy)
    g = load("step")
    pos = g["initial_position"]
    n0 = len(pos)
    geom = GridGeometry.centered_cube(float(g["box"]), float(g["dx"]))
    ox_p = mk.STEP_OXYGEN
    ox = OxygenField(geom, OxygenParams(ox_p["D"], ox_p["q"], ox_p["K"], ox_p["O_b"]),
                     SolverSettings(omega=ox_p["omega"], tolerance=ox_p["tol"], max_sweeps=ox_p["max_sweeps"],
                                    check_every=ox_p["check"]), device=device)
    pop = CellPopulation.from_numpy(pos, np.full(n0, float(g["radius0"])), capacity=int(g["capacity"]),
                                    device=device, seed=int(g["seed"]))
    contact = ContactParams(**mk.STEP_CONTACT)
    grid = make_neighbor_grid(contact, device)
    lifecycle = LifecycleParams(**mk.STEP_LIFECYCLE)
    stepping = TimeStepping(dt_cells=0.1, dt_mechanics=0.025)
    daughters, sweeps = [], []
    for _ in range(3):
        rep = cell_step(pop, grid, lifecycle, contact, stepping, oxygen=ox)
        daughters.append(rep.daughters)
        sweeps.append(rep.field_sweeps)
    np.testing.assert_array_equal(daughters, g["daughters"])
    np.testing.assert_array_equal(sweeps, g["sweeps"])
    assert pop.count == int(g["count"])
    n = pop.count
    np.testing.assert_array_equal(pop.positions_numpy(), g["position"])
    np.testing.assert_array_equal(pop.cell_state.cpu().numpy()[:n], g["cell_state"])
    np.testing.assert_array_equal(pop.age.cpu().numpy()[:n], g["age"])
    np.testing.assert_array_equal(pop.oxygen_local.cpu().numpy()[:n], g["oxygen_local"])
    np.testing.assert_array_equal(pop.neighbor_counts_numpy(), g["neighbor_count"])
    np.testing.assert_array_equal(pop.rng_numpy(), g["rng_state"])
    np.testing.assert_array_equal(ox.numpy(), g["field"])
