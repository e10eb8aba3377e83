"""Parity at the benchmark's full size (100k cells, 161^3 grid) through checks that stay cheap
for the oracle: a few fused sweeps compared bit for bit, the residual the GPU reports recomputed
by the oracle on the GPU's own final field, deposit conservation, and the contact gather."""

import numpy as np
import pytest
import torch

from warpbiocell_amd import _lib
from warpbiocell_amd import workloads as W
from warpbiocell_amd.cells.mechanics import ContactParams, contact_substep, make_neighbor_grid
from warpbiocell_amd.cells.state import CellPopulation
from warpbiocell_amd.fields.diffusion import SolverSettings
from warpbiocell_amd.fields.oxygen import OxygenField, OxygenParams
from warpbiocell_amd.fields.scalar_field import GridGeometry

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(device):
    w = W.spheroid_100k_161()
    geom = GridGeometry.centered_cube(w.box_um, w.dx_um)
    params = OxygenParams(W.D_UM2_PER_H, W.UPTAKE_MMHG_UM3_PER_H, W.MICHAELIS_K, w.boundary_mmHg)
    settings = SolverSettings(omega=w.omega, tolerance=w.tolerance, max_sweeps=w.max_sweeps, check_every=w.check_every)
    pop = CellPopulation.from_numpy(w.positions, w.radii, device=device)
    ox = OxygenField(geom, params, settings, device=device)
    return w, geom, params, settings, pop, ox


def _args(oracle, geom, params):
    return oracle.SolveArgs(dx=geom.dx, diffusion=params.diffusion_coefficient, uptake_max=params.uptake_max,
                            michaelis_k=params.michaelis_k, boundary_value=params.boundary_value)


def test_deposit_conserves_cells_at_full_size(oracle, c1):
    w, geom, params, settings, pop, ox = c1
    ox.deposit_uptake(pop)
    acc = ox.densities.accumulator.cpu().numpy()
    inside = geom.contains(w.positions)
    total = int(acc.sum())
    assert abs(total - int(inside.sum()) * 2 ** 24) <= 4 * int(inside.sum())
    expected = oracle.deposit_accumulator(w.positions, np.zeros(w.n, np.int32), geom.origin, geom.dx, geom.shape)
    np.testing.assert_array_equal(acc, expected)


def test_fused_sweeps_bit_exact_at_full_size(oracle, c1, device):
    w, geom, params, settings, pop, ox = c1
    ox.deposit_uptake(pop)
    a = _args(oracle, geom, params)
    density = ox.density.cpu().numpy()
    start = ox.numpy()
    ref = start.copy()
    for sweep in range(2):
        for color in (0, 1):
            oracle.sor_sweep(ref, density, a, settings.omega, color)
    nbytes = _lib.query("wbc_sor_sweeps_workspace_bytes", *geom.shape)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
    for cfg in range(_lib.query("wbc_sor_fused_config_count")):
        bufs = [torch.from_numpy(start).to(device), torch.empty_like(ox.field.values)]
        _lib.call("wbc_sor_sweeps_cfg", cfg, 2, settings.omega, a.dx2_over_D, a.uptake_max, a.michaelis_k, None,
                  ox.density.data_ptr(), None, 0, None, 1.0, bufs[0].data_ptr(), bufs[1].data_ptr(), *geom.shape,
                  ws.data_ptr(), nbytes, _lib.stream_of(device))
        np.testing.assert_array_equal(bufs[1].cpu().numpy(), ref, err_msg=f"config {cfg}")


def test_reported_residual_is_the_oracle_residual_of_the_gpu_field(oracle, c1):
    w, geom, params, settings, pop, ox = c1
    report = ox.update(pop)
    assert report.converged and report.sweeps % settings.check_every == 0
    field = ox.numpy()
    again = oracle.residual(field, ox.density.cpu().numpy(), _args(oracle, geom, params))
    assert np.float32(again) == np.float32(report.residual) and report.residual < settings.tolerance
    o = pop.oxygen_numpy()
    assert o.min() >= 0.0 and o.max() <= params.boundary_value
    np.testing.assert_array_equal(o, oracle.sample(field, w.positions, geom.origin, geom.dx))
    print(f"161^3 cold solve: {report.sweeps} sweeps, residual {report.residual:.3e}, core min {field.min():.3e}")


def test_contact_gather_bit_exact_at_full_size(oracle, c1, device):
    w, geom, params, settings, pop, ox = c1
    contact = ContactParams(stiffness=W.CONTACT_RATE, damping=1.0, max_radius=float(w.radii.max()))
    contact_substep(pop, make_neighbor_grid(contact, device), contact, dt=0.0)
    vel, cnt = oracle.contact_velocities(w.positions, w.radii, np.float32(contact.query_radius), 128,
                                         contact.query_radius, contact.rate)
    np.testing.assert_array_equal(pop.neighbor_counts_numpy(), cnt)
    np.testing.assert_array_equal(pop.velocity.cpu().numpy()[: w.n], vel)
