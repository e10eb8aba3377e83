"""GPU parity of the field path (deposit, density, SOR, residual, solve loop, sampling) against
the oracle on the same inputs. The HIP kernels and the oracle evaluate the same expressions with
one rounding per operation, so integer AND float results are compared bit for bit; the error
against a float64 solve is reported alongside."""

import numpy as np
import pytest
import torch

from warpbiocell_amd import workloads as W
from warpbiocell_amd.cells.initialization import spherical_cluster
from warpbiocell_amd.cells.state import CellPopulation
from warpbiocell_amd.fields import diffusion as Dn
from warpbiocell_amd.fields.densities import CellDensities
from warpbiocell_amd.fields.diffusion import SolverSettings, UptakeKinetics
from warpbiocell_amd.fields.oxygen import OxygenField, OxygenParams
from warpbiocell_amd.fields.scalar_field import Boundary, GridGeometry, ScalarField

pytestmark = pytest.mark.gpu

MIXED = (Boundary.DIRICHLET, Boundary.NEUMANN, Boundary.DIRICHLET)


def _cluster_with_states(n, seed, device, outside=0):
    rng = np.random.default_rng(seed)
    pos = spherical_cluster(n, 8.0, spacing_factor=0.9, jitter=0.3, seed=seed)
    if outside:
        pos[:outside] += np.float32(5000.0)
    states = rng.integers(0, 4, n).astype(np.int32)
    return pos, states, CellPopulation.from_numpy(pos, np.full(n, 8.0), capacity=n + 17, device=device, states=states)


def test_deposit_accumulator_is_bit_exact_and_order_independent(oracle, device):
    pos, states, pop = _cluster_with_states(5000, 21, device, outside=37)
    geom = GridGeometry.centered_cube(600.0, 20.0)
    dens = CellDensities(geom, device=device)
    expected = oracle.deposit_accumulator(pos, states, geom.origin, geom.dx, geom.shape)
    for _ in range(3):
        dens.deposit(pop)
        np.testing.assert_array_equal(dens.accumulator.cpu().numpy(), expected)
    np.testing.assert_array_equal(dens.density.cpu().numpy(), oracle.fixed_point_to_density(expected, geom.dx))


def test_deposit_layer_tracking_follows_states_that_come_and_go(oracle, device):
    """The deposit zeroes and reads only the accumulator layers some deposit wrote (a bit per
    state on the device). Cells move between states across deposits: first all in one state, then
    spread over three, then all in the fourth; every accumulator, per-state density and combined
    density equals the oracle's, so a layer the cells left is zeroed and a layer never written
    stays zero. Empty and zero-cell deposits keep the layers' bits."""
    pos, _, pop = _cluster_with_states(3000, 31, device, outside=11)
    geom = GridGeometry.centered_cube(600.0, 20.0)
    dens = CellDensities(geom, device=device)
    rng = np.random.default_rng(5)
    weights = (1.0, 0.25, 0.5, 2.0)
    for step, states in enumerate((np.zeros(3000, np.int32), rng.choice([0, 1, 2], 3000).astype(np.int32),
                                   np.full(3000, 3, np.int32), np.full(3000, 1, np.int32))):
        pop.cell_state[:3000].copy_(torch.from_numpy(states))
        expected = oracle.deposit_accumulator(pos, states, geom.origin, geom.dx, geom.shape)
        combined = torch.empty(geom.shape, dtype=torch.float32, device=device)
        dens.deposit_combined(pop, weights, combined)
        np.testing.assert_array_equal(dens.accumulator.cpu().numpy(), expected, err_msg=f"deposit {step}")
        per_state = oracle.fixed_point_to_density(expected, geom.dx)
        np.testing.assert_array_equal(combined.cpu().numpy(), oracle.combine(weights, per_state),
                                      err_msg=f"combined {step}")
        np.testing.assert_array_equal(dens.density.cpu().numpy(), per_state, err_msg=f"density {step}")
    assert int(dens._layers.item()) == 0b1111
    pop.count = 0
    dens.deposit(pop)
    assert not dens.accumulator.any()
    assert int(dens._layers.item()) == 0b1111


def test_deposit_in_any_cell_order_gives_the_same_accumulator(oracle, device):
    """wbc_deposit_trilinear_i64_ordered visits cells in a given permutation of the live slots:
    reversed, random and the hash grid's bucket order (recorded by a contact relaxation) all give
    the oracle's accumulator bit for bit. Adding cells invalidates the recorded order, and the
    deposit falls back to slot order."""
    from warpbiocell_amd import _lib
    from warpbiocell_amd.cells.lifecycle import LifecycleParams, commit_divisions
    from warpbiocell_amd.cells.mechanics import ContactParams, make_neighbor_grid, relax_contacts

    pos, states, pop = _cluster_with_states(5000, 23, device, outside=41)
    geom = GridGeometry.centered_cube(600.0, 20.0)
    dens = CellDensities(geom, device=device)
    expected = oracle.deposit_accumulator(pos, states, geom.origin, geom.dx, geom.shape)
    nx, ny, nz = geom.shape
    origin, inv_dx, _ = dens._scalars()
    acc = torch.empty_like(dens.accumulator)
    perms = {"reversed": np.arange(5000)[::-1], "random": np.random.default_rng(4).permutation(5000)}
    for name, perm in perms.items():
        order = torch.from_numpy(np.ascontiguousarray(perm, dtype=np.int32)).to(device)
        acc.zero_()
        _lib.call("wbc_deposit_trilinear_i64_ordered", pop.position.data_ptr(), pop.cell_state.data_ptr(),
                  order.data_ptr(), 5000, origin[0], origin[1], origin[2], inv_dx, nx, ny, nz, acc.data_ptr(),
                  pop.stream)
        np.testing.assert_array_equal(acc.cpu().numpy(), expected, err_msg=name)

    assert pop.deposit_order() is None
    params = ContactParams(stiffness=1.0, damping=1.0, max_radius=8.0, margin=0.5)
    relax_contacts(pop, make_neighbor_grid(params, device), params, dt=0.01, substeps=2)
    order = pop.deposit_order()
    assert order is not None
    assert sorted(order[:5000].cpu().numpy().tolist()) == list(range(5000))
    assert not np.array_equal(order[:5000].cpu().numpy(), np.arange(5000))
    moved = pop.positions_numpy()
    dens.deposit(pop)
    np.testing.assert_array_equal(dens.accumulator.cpu().numpy(),
                                  oracle.deposit_accumulator(moved, states, geom.origin, geom.dx, geom.shape))

    pop.divide_flag[:5000] = 0
    pop.divide_flag[:3] = 1
    assert commit_divisions(pop, LifecycleParams()) == 3
    assert pop.deposit_order() is None
    dens.deposit(pop)
    grown = pop.positions_numpy()
    np.testing.assert_array_equal(dens.accumulator.cpu().numpy(),
                                  oracle.deposit_accumulator(grown, pop.cell_state.cpu().numpy()[:5003], geom.origin,
                                                             geom.dx, geom.shape))


def test_fused_density_combine_is_bit_exact(oracle, device):
    pos, states, pop = _cluster_with_states(3000, 5, device)
    geom = GridGeometry.centered_cube(500.0, 12.5)
    dens = CellDensities(geom, device=device)
    out = torch.zeros(geom.shape, dtype=torch.float32, device=device)
    w = (1.0, 0.75, 0.5, 0.0)
    dens.deposit_combined(pop, w, out)
    per_state = oracle.fixed_point_to_density(oracle.deposit_accumulator(pos, states, geom.origin, geom.dx,
                                                                         geom.shape), geom.dx)
    np.testing.assert_array_equal(out.cpu().numpy(), oracle.combine(w, per_state))
    np.testing.assert_array_equal(dens.density.cpu().numpy(), per_state)  # lazily materialised
    out2 = torch.zeros_like(out)
    dens.combine((0.3, 0.0, 1.0, 2.0), out2)
    np.testing.assert_array_equal(out2.cpu().numpy(), oracle.combine((0.3, 0.0, 1.0, 2.0), per_state))


def test_sampling_is_bit_exact_and_clamps(oracle, device):
    geom = GridGeometry.centered_cube(200.0, 25.0)
    rng = np.random.default_rng(1)
    values = rng.uniform(0.0, 50.0, geom.shape).astype(np.float32)
    pos = rng.uniform(-130.0, 130.0, size=(4000, 3)).astype(np.float32)
    ox = OxygenField(geom, OxygenParams(), boundary=(Boundary.NEUMANN,) * 3, device=device)
    ox.field.assign(values)
    pop = CellPopulation.from_numpy(pos, np.full(4000, 8.0), device=device)
    ox.sample_at_cells(pop)
    np.testing.assert_array_equal(pop.oxygen_numpy(), oracle.sample(values, pos, geom.origin, geom.dx))


@pytest.mark.parametrize("boundary", [(0, 0, 0), (0, 1, 0), (1, 1, 1)])
@pytest.mark.parametrize("pinned", [False, True])
def test_colour_passes_and_residual_are_bit_exact(oracle, device, boundary, pinned):
    geom = GridGeometry(origin=(0.0, 0.0, 0.0), dx=10.0, shape=(19, 23, 37))
    rng = np.random.default_rng(7)
    values = rng.uniform(5.0, 20.0, geom.shape).astype(np.float32)
    density = rng.uniform(0.0, 2.0e-3, geom.shape).astype(np.float32)
    production = rng.uniform(0.0, 10.0, geom.shape).astype(np.float32) if pinned else None
    fixed = (rng.random(geom.shape) < 0.05) if pinned else None
    kin = UptakeKinetics(2.0e3, 5.0e4, 3.0)
    f = ScalarField.create(geom, 20.0, boundary=tuple(Boundary(b) for b in boundary), device=device)
    f.assign(values)
    if pinned:
        f.set_fixed(fixed)
    ref = f.numpy()
    a = oracle.SolveArgs(dx=10.0, diffusion=2.0e3, uptake_max=5.0e4, michaelis_k=3.0, boundary_value=20.0,
                         boundary=boundary)
    d_t = torch.from_numpy(density).to(device)
    p_t = None if production is None else torch.from_numpy(production).to(device)
    fx = None if fixed is None else fixed.astype(np.int32)
    for sweep in range(6):
        for color in (0, 1):
            Dn.sor_sweep(f, d_t, kin, 1.7, color, production=p_t)
            oracle.sor_sweep(ref, density, a, 1.7, color, fixed=fx, production=production)
        np.testing.assert_array_equal(f.numpy(), ref, err_msg=f"sweep {sweep}")
    res_buf = torch.zeros(1, dtype=torch.float32, device=device)
    gpu_res = Dn.steady_state_residual(f, d_t, kin, res_buf, production=p_t)
    assert np.float32(gpu_res) == np.float32(oracle.residual(ref, density, a, fixed=fx, production=production))


@pytest.mark.parametrize("shape_name", ["tile60x8x8", "copy60x4x8", "packed/2", "packed/8", "packed/quad",
                                        "packed/quad2", "packed/quad4"])
def test_solve_loop_matches_the_oracle_report_and_field(oracle, device, shape_name):
    """The whole chunked loop: same sweeps, same residual, same field bits (41^3 spheroid), with
    the solve loop's sweeps on a one-launch tile and on the colour-packed arrays."""
    from warpbiocell_amd import _lib
    from warpbiocell_amd.autotune import config_shapes

    previous = _lib.query("wbc_sor_fused_get_config")
    _lib.call("wbc_sor_fused_set_config", config_shapes().index(shape_name))
    try:
        _solve_loop_matches(oracle, device)
    finally:
        _lib.call("wbc_sor_fused_set_config", previous)


def _solve_loop_matches(oracle, device):
    w = W.spheroid_1k_41()
    geom = GridGeometry.centered_cube(w.box_um, w.dx_um)
    settings = SolverSettings(omega=1.8, tolerance=1e-6, max_sweeps=2000, check_every=10)
    params = OxygenParams(W.D_UM2_PER_H, W.UPTAKE_MMHG_UM3_PER_H, W.MICHAELIS_K, 40.0)
    ox = OxygenField(geom, params, settings, device=device)
    pop = CellPopulation.from_numpy(w.positions, w.radii, device=device)
    a = oracle.SolveArgs(dx=geom.dx, diffusion=W.D_UM2_PER_H, uptake_max=W.UPTAKE_MMHG_UM3_PER_H,
                         michaelis_k=W.MICHAELIS_K, boundary_value=40.0)
    ref = oracle.OracleOxygen(shape=geom.shape, origin=geom.origin, args=a, omega=1.8, max_sweeps=2000,
                              tolerance=1e-6, check_every=10)
    opop = oracle.OraclePopulation.from_numpy(w.positions, w.radii)
    for step in range(3):
        report = ox.update(pop)
        expected = ref.update(opop)
        assert (report.sweeps, np.float32(report.residual), report.converged) == \
               (expected[0], np.float32(expected[1]), expected[2]), f"step {step}"
        np.testing.assert_array_equal(ox.numpy(), ref.values)
        np.testing.assert_array_equal(pop.oxygen_numpy(), opop.oxygen_local[: opop.count])
    # Error of the FP32 field against an FP64 solve of the same discrete problem.
    f64 = ref.values.astype(np.float64)
    oracle.solve(f64, ref.density.astype(np.float64), a, 1.8, 4000, 1e-12, 10)
    err = np.abs(ox.numpy() - f64).max() / 40.0
    print(f"41^3 field: max |fp32 - fp64| / O_b = {err:.3e}")
    assert err < 1e-4


@pytest.mark.parametrize("shape_name", ["tile60x8x8", "packed/2", "packed/4", "packed/8", "packed/quad",
                                        "packed/quad2", "packed/quad4"])
def test_max_sweeps_and_pipelined_checks(oracle, device, shape_name):
    """Budget exhaustion mid-chunk, a last chunk shorter than the others, one sweep per check and
    convergence inside a queued-ahead chunk, on the one-launch tile and on the colour-packed loop
    (whose checks ride in the colour passes and whose red field alternates between two arrays)."""
    from warpbiocell_amd import _lib
    from warpbiocell_amd.autotune import config_shapes

    geom = GridGeometry.centered_cube(300.0, 20.0)
    kin = UptakeKinetics(7.2e6, 1.4e8, 3.4)
    a = oracle.SolveArgs(dx=20.0, diffusion=7.2e6, uptake_max=1.4e8, michaelis_k=3.4, boundary_value=38.0)
    rng = np.random.default_rng(11)
    previous = _lib.query("wbc_sor_fused_get_config")
    _lib.call("wbc_sor_fused_set_config", config_shapes().index(shape_name))
    try:
        for max_sweeps, tol, check, dense in ((25, 1e-9, 10, False), (2000, 1e-5, 7, False), (3, 1e-5, 1, False),
                                              (1, 1e-9, 1, True), (7, 1e-9, 3, True), (60, 2e-3, 1, True),
                                              (2000, 1e-6, 10, True), (13, 1e-9, 5, True)):
            f = ScalarField.create(geom, 38.0, initial_value=0.0, device=device)
            ref = f.numpy()
            density = (rng.random(geom.shape) * 2e-4).astype(np.float32) if dense else np.zeros(geom.shape, np.float32)
            rep = Dn.solve_steady_state(f, torch.from_numpy(density).to(device), kin,
                                        SolverSettings(1.8, max_sweeps, tol, check))
            exp = oracle.solve(ref, density, a, 1.8, max_sweeps, tol, check)
            case = (shape_name, max_sweeps, tol, check, dense)
            assert (rep.sweeps, np.float32(rep.residual), rep.converged) == \
                   (exp[0], np.float32(exp[1]), exp[2]), case
            np.testing.assert_array_equal(f.numpy(), ref, err_msg=str(case))
    finally:
        _lib.call("wbc_sor_fused_set_config", previous)


def test_cosh_profile_on_the_gpu(device):
    L, nx = 200.0, 81
    geom = GridGeometry(origin=(-L, 0.0, 0.0), dx=2 * L / (nx - 1), shape=(nx, 3, 3))
    D, kappa, K = 1.0e4, 3.0 / L, 1.0e9
    params = OxygenParams(diffusion_coefficient=D, uptake_max=kappa ** 2 * D * K, michaelis_k=K, boundary_value=1.0)
    ox = OxygenField(geom, params, SolverSettings(omega=1.5, tolerance=1e-5, max_sweeps=20000, check_every=10),
                     boundary=(Boundary.DIRICHLET, Boundary.NEUMANN, Boundary.NEUMANN), device=device)
    ox.density.fill_(1.0)
    report = ox.solve_steady_state()
    x = geom.node_coordinates()[0]
    assert report.converged
    np.testing.assert_allclose(ox.numpy()[:, 1, 1], np.cosh(kappa * x) / np.cosh(kappa * L), atol=2e-3)


@pytest.mark.parametrize("shape", [(19, 23, 37), (9, 9, 33), (17, 10, 66), (3, 3, 3), (41, 41, 41), (35, 30, 130)])
@pytest.mark.parametrize("pinned", [False, True])
def test_fused_sweep_equals_two_colour_passes(oracle, device, shape, pinned):
    """One fused launch per sweep (src -> dst) is bit-identical to the two in-place colour passes."""
    from warpbiocell_amd import _lib

    rng = np.random.default_rng(sum(shape) + pinned)
    values = rng.uniform(0.0, 40.0, shape).astype(np.float32)
    density = np.where(rng.random(shape) < 0.5, rng.uniform(0.0, 2e-3, shape), 0.0).astype(np.float32)
    production = rng.uniform(0.0, 10.0, shape).astype(np.float32) if pinned else None
    fixed = (rng.random(shape) < 0.05).astype(np.int32) if pinned else None
    a = oracle.SolveArgs(dx=10.0, diffusion=2.0e3, uptake_max=5.0e4, michaelis_k=3.0, boundary_value=20.0)
    bufs = [torch.from_numpy(values).to(device), torch.zeros(shape, dtype=torch.float32, device=device)]
    d_t = torch.from_numpy(density).to(device)
    p_t = None if production is None else torch.from_numpy(production).to(device)
    f_t = None if fixed is None else torch.from_numpy(fixed).to(device)
    refs = [values.copy()]
    for sweep in range(4):
        r = refs[-1].copy()
        for color in (0, 1):
            oracle.sor_sweep(r, density, a, 1.8, color, fixed=fixed, production=production)
        refs.append(r)
    from warpbiocell_amd.autotune import config_shapes

    nbytes = _lib.query("wbc_sor_sweeps_workspace_bytes", *shape)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
    for cfg, name in enumerate(config_shapes()):
        if name.startswith("packed") and pinned:
            continue  # the colour-packed entry runs only without pinned nodes and production
        if not name.startswith("packed"):
            # one launch per sweep, src -> dst
            bufs[0].copy_(torch.from_numpy(values))
            for sweep in range(4):
                _lib.call("wbc_sor_sweep_fused_cfg", cfg, 1.8, a.dx2_over_D, a.uptake_max, a.michaelis_k,
                          _lib.ptr(f_t), d_t.data_ptr(), _lib.ptr(p_t), 0, None, 1.0, bufs[sweep % 2].data_ptr(),
                          bufs[(sweep + 1) % 2].data_ptr(), *shape, _lib.stream_of(device))
                np.testing.assert_array_equal(bufs[(sweep + 1) % 2].cpu().numpy(), refs[sweep + 1],
                                              err_msg=f"config {name}, sweep {sweep}")
        # several sweeps in one call (packing once for the colour-packed entry)
        for sweeps in (1, 2, 3):
            bufs[0].copy_(torch.from_numpy(values))
            _lib.call("wbc_sor_sweeps_cfg", cfg, sweeps, 1.8, a.dx2_over_D, a.uptake_max, a.michaelis_k,
                      _lib.ptr(f_t), d_t.data_ptr(), _lib.ptr(p_t), 0, None, 1.0, bufs[0].data_ptr(),
                      bufs[1].data_ptr(), *shape, ws.data_ptr(), nbytes, _lib.stream_of(device))
            np.testing.assert_array_equal(bufs[1].cpu().numpy(), refs[sweeps],
                                          err_msg=f"config {name}, {sweeps} sweeps in one call")


def test_depleted_core_solve_is_bit_exact(oracle, device):
    """Strong uptake drives the core to (near) zero: exercises the tiny-value division paths."""
    geom = GridGeometry.centered_cube(300.0, 20.0)
    params = OxygenParams(uptake_max=1.0e12)
    a = oracle.SolveArgs(dx=20.0, diffusion=params.diffusion_coefficient, uptake_max=1.0e12,
                         michaelis_k=params.michaelis_k, boundary_value=params.boundary_value)
    ox = OxygenField(geom, params, SolverSettings(omega=1.9, tolerance=1e-5, check_every=10), device=device)
    density = np.zeros(geom.shape, np.float32)
    density[3:-3, 3:-3, 3:-3] = 1e-3
    ox.density.copy_(torch.from_numpy(density))
    ref = ox.numpy()
    report = ox.solve_steady_state()
    expected = oracle.solve(ref, density, a, 1.9, 2000, 1e-5, 10)
    assert (report.sweeps, report.converged) == (expected[0], expected[2])
    np.testing.assert_array_equal(ox.numpy(), ref)
    assert ref.min() >= 0.0 and ref[8, 8, 8] < 1e-3
