"""The quad colour pass skips the density load of lanes outside the box of nonzero density that
the pack computes. These cases put the nonzero density where the box is tightest or degenerate
(none at all, one interior corner node, the Dirichlet faces only, negative zeros everywhere, the
last packed index of a row, infinite uptake) and compare the multi-sweep entry and the solve loop
on `packed/quad` with the oracle bit for bit."""

import numpy as np
import pytest
import torch

from warpbiocell_amd.fields import diffusion as Dn
from warpbiocell_amd.fields.diffusion import SolverSettings, UptakeKinetics
from warpbiocell_amd.fields.scalar_field import GridGeometry, ScalarField

pytestmark = pytest.mark.gpu

SHAPES = [(19, 23, 37), (17, 10, 66), (33, 33, 33), (9, 12, 16)]


def _density(pattern, shape, rng):
    d = np.zeros(shape, np.float32)
    nx, ny, nz = shape
    if pattern == "empty":
        pass
    elif pattern == "low_corner":
        d[1, 1, 1] = 3e-3
    elif pattern == "high_corner":
        d[nx - 2, ny - 2, nz - 2] = 3e-3
    elif pattern == "faces_only":
        d[0, :, :] = 2e-3
        d[:, -1, :] = 2e-3
        d[:, :, 0] = 2e-3
    elif pattern == "negative_zero_and_cluster":
        d[:] = np.float32(-0.0)
        c = (nx // 2, ny // 2, nz // 2)
        d[c[0] - 1:c[0] + 2, c[1] - 1:c[1] + 2, c[2] - 1:c[2] + 2] = 1e-3
    elif pattern == "row_ends":
        d[nx // 2, 1:-1, nz - 2] = 2e-3
        d[1:-1, ny // 3, 1] = 2e-3
    elif pattern == "sparse_random":
        d[:] = np.where(rng.random(shape) < 0.01, rng.uniform(0.0, 2e-3, shape), 0.0)
    else:
        raise ValueError(pattern)
    return d


PATTERNS = ["empty", "low_corner", "high_corner", "faces_only", "negative_zero_and_cluster", "row_ends",
            "sparse_random"]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("uptake_max", [5.0e4, np.inf])
@pytest.mark.parametrize("shape_name", ["packed/quad", "packed/quad2", "packed/quad4"])
def test_quad_sweeps_with_the_density_box_match_the_oracle(oracle, device, shape, pattern, uptake_max, shape_name):
    from warpbiocell_amd import _lib
    from warpbiocell_amd.autotune import config_shapes

    rng = np.random.default_rng(sum(shape))
    values = rng.uniform(0.0, 40.0, shape).astype(np.float32)
    density = _density(pattern, shape, rng)
    a = oracle.SolveArgs(dx=10.0, diffusion=2.0e3, uptake_max=float(uptake_max), michaelis_k=3.0,
                         boundary_value=20.0)
    refs = [values.copy()]
    for _ in range(3):
        r = refs[-1].copy()
        for color in (0, 1):
            oracle.sor_sweep(r, density, a, 1.8, color)
        refs.append(r)
    cfg = config_shapes().index(shape_name)
    nbytes = _lib.query("wbc_sor_sweeps_workspace_bytes", *shape)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
    src = torch.from_numpy(values).to(device)
    dst = torch.zeros(shape, dtype=torch.float32, device=device)
    d_t = torch.from_numpy(density).to(device)
    for sweeps in (1, 3):
        _lib.call("wbc_sor_sweeps_cfg", cfg, sweeps, 1.8, a.dx2_over_D, a.uptake_max, a.michaelis_k, None,
                  d_t.data_ptr(), None, 0, None, 1.0, src.data_ptr(), dst.data_ptr(), *shape, ws.data_ptr(), nbytes,
                  _lib.stream_of(device))
        np.testing.assert_array_equal(dst.cpu().numpy(), refs[sweeps],
                                      err_msg=f"{shape_name} {pattern} {shape} uptake {uptake_max}, {sweeps} sweeps")


@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("shape_name", ["packed/quad", "packed/quad2"])
def test_quad_solve_loop_with_the_density_box_matches_the_oracle(oracle, device, pattern, shape_name):
    from warpbiocell_amd import _lib
    from warpbiocell_amd.autotune import config_shapes

    geom = GridGeometry.centered_cube(300.0, 20.0)  # 16^3 nodes
    kin = UptakeKinetics(7.2e6, 1.4e8, 3.4)
    a = oracle.SolveArgs(dx=20.0, diffusion=7.2e6, uptake_max=1.4e8, michaelis_k=3.4, boundary_value=38.0)
    rng = np.random.default_rng(5)
    density = _density(pattern, geom.shape, rng) * np.float32(0.1)
    previous = _lib.query("wbc_sor_fused_get_config")
    _lib.call("wbc_sor_fused_set_config", config_shapes().index(shape_name))
    try:
        for max_sweeps, tol, check in ((2000, 1e-6, 10), (13, 1e-9, 5)):
            f = ScalarField.create(geom, 38.0, initial_value=0.0, device=device)
            ref = f.numpy()
            rep = Dn.solve_steady_state(f, torch.from_numpy(density).to(device), kin,
                                        SolverSettings(1.8, max_sweeps, tol, check))
            exp = oracle.solve(ref, density, a, 1.8, max_sweeps, tol, check)
            case = (shape_name, pattern, max_sweeps, tol, check)
            assert (rep.sweeps, np.float32(rep.residual), rep.converged) == \
                   (exp[0], np.float32(exp[1]), exp[2]), case
            np.testing.assert_array_equal(f.numpy(), ref, err_msg=str(case))
    finally:
        _lib.call("wbc_sor_fused_set_config", previous)
