"""GPU adjoint of the steady-state solve: against the FP64 adjoint oracle (itself pinned to
central differences on the CPU), through torch.autograd, and against FP32 central differences
(step 1e-3 relative) at the full 100k-cell, 129^3 size."""

import time

import numpy as np
import pytest
import torch

from warpbiocell_amd import workloads as W
from warpbiocell_amd.adjoint import AdjointSettings, OxygenAdjointProblem, sampled_oxygen
from warpbiocell_amd.fields.diffusion import SolverSettings
from warpbiocell_amd.fields.scalar_field import GridGeometry

pytestmark = pytest.mark.gpu


def _moderate(seed=3, n=3000):
    rng = np.random.default_rng(seed)
    pos = W.ball_uniform(n, 150.0, rng)
    rates = (W.UPTAKE_MMHG_UM3_PER_H * np.exp(0.3 * rng.standard_normal(n))).astype(np.float32)
    targets = rng.uniform(0.0, 40.0, n).astype(np.float32)
    return pos, rates, targets, GridGeometry.centered_cube(600.0, 12.5)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_gradients_match_the_fp64_adjoint(oracle, device):
    pos, rates, targets, geom = _moderate()
    D, Ob, K = W.D_UM2_PER_H, 40.0, 1.0
    prob = OxygenAdjointProblem(geom, torch.from_numpy(pos), K, W.UPTAKE_MMHG_UM3_PER_H,
                                SolverSettings(omega=1.8, tolerance=1e-6, max_sweeps=20000, check_every=10),
                                AdjointSettings(), device=device)
    S = prob.forward(torch.from_numpy(rates), D, Ob)
    gq, gD, gOb = prob.backward(2.0 * (S - torch.from_numpy(targets).to(device)))
    f64, Q64, S64, rep = oracle.forward_f64(pos, rates.astype(np.float64), D, Ob, K, geom.origin, geom.dx, geom.shape)
    assert rep[2]
    rq, rD, rOb, _ = oracle.adjoint_f64(pos, 2.0 * (S64 - targets), f64, Q64, D, K, geom.origin, geom.dx, geom.shape,
                                        sweeps=6000)
    errs = {"S": _rel(S.cpu().numpy(), S64), "dq": _rel(gq.cpu().numpy(), rq), "dD": abs(gD - rD) / abs(rD),
            "dOb": abs(gOb - rOb) / abs(rOb)}
    print("relative error vs FP64 adjoint:", {k: f"{v:.2e}" for k, v in errs.items()},
          "forward sweeps", prob.last_forward.sweeps, "adjoint sweeps", prob.last_backward[0])
    assert errs["S"] < 1e-4 and errs["dq"] < 2e-3 and errs["dD"] < 2e-3 and errs["dOb"] < 2e-3


def test_packed_adjoint_sweeps_equal_the_grid_layout_ones(device):
    """wbc_adjoint_solve on the colour-packed quad lanes and on the grid layout, from the same
    forward field, rates and g: mu, the rate gradients, both sums, the sweep count and the residual
    are equal bit for bit, also with pinned interior nodes (mu held at 0) and with a budget that
    stops the solve unconverged."""
    import ctypes

    from warpbiocell_amd import _lib
    from warpbiocell_amd.fields.diffusion import float32_threshold

    pos, rates, targets, geom = _moderate(n=2000)
    prob = OxygenAdjointProblem(geom, torch.from_numpy(pos), 1.0, W.UPTAKE_MMHG_UM3_PER_H, device=device)
    S = prob.forward(torch.from_numpy(rates), W.D_UM2_PER_H, 40.0)
    prob.backward(2.0 * (S - torch.from_numpy(targets).to(device)))  # leaves g in prob._g
    g = prob._g.clone()
    state = prob.state
    nx, ny, nz = geom.shape
    h = geom.dx ** 2 / state.diffusion
    omega = 2.0 / (1.0 + np.sin(np.pi / (max(geom.shape) - 1)))
    rng = np.random.default_rng(7)
    fixed = np.zeros(geom.shape, np.int32)
    fixed[tuple(rng.integers(1, n - 1, 400) for n in geom.shape)] = 1
    fixed_d = torch.from_numpy(fixed).to(device)
    s = torch.cuda.current_stream(device).cuda_stream

    def solve(packed, pinned, max_sweeps):
        _lib.call("wbc_adjoint_set_packed", packed)
        mu = torch.full(geom.shape, 7.0, dtype=torch.float32, device=device)  # overwritten entirely
        grad = torch.empty(prob.n, dtype=torch.float32, device=device)
        sums = (ctypes.c_double * 2)()
        sweeps, residual = ctypes.c_int32(), ctypes.c_float()
        _lib.call("wbc_adjoint_solve", omega, max_sweeps, float32_threshold(1e-6), 10, h, prob.q_ref,
                  prob.michaelis_k, fixed_d.data_ptr() if pinned else None, state.field.values.data_ptr(),
                  state.rate_density.data_ptr(), None, g.data_ptr(), mu.data_ptr(), prob.positions.data_ptr(),
                  prob.n, geom.origin[0], geom.origin[1], geom.origin[2], geom.dx, nx, ny, nz, grad.data_ptr(),
                  ctypes.byref(sums), ctypes.byref(sweeps), ctypes.byref(residual), prob._ws.data_ptr(),
                  prob._ws.numel(), s)
        torch.cuda.synchronize()
        return mu.cpu().numpy(), grad.cpu().numpy(), tuple(sums), sweeps.value, residual.value

    try:
        for pinned, budget in ((False, 20000), (True, 20000), (False, 35)):
            print(f"adjoint layouts: pinned {pinned}, max_sweeps {budget}", flush=True)
            a, b = solve(1, pinned, budget), solve(0, pinned, budget)
            np.testing.assert_array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
            np.testing.assert_array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
            assert a[2:] == b[2:], (a[2:], b[2:])
            if pinned:
                assert (a[0][fixed == 1] == 0.0).all()
            assert a[3] > 0 and (budget != 35 or a[3] == 35)
    finally:
        _lib.call("wbc_adjoint_set_packed", 1)


def test_two_forwards_then_both_backwards(device):
    """Two sampled_oxygen calls on one problem (different D and O_b) before either backward: each
    gradient is that of its own forward, equal to a forward+backward done alone within the adjoint
    solve's tolerance (the sampling adjoint scatters dL/dS into the grid with float atomics, whose
    order moves the last bits); the two forwards' gradients differ by far more."""
    pos, rates, targets, geom = _moderate(n=800)
    prob = OxygenAdjointProblem(geom, torch.from_numpy(pos), 1.0, W.UPTAKE_MMHG_UM3_PER_H, device=device)
    t = torch.from_numpy(targets).to(device)
    cases = [(W.D_UM2_PER_H, 40.0), (0.5 * W.D_UM2_PER_H, 30.0)]

    def alone(D, Ob):
        q = torch.from_numpy(rates).to(device).requires_grad_(True)
        Dt = torch.tensor(D, dtype=torch.float64, requires_grad=True)
        Bt = torch.tensor(Ob, dtype=torch.float64, requires_grad=True)
        ((sampled_oxygen(q, Dt, Bt, prob) - t) ** 2).sum().backward()
        return q.grad.cpu().numpy(), float(Dt.grad), float(Bt.grad)

    expected = [alone(D, Ob) for D, Ob in cases]
    leaves, losses = [], []
    for D, Ob in cases:
        q = torch.from_numpy(rates).to(device).requires_grad_(True)
        Dt = torch.tensor(D, dtype=torch.float64, requires_grad=True)
        Bt = torch.tensor(Ob, dtype=torch.float64, requires_grad=True)
        leaves.append((q, Dt, Bt))
        losses.append(((sampled_oxygen(q, Dt, Bt, prob) - t) ** 2).sum())
    for loss in losses:
        loss.backward()
    for (q, Dt, Bt), (eq, eD, eB) in zip(leaves, expected):
        assert _rel(q.grad.cpu().numpy(), eq) < 1e-4
        assert abs(float(Dt.grad) - eD) <= 1e-4 * abs(eD) and abs(float(Bt.grad) - eB) <= 1e-4 * abs(eB)
    assert _rel(expected[0][0], expected[1][0]) > 0.05


def test_autograd_function_routes_the_adjoint(device):
    pos, rates, targets, geom = _moderate(n=800)
    prob = OxygenAdjointProblem(geom, torch.from_numpy(pos), 1.0, W.UPTAKE_MMHG_UM3_PER_H, device=device)
    q = torch.from_numpy(rates).to(device).requires_grad_(True)
    D = torch.tensor(W.D_UM2_PER_H, dtype=torch.float64, requires_grad=True)
    Ob = torch.tensor(40.0, dtype=torch.float64, requires_grad=True)
    S = sampled_oxygen(q, D, Ob, prob)
    loss = ((S - torch.from_numpy(targets).to(device)) ** 2).sum()
    loss.backward()
    gq, gD, gOb = prob.backward(2.0 * (S.detach() - torch.from_numpy(targets).to(device)))
    np.testing.assert_allclose(q.grad.cpu().numpy(), gq.cpu().numpy(), rtol=1e-5, atol=1e-30)
    assert D.grad.item() == pytest.approx(gD, rel=1e-6) and Ob.grad.item() == pytest.approx(gOb, rel=1e-6)


def test_full_size_gradients_against_central_differences(device):
    w, rates, targets = W.adjoint_100k_129()
    geom = GridGeometry.centered_cube(w.box_um, w.dx_um)
    D, Ob = W.D_UM2_PER_H, w.boundary_mmHg
    prob = OxygenAdjointProblem(geom, torch.from_numpy(w.positions), W.MICHAELIS_K, W.UPTAKE_MMHG_UM3_PER_H,
                                SolverSettings(omega=1.8, tolerance=1e-6, max_sweeps=20000, check_every=10),
                                device=device)
    T = torch.from_numpy(targets).to(device)
    q = torch.from_numpy(rates).to(device)

    def loss(qq, d, ob):
        S = prob.forward(qq, d, ob)
        return float(((S.double() - T.double()) ** 2).sum()), S

    L0, S = loss(q, D, Ob)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L0, S = loss(q, D, Ob)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    gq, gD, gOb = prob.backward(2.0 * (S - T))
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    errs = {}
    for name, grad, fn, base in (("D", gD, lambda e: loss(q, D + e, Ob)[0], D),
                                 ("O_b", gOb, lambda e: loss(q, D, Ob + e)[0], Ob)):
        eps = 1e-3 * base
        fd = (fn(eps) - fn(-eps)) / (2 * eps)
        errs[name] = abs(fd - grad) / abs(fd)
    # Directional central difference over all 100k rates (step 1e-3 relative, positive random
    # weights so that the contributions add up instead of cancelling below the FP32 solve's noise).
    v = torch.from_numpy(np.random.default_rng(5).uniform(0.5, 1.5, w.n).astype(np.float32)).to(device)
    dq = 1e-3 * q * v
    fd = (loss(q + dq, D, Ob)[0] - loss(q - dq, D, Ob)[0]) / 2.0
    adj = float((gq.double() * dq.double()).sum())
    errs["rates (directional)"] = abs(fd - adj) / abs(fd)
    print(f"129^3 adjoint: forward {t_fwd * 1e3:.1f} ms ({prob.last_forward.sweeps} sweeps), backward "
          f"{t_bwd * 1e3:.1f} ms ({prob.last_backward[0]} sweeps), ratio {t_bwd / t_fwd:.2f}; FD relative errors "
          + ", ".join(f"{k} {e:.2e}" for k, e in errs.items()))
    assert all(e < 2e-2 for e in errs.values()), errs


def test_full_size_gradients_against_the_fp64_adjoint_and_central_differences(oracle, device):
    """The 100k-cell, 129^3 problem: the GPU gradient (all rates, D, O_b) against the float64
    adjoint at the float64 solution, and on 64 random rates against central differences of float64
    solves (step 1e-3 relative), which also pin the float64 adjoint at this size."""
    w, rates, targets = W.adjoint_100k_129()
    geom = GridGeometry.centered_cube(w.box_um, w.dx_um)
    D, Ob, K = W.D_UM2_PER_H, w.boundary_mmHg, W.MICHAELIS_K
    prob = OxygenAdjointProblem(geom, torch.from_numpy(w.positions), K, W.UPTAKE_MMHG_UM3_PER_H,
                                SolverSettings(omega=w.omega, tolerance=w.tolerance, max_sweeps=w.max_sweeps,
                                               check_every=w.check_every), device=device)
    S = prob.forward(torch.from_numpy(rates).to(device), D, Ob)
    gq, gD, gOb = prob.backward(2.0 * (S - torch.from_numpy(targets).to(device)))
    oracle.set_threads(oracle.get_threads())
    r = oracle.adjoint_check(w.positions, rates, targets, D, Ob, K, geom.origin, geom.dx, geom.shape,
                             gq.cpu().numpy(), gD, gOb, fd_cells=64)
    fd_fp64 = np.asarray(r["fd_vs_fp64_rel_err"])
    gpu_fd = np.asarray(r["gradient_vs_fd_rel_err"])
    print("129^3 / 100k gradient check:", {k: (f"{v:.3e}" if isinstance(v, float) else v) for k, v in r.items()
                                            if not isinstance(v, list)},
          f"FD vs FP64 adjoint per rate: max {fd_fp64.max():.2e}; GPU vs FD per rate: median "
          f"{np.median(gpu_fd):.2e}, max {gpu_fd.max():.2e}")
    # Measured on MI355X: FD vs FP64 adjoint 6.4e-5 at most; GPU vs FP64 adjoint 7.4e-5 (2-norm),
    # 3.1e-4 of the largest gradient at most, D 7.3e-5, O_b 9.4e-5; GPU vs FD median 9.7e-5.
    assert fd_fp64.max() < 1e-3  # the differences pin the float64 adjoint rate by rate
    assert r["rates_rel_l2"] < 1e-3 and r["rates_max_err_over_max_grad"] < 2e-3
    assert r["rates_rel_err_p99"] < 1e-2
    assert r["D_rel_err"] < 1e-3 and r["O_b_rel_err"] < 1e-3
    assert np.median(gpu_fd) < 1e-3 and gpu_fd.max() < 1e-2
