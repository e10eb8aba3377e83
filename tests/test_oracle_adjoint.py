"""Pin the FP64 adjoint oracle against central finite differences of the FP64 forward solve
(step 1e-3 relative), on a small spheroid: 64 random per-cell rates, D and the boundary value."""

import numpy as np

from warpbiocell_amd import workloads as W
from warpbiocell_amd.fields.scalar_field import GridGeometry


def _problem():
    rng = np.random.default_rng(7)
    n = 300
    pos = W.ball_uniform(n, 70.0, rng)
    rates = (W.UPTAKE_MMHG_UM3_PER_H * np.exp(0.3 * rng.standard_normal(n))).astype(np.float64)
    targets = rng.uniform(0.0, 40.0, n)
    geom = GridGeometry.centered_cube(400.0, 16.0)
    return pos, rates, targets, geom


def test_adjoint_gradients_match_central_differences(oracle):
    pos, rates, targets, geom = _problem()
    D, Ob, K = W.D_UM2_PER_H, 40.0, 1.0
    args = dict(michaelis_k=K, origin=geom.origin, dx=geom.dx, shape=geom.shape)

    def loss(q, d, ob, start):
        f, Q, S, rep = oracle.forward_f64(pos, q, d, ob, start=start, **args)
        assert rep[2], rep
        return float(np.sum((S - targets) ** 2)), f, Q, S

    L0, field, Q, S = loss(rates, D, Ob, None)
    gq, gD, gOb, _ = oracle.adjoint_f64(pos, 2.0 * (S - targets), field, Q, D, K, geom.origin, geom.dx, geom.shape)
    assert np.abs(S - targets).max() > 1.0 and np.abs(gq).max() > 0.0  # a non-trivial problem
    pick = np.random.default_rng(1).choice(len(rates), 64, replace=False)
    for c in pick:
        eps = 1e-3 * rates[c]
        qp, qm = rates.copy(), rates.copy()
        qp[c] += eps
        qm[c] -= eps
        fd = (loss(qp, D, Ob, field)[0] - loss(qm, D, Ob, field)[0]) / (2 * eps)
        assert abs(fd - gq[c]) <= 2e-3 * abs(gq).max(), (c, fd, gq[c])
    for name, grad, fn in (("D", gD, lambda e: loss(rates, D + e, Ob, field)[0]),
                           ("O_b", gOb, lambda e: loss(rates, D, Ob + e, None)[0])):
        eps = 1e-3 * (D if name == "D" else Ob)
        fd = (fn(eps) - fn(-eps)) / (2 * eps)
        assert abs(fd - grad) <= 2e-3 * abs(fd), (name, fd, grad)


def test_warm_started_central_differences_match_the_adjoint(oracle):
    """fd_rate_gradients (fixed-sweep solves from the converged field at the optimal omega, loss
    difference summed term by term) against the FP64 adjoint and against central differences of
    fully converged solves, on the small spheroid."""
    pos, rates, targets, geom = _problem()
    D, Ob, K = W.D_UM2_PER_H, 40.0, 1.0
    args = dict(michaelis_k=K, origin=geom.origin, dx=geom.dx, shape=geom.shape)
    field, Q, S, rep = oracle.forward_f64(pos, rates, D, Ob, tolerance=1e-13, omega=oracle.optimal_omega(geom.shape),
                                          **args)
    assert rep[2]
    gq, _, _, _ = oracle.adjoint_f64(pos, 2.0 * (S - targets), field, Q, D, K, geom.origin, geom.dx, geom.shape)
    cells = np.random.default_rng(2).choice(len(rates), 16, replace=False)
    fd = oracle.fd_rate_gradients(pos, rates, targets, D, Ob, K, geom.origin, geom.dx, geom.shape, cells, field,
                                  sweeps=300)
    scale = np.abs(gq).max()
    assert np.abs(fd - gq[cells]).max() <= 1e-6 * scale
    c = cells[0]
    eps = 1e-3 * rates[c]
    ls = []
    for sign in (1, -1):
        q = rates.copy()
        q[c] += sign * eps
        ls.append(float(np.sum((oracle.forward_f64(pos, q, D, Ob, tolerance=1e-13, start=field, **args)[2] - targets)
                               ** 2)))
    assert abs((ls[0] - ls[1]) / (2 * eps) - fd[0]) <= 1e-4 * scale
