"""Independent numpy restatements used to pin the oracle (brute force, dense solves, closed
forms). Written in float64 from the mathematics, not from the oracle's code."""

from __future__ import annotations

import numpy as np


def trilinear_deposit_f64(positions, consuming, origin, dx, shape):
    """Cells / dx^3 from unit trilinear weights (the reference's deposit_reference)."""
    u = (np.asarray(positions, dtype=np.float64) - np.asarray(origin)) / dx
    base = np.floor(u).astype(int)
    frac = u - base
    out = np.zeros(shape)
    for c in range(len(u)):
        i0, j0, k0 = base[c]
        if min(i0, j0, k0) < 0 or i0 + 1 >= shape[0] or j0 + 1 >= shape[1] or k0 + 1 >= shape[2]:
            continue
        if not consuming[c]:
            continue
        fx, fy, fz = frac[c]
        for di, wx in ((0, 1 - fx), (1, fx)):
            for dj, wy in ((0, 1 - fy), (1, fy)):
                for dk, wz in ((0, 1 - fz), (1, fz)):
                    out[i0 + di, j0 + dj, k0 + dk] += wx * wy * wz
    return out / dx ** 3


def _mirror(idx, n):
    return 1 if idx < 0 else (n - 2 if idx > n - 1 else idx)


def dirichlet_mask(shape, boundary):
    mask = np.zeros(shape, dtype=bool)
    for axis, b in enumerate(boundary):
        if b == 0:
            idx = [slice(None)] * 3
            idx[axis] = 0
            mask[tuple(idx)] = True
            idx[axis] = -1
            mask[tuple(idx)] = True
    return mask


def steady_state_dense(values0, density, D, q, K, dx, boundary, picard=40, tol=1e-12, fixed_extra=None):
    """Discrete steady state by a dense linear solve inside a Picard loop on the uptake; nodes of
    ``fixed_extra`` (and the Dirichlet faces) keep their ``values0``."""
    v = np.asarray(values0, dtype=np.float64).copy()
    dens = np.asarray(density, dtype=np.float64)
    shape = v.shape
    fixed = dirichlet_mask(shape, boundary)
    if fixed_extra is not None:
        fixed = fixed | np.asarray(fixed_extra, dtype=bool)
    unknown = np.argwhere(~fixed)
    index = -np.ones(shape, dtype=int)
    for row, (i, j, k) in enumerate(unknown):
        index[i, j, k] = row
    m = len(unknown)
    nx, ny, nz = shape
    for _ in range(picard):
        a = np.zeros((m, m))
        rhs = np.zeros(m)
        for row, (i, j, k) in enumerate(unknown):
            c = dens[i, j, k] * q / (K + v[i, j, k])
            a[row, row] = -(6.0 + dx ** 2 * c / D)
            for nb in ((_mirror(i - 1, nx), j, k), (_mirror(i + 1, nx), j, k), (i, _mirror(j - 1, ny), k),
                       (i, _mirror(j + 1, ny), k), (i, j, _mirror(k - 1, nz)), (i, j, _mirror(k + 1, nz))):
                col = index[nb]
                if col >= 0:
                    a[row, col] += 1.0
                else:
                    rhs[row] -= v[nb]
        sol = np.linalg.solve(a, rhs)
        new = v.copy()
        new[~fixed] = sol
        change = np.abs(new - v).max()
        v = new
        if change < tol:
            break
    return v


def contact_brute_force(positions, radii, rate, query_radius, eps=1e-6):
    x = np.asarray(positions, dtype=np.float64)
    r = np.asarray(radii, dtype=np.float64)
    d_vec = x[:, None, :] - x[None, :, :]
    d = np.linalg.norm(d_vec, axis=2)
    np.fill_diagonal(d, np.inf)
    near = d < query_radius
    overlap = r[:, None] + r[None, :] - d
    active = near & (overlap > 0.0) & (d > eps)
    with np.errstate(divide="ignore", invalid="ignore"):
        coef = np.where(active, rate * overlap / d, 0.0)
    return np.einsum("ij,ijk->ik", coef, d_vec), near.sum(axis=1).astype(np.int32)


def pcg_py(v: int) -> int:
    """PCG-RXS-M-XS 32-bit hash with Python integers (Jarzynski & Olano 2020, listing pcg_hash)."""
    state = (v * 747796405 + 2891336453) % 2 ** 32
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) % 2 ** 32
    return (word >> 22) ^ word


def residual_terms_f64(field, density, dx2_over_D, uptake_max, michaelis_k, reference, boundary, fixed=None,
                       production=None, source=None, source_k=1.0):
    """Per-node residual terms |r| / max(6 O, O_ref) of kernels/field_kernels.py:215-243 in float64,
    vectorised over the whole grid, and the mask of nodes the residual runs over (not pinned, not
    on a Dirichlet face). r = sum of the six mirrored neighbours + dx^2/D P S/(K_S + S)
    - (6 + dx^2/D rho q / (K + O)) O; the mirror across a face is numpy's reflect padding
    (value at -1 = value at 1)."""
    f = np.asarray(field, dtype=np.float64)
    pad = np.pad(f, 1, mode="reflect")
    nb = (pad[:-2, 1:-1, 1:-1] + pad[2:, 1:-1, 1:-1] + pad[1:-1, :-2, 1:-1] + pad[1:-1, 2:, 1:-1]
          + pad[1:-1, 1:-1, :-2] + pad[1:-1, 1:-1, 2:])
    c = np.asarray(density, dtype=np.float64) * uptake_max / (michaelis_k + f)
    prod = np.zeros_like(f)
    if production is not None:
        prod = np.asarray(production, dtype=np.float64)
        if source is not None:
            s = np.asarray(source, dtype=np.float64)
            prod = prod * (s / (source_k + s))
    r = nb + dx2_over_D * prod - (6.0 + dx2_over_D * c) * f
    terms = np.abs(r) / np.maximum(6.0 * f, reference)
    mask = ~dirichlet_mask(f.shape, boundary)
    if fixed is not None:
        mask &= np.asarray(fixed) == 0
    scale = (np.abs(nb) + np.abs(dx2_over_D * prod) + np.abs((6.0 + dx2_over_D * c) * f)) / np.maximum(6.0 * f,
                                                                                                     reference)
    return terms, mask, scale
