"""The contact gather's prefilter keeps every candidate the exact test keeps.

The gather first lists candidates whose squared distance, summed with fused multiply-adds, is at
most q^2 (1 + 2^-18), then applies the reference's test sqrt(d^2) < q (unfused sum, rounded square
root) to the listed ones. This checks the superset property on candidates placed at the query
radius to within a few units in the last place, in float32 arithmetic emulated exactly on the host
(a product of two float32 values is exact in float64; the fused sums are rounded once)."""

import numpy as np


def _f32(x):
    return np.float32(x)


def _fma32(a, b, c):
    # a*b is exact in float64; the sum is rounded to float64 and then to float32, which equals the
    # single rounding except on exact float64 ties, far from the bound this test probes.
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def test_prefilter_bound_is_a_superset_of_the_exact_test():
    rng = np.random.default_rng(7)
    kept = 0
    for q in (np.float32(16.5), np.float32(15.4321), np.float32(3.0), np.float32(1000.0)):
        bound = q * q * np.float32(1.0 + 2.0 ** -18)
        d = rng.normal(size=(40000, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        scale = q * (1.0 + rng.uniform(-4e-7, 4e-7, size=(40000, 1)))
        p = rng.uniform(-200.0, 200.0, size=(40000, 3)).astype(np.float32)
        qpos = (p - d * scale).astype(np.float32)
        for a, b in zip(p, qpos):
            dx, dy, dz = _f32(a[0] - b[0]), _f32(a[1] - b[1]), _f32(a[2] - b[2])
            s = _f32(_f32(_f32(dx * dx) + _f32(dy * dy)) + _f32(dz * dz))
            if np.sqrt(s, dtype=np.float32) < q:
                kept += 1
                fused = _fma32(dz, dz, _fma32(dy, dy, _f32(dx * dx)))
                assert fused <= bound, (q, dx, dy, dz)
    assert kept > 1000
