"""Write the golden fixtures in this directory from the CPU oracle (the C restatement of the
reference's hot-path functions in oracle/). Run from the repository root:

    python tests/golden/make_golden.py

Each fixture is an .npz of inputs and expected outputs (numeric arrays only, no pickles). The CPU
tests (tests/test_golden.py) recompute them with the oracle and must match bit for bit, so a change
to the oracle shows; the GPU tests (tests/test_gpu_golden.py) run the HIP path on the stored inputs
and must match the stored outputs bit for bit. Every case is small (the whole set is under 0.1 MB).
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402
from warpbiocell_amd.cells.initialization import spherical_cluster  # noqa: E402
from warpbiocell_amd.cells.lifecycle import LifecycleParams  # noqa: E402
from warpbiocell_amd.cells.mechanics import ContactParams  # noqa: E402
from warpbiocell_amd.fields.scalar_field import GridGeometry  # noqa: E402
from warpbiocell_amd.geometry.shapes import Sphere, evaluate_on_grid  # noqa: E402
from warpbiocell_amd.simulation.simulator import TimeStepping  # noqa: E402

# Case parameters shared with the tests (the tests read them back from the files).
STEP_LIFECYCLE = dict(division_rate=0.6, death_rate=0.02, anoxic_death_rate=0.5, hypoxia_threshold=8.0,
                      death_threshold=2.0, inhibition_threshold=8)
STEP_CONTACT = dict(stiffness=10.0, damping=1.0, max_radius=5.0, margin=0.5)
STEP_OXYGEN = dict(D=7.2e6, q=5.538e7, K=3.4, O_b=40.0, omega=1.8, tol=1e-5, max_sweeps=2000, check=10)


def rng_case():
    states = O.init_rng(96, 7)
    seq, _ = O.randf_sequence(int(states[5]), 16)
    return dict(capacity=np.int64(96), seed=np.int64(7), rng_state=states, randf_from_slot5=seq)


def deposit_case():
    rng = np.random.default_rng(1)
    pos = rng.uniform(-45.0, 45.0, size=(80, 3)).astype(np.float32)
    states = rng.integers(0, 4, 80).astype(np.int32)
    geom = GridGeometry(origin=(-40.0, -40.0, -40.0), dx=10.0, shape=(9, 9, 9))
    acc = O.deposit_accumulator(pos, states, geom.origin, geom.dx, geom.shape)
    return dict(position=pos, cell_state=states, origin=np.array(geom.origin), dx=np.float64(geom.dx),
                shape=np.array(geom.shape), accumulator=acc)


def solve_case():
    geom = GridGeometry(origin=(-50.0, -50.0, -50.0), dx=10.0, shape=(11, 11, 11))
    pos = spherical_cluster(300, 5.0, spacing_factor=1.0, jitter=0.2, seed=2)
    acc = O.deposit_accumulator(pos, np.zeros(300, np.int32), geom.origin, geom.dx, geom.shape)
    density = O.combine((1.0, 1.0, 1.0, 0.0), O.fixed_point_to_density(acc, geom.dx))
    a = O.SolveArgs(dx=geom.dx, diffusion=STEP_OXYGEN["D"], uptake_max=STEP_OXYGEN["q"],
                    michaelis_k=STEP_OXYGEN["K"], boundary_value=STEP_OXYGEN["O_b"])
    field = np.full(geom.shape, STEP_OXYGEN["O_b"], dtype=np.float32)
    sweeps, residual, converged = O.solve(field, density, a, 1.8, 2000, 1e-6, 10)
    return dict(origin=np.array(geom.origin), dx=np.float64(geom.dx), shape=np.array(geom.shape), density=density,
                sweeps=np.int64(sweeps), residual=np.float32(residual), converged=np.int64(converged), field=field)


def contact_case():
    rng = np.random.default_rng(3)
    pos = spherical_cluster(400, 8.0, spacing_factor=0.85, jitter=0.3, seed=3)
    radii = rng.uniform(5.0, 9.0, 400).astype(np.float32)
    q = 2 * 9.0 + 0.5
    start, ids = O.hashgrid_build(pos, np.float32(q), 128)
    vel, cnt = O.contact_velocities(pos, radii, np.float32(q), 128, q, 1.0)
    moved, v5, c5 = pos.copy(), np.zeros_like(pos), np.zeros(400, np.int32)
    O.relax_contacts(moved, radii, np.float32(q), 128, q, 10.0, 0.01, 5, v5, c5)
    return dict(position=pos, radius=radii, query_radius=np.float64(q), point_ids=ids,
                bucket_start_nonzero=np.flatnonzero(np.diff(start)).astype(np.int64), velocity=vel,
                neighbor_count=cnt, relaxed_position=moved, relaxed_count=c5)


def confinement_case():
    geom = GridGeometry.centered_cube(240.0, 20.0)
    sdf = evaluate_on_grid(Sphere(radius=80.0), geom)
    region = O.OracleRegion.from_sdf(sdf, geom.origin, geom.dx)
    rng = np.random.default_rng(4)
    d = rng.normal(size=(150, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = (d * rng.uniform(40.0, 120.0, (150, 1))).astype(np.float32)
    radii = rng.uniform(5.0, 9.0, 150).astype(np.float32)
    vel = rng.normal(size=(150, 3)).astype(np.float32)
    pushed = vel.copy()
    O.confinement_velocities(region, 12.0, pos, radii, pushed)
    return dict(origin=np.array(geom.origin), dx=np.float64(geom.dx), sdf=sdf, gradient=region.gradient,
                position=pos, radius=radii, velocity=vel, wall_rate=np.float64(12.0), pushed=pushed)


def step_case():
    n0 = 250
    pos = spherical_cluster(n0, 5.0, spacing_factor=1.0, jitter=0.1, seed=6)
    radii = np.full(n0, 5.0, np.float32)
    geom = GridGeometry.centered_cube(200.0, 10.0)
    lifecycle = LifecycleParams(**STEP_LIFECYCLE)
    contact = ContactParams(**STEP_CONTACT)
    stepping = TimeStepping(dt_cells=0.1, dt_mechanics=0.025)
    ref = O.OraclePopulation.from_numpy(pos, radii, capacity=4 * n0, seed=11)
    a = O.SolveArgs(dx=geom.dx, diffusion=STEP_OXYGEN["D"], uptake_max=STEP_OXYGEN["q"],
                    michaelis_k=STEP_OXYGEN["K"], boundary_value=STEP_OXYGEN["O_b"])
    rox = O.OracleOxygen(shape=geom.shape, origin=geom.origin, args=a, omega=STEP_OXYGEN["omega"],
                         max_sweeps=STEP_OXYGEN["max_sweeps"], tolerance=STEP_OXYGEN["tol"],
                         check_every=STEP_OXYGEN["check"])
    daughters, sweeps = [], []
    for _ in range(3):
        d, f = O.oracle_cell_step(ref, lifecycle, contact, stepping, oxygen=rox)
        daughters.append(d)
        sweeps.append(f[0])
    n = ref.count
    return dict(initial_position=pos, radius0=np.float64(5.0), capacity=np.int64(4 * n0), seed=np.int64(11),
                box=np.float64(200.0), dx=np.float64(10.0), daughters=np.array(daughters, np.int64),
                sweeps=np.array(sweeps, np.int64), count=np.int64(n), position=ref.position[:n].copy(),
                cell_state=ref.cell_state[:n].copy(), age=ref.age[:n].copy(),
                oxygen_local=ref.oxygen_local[:n].copy(), neighbor_count=ref.neighbor_count[:n].copy(),
                rng_state=ref.rng_state.copy(), field=rox.values.copy())


CASES = {"rng": rng_case, "deposit": deposit_case, "solve": solve_case, "contact": contact_case,
         "confinement": confinement_case, "step": step_case}


def main() -> None:
    O.set_threads(1)
    for name, make in CASES.items():
        np.savez_compressed(HERE / f"{name}.npz", **make())
        print(f"wrote {name}.npz")


if __name__ == "__main__":
    main()
