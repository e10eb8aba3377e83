"""Synthetic spheroid workloads (numpy only) used by the benchmark, the smoke check and tests.

Physics shared by every workload: D = 2000 um^2/s = 7.2e6 um^2/h; uptake 20 amol/s per cell
= 2e-17 mol/s * 3600 s/h / (1.3e-21 mol um^-3 mmHg^-1) = 5.538e7 mmHg um^3/h; the uptake is
zero-order in intent, written as Michaelis-Menten with K = 1e-3 mmHg (the solver requires
K > 0); contact rate k/gamma = 10 /h; dt_cells = 0.1 h with dt_mechanics = 0.025 h (4 contact
substeps per step). Cells are spheres with uniformly random centres in a ball.

  spheroid_1k_41      1,000 cells of radius 5 um in a 60 um ball; 41^3 grid over 800 um;
                      omega 1.8, tolerance 1e-6, check every 10, at most 2000 sweeps.
  spheroid_100k_161   100,000 cells, radii log-normal (median 5 um, sigma 0.1), ball sized for
                      0.6 volume fraction (R = 279 um); 161^3 grid over 1600 um; same solver.
  spheroid_1m_257     1,000,000 cells, radii uniform in [4, 6] um, jittered lattice packed at
                      0.64 volume fraction (R = 588 um, Gaussian jitter 0.5 um); 257^3 grid over
                      1280 um; fixed 20 sweeps per step; 5 % of cells divide per step.
  adjoint_100k_129    100,000 cells as spheroid_100k_161 on a 129^3 grid over 1280 um, per-cell
                      uptake log-normal (median 5.538e7, sigma 0.3), targets uniform in [0, 40].
  sweep_15            15 replicas = boundary {20, 40, 60, 100, 150} mmHg x seeds {0, 1, 2}; cell
                      counts log-uniform in [2,000, 200,000] and grids drawn from {41, 81, 121,
                      161}^3 over a 1600 um box (one draw per replica from a fixed generator);
                      cells as spheroid_100k_161; solver as spheroid_1k_41; 100 steps; the
                      lifecycle of the reference's sweep base (configs/tumor_spheroid.yaml:31-38:
                      division 0.0289 /h, anoxic death 0.5 /h below 2 mmHg, hypoxia below 8 mmHg,
                      contact inhibition at 8 neighbours), capacity 1.4 n + 64 for the growth.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

D_UM2_PER_H = 7.2e6
UPTAKE_MMHG_UM3_PER_H = 2.0e-17 * 3600.0 / 1.3e-21
MICHAELIS_K = 1.0e-3
CONTACT_RATE = 10.0
DT_CELLS = 0.1
DT_MECHANICS = 0.025


@dataclass
class Workload:
    name: str
    positions: np.ndarray  # (n, 3) float32
    radii: np.ndarray  # (n,) float32
    box_um: float
    dx_um: float
    boundary_mmHg: float
    omega: float = 1.8
    tolerance: float = 1.0e-6
    check_every: int = 10
    max_sweeps: int = 2000
    steps: int = 100
    division_rate: float = 0.0
    capacity: int | None = None
    seed: int = 0
    grid_nodes: int | None = None  # nodes per axis when set (dx = box / (nodes - 1)), else from dx
    lifecycle: dict | None = None  # LifecycleParams keyword arguments, else divisions only (no deaths)

    @property
    def n(self) -> int:
        return len(self.radii)

    @property
    def grid_n(self) -> int:
        if self.grid_nodes is not None:
            return int(self.grid_nodes)
        return int(math.ceil(self.box_um / self.dx_um)) + 1


def ball_uniform(n: int, radius: float, rng: np.random.Generator) -> np.ndarray:
    """Uniform points in a ball: Gaussian direction, radius R u^(1/3)."""
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius * rng.random(n) ** (1.0 / 3.0)
    return (d * r[:, None]).astype(np.float32)


def ball_radius_for_fraction(n: int, mean_r3: float, fraction: float) -> float:
    return (n * mean_r3 / fraction) ** (1.0 / 3.0)


def spheroid_1k_41(seed: int = 0) -> Workload:
    rng = np.random.default_rng(seed)
    pos = ball_uniform(1000, 60.0, rng)
    return Workload("spheroid_1k_41", pos, np.full(1000, 5.0, dtype=np.float32), 800.0, 20.0, 40.0, steps=50,
                    seed=seed)


def spheroid_100k_161(seed: int = 0, n: int = 100_000) -> Workload:
    rng = np.random.default_rng(seed)
    sigma = 0.1
    radii = (5.0 * np.exp(sigma * rng.standard_normal(n))).astype(np.float32)
    mean_r3 = 125.0 * math.exp(4.5 * sigma ** 2)
    R = ball_radius_for_fraction(n, mean_r3, 0.6)
    pos = ball_uniform(n, R, rng)
    return Workload("spheroid_100k_161", pos, radii, 1600.0, 10.0, 40.0, steps=100, seed=seed)


def spheroid_1m_257(seed: int = 0, n: int = 1_000_000) -> Workload:
    rng = np.random.default_rng(seed)
    radii = rng.uniform(4.0, 6.0, n).astype(np.float32)
    mean_r3 = (6.0 ** 4 - 4.0 ** 4) / (4.0 * 2.0)
    R = ball_radius_for_fraction(n, mean_r3, 0.64)
    spacing = (4.0 / 3.0 * math.pi * mean_r3 / 0.64) ** (1.0 / 3.0)
    half = int(math.ceil(R / spacing)) + 2
    ticks = np.arange(-half, half + 1) * spacing
    lattice = np.stack(np.meshgrid(ticks, ticks, ticks, indexing="ij"), -1).reshape(-1, 3)
    keep = np.argsort(np.einsum("ij,ij->i", lattice, lattice), kind="stable")[:n]
    pos = (lattice[keep] + rng.normal(0.0, 0.5, (n, 3))).astype(np.float32)
    daughters_cap = int(n * 1.05 ** 21) + 1024
    return Workload("spheroid_1m_257", pos, radii, 1280.0, 5.0, 40.0, steps=20, max_sweeps=20, check_every=20,
                    tolerance=1e-30, division_rate=-math.log(0.95) / DT_CELLS, capacity=daughters_cap, seed=seed)


def adjoint_100k_129(seed: int = 0, n: int = 100_000):
    w = spheroid_100k_161(seed, n)
    w.name = "adjoint_100k_129"
    w.box_um, w.dx_um = 1280.0, 10.0
    rng = np.random.default_rng(seed + 1000)
    rates = (UPTAKE_MMHG_UM3_PER_H * np.exp(0.3 * rng.standard_normal(n))).astype(np.float32)
    targets = rng.uniform(0.0, 40.0, n).astype(np.float32)
    return w, rates, targets


SWEEP_BOUNDARIES = (20.0, 40.0, 60.0, 100.0, 150.0)
# The lifecycle block of the sweep's base experiment (configs/tumor_spheroid.yaml:31-38), in
# LifecycleParams keywords.
SWEEP_LIFECYCLE = dict(division_rate=0.0289, death_rate=0.0, anoxic_death_rate=0.5, inhibition_threshold=8,
                       hypoxia_threshold=8.0, death_threshold=2.0, placement_factor=1.0)
SWEEP_SEEDS = (0, 1, 2)
SWEEP_GRIDS = (41, 81, 121, 161)


def sweep_15(draw_seed: int = 0, box_um: float = 1600.0) -> list[Workload]:
    """The replicas of the 15-run parameter sweep, in the order of the reference's sweep runs
    (grid value outer, seed inner) and labelled as it labels them (``boundary_mmHg=20_seed=0``)."""
    rng = np.random.default_rng(10_000 + draw_seed)
    out = []
    for ob in SWEEP_BOUNDARIES:
        for seed in SWEEP_SEEDS:
            n = int(round(math.exp(rng.uniform(math.log(2_000), math.log(200_000)))))
            g = int(rng.choice(SWEEP_GRIDS))
            w = spheroid_100k_161(seed=seed, n=n)
            w.name = f"boundary_mmHg={ob:g}_seed={seed}"
            w.boundary_mmHg = ob
            w.box_um, w.dx_um, w.grid_nodes = box_um, box_um / (g - 1), g
            w.lifecycle = dict(SWEEP_LIFECYCLE)
            w.division_rate = SWEEP_LIFECYCLE["division_rate"]
            w.capacity = int(1.4 * n) + 64
            out.append(w)
    return out
