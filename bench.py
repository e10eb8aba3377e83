#!/usr/bin/env python3
"""Benchmark: cell-steps/s of the full hot-path step on the 100k-cell, 161^3 spheroid.

A step is one ``cell_step`` (simulation/simulator.py:58-81): deposit -> red-black SOR to a
max-norm residual of 1e-6 (omega 1.8, checked every 10 sweeps) -> sampling -> lifecycle
decisions and slot allocation -> 4 contact substeps (hash-grid build, gather, integrate).
Inputs are synthetic (``warpbiocell_amd.workloads.spheroid_100k_161``) and resident on the GPU
before the timed region. One JSON line goes to stdout; progress goes to stderr.

Roofline: the SOR shape the autotuner selects (on the headline grid, the two colour-pass
launches of `packed/quad`), timed live with HIP events on the stream it runs on over 1000 sweeps,
priced at 8 B per node per launch (16 B per sweep: each of the two colour passes reads the field
once, the sweep writes it once and reads the density once) against MI355X's 8 TB/s HBM peak. The
quad passes load the density only inside the box of nonzero density, so they move 12 B per node
plus 16 B per lane inside it: `achieved_moved` / `frac_moved` price the same time at those bytes.
`--sweep-grid N` times the same on an N^3 grid.

CPU leg (the only part of the benchmark that uses the oracle, the C restatement of the reference,
and only as the checker): the oracle runs the first CPU_STEPS steps of the timed region from the
exact state they started at, on every host core the process may use (and the first two on one
thread); the GPU re-runs the same steps from a device copy of that state, so both sides are timed
on the same inputs and sweeps. The same leg reports accuracy against float64: the final headline
field against a float64 solve of the same discrete problem started from it, the contact gather at
the final positions against a float64 restatement over every pair, and the gradient of the 129^3
adjoint workload against the float64 adjoint and central differences of float64 solves.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0
CPU_STEPS = 10  # timed steps the CPU leg repeats on the oracle and the GPU
METRIC = "cell-steps/s @100k cells,161\u00b3 SOR; batched-sweep & adjoint speedup vs plain"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def setup(workload_name: str, device, grid_nodes: int | None = None):
    from warpbiocell_amd import workloads as W
    from warpbiocell_amd.simulation import scenario

    w = getattr(W, workload_name)()
    if grid_nodes is not None:  # the same cells and box on grid_nodes^3 nodes
        w.grid_nodes = grid_nodes
    sc = scenario.build(w, device)
    return (sc.workload, sc.geometry, sc.params, sc.settings, sc.lifecycle, sc.contact, sc.stepping,
            sc.population, sc.oxygen, sc.grid)


KERNEL_OF_KIND = {"tile": "sor_fused_kernel<{tj}, {ti}>", "march": "sor_march_kernel<{tj}, {ti}>",
                  "pipe": "sor_pipe_kernel<{tj}, {ti}>", "copy": "sor_copy_kernel<{tj}, {ti}>",
                  "packed": "sor_packed_kernel<C, {tj}, 0>", "quad": "sor_quad_kernel<C, 0, {q}>"}
COUNTER_TABLE = ROOT / "profiles" / "sor_counters_spheroid_100k_161.md"


def measured_traffic(shape: str, workload: str, grid) -> int | None:
    """HBM bytes of one red+black sweep of tile shape `shape` (both launches of a two-pass shape)
    from the committed FETCH_SIZE / WRITE_SIZE table (scaled by the copy calibration), or None when
    the table does not cover this workload/shape."""
    if workload != "spheroid_100k_161" or tuple(grid) != (161, 161, 161) or not COUNTER_TABLE.exists():
        return None
    for line in COUNTER_TABLE.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 6 and cells[0] == shape:
            return int(cells[5])
    return None


def quad_box_lanes(density, shape) -> int:
    """Lanes of one quad colour pass that load their density: those inside the box of the nodes
    whose density is not +-0 (NaN counts), in the pass's lane coordinates (plane i - 1, row j - 1,
    group k // 8), clipped to the interior lanes. Each loads 16 B."""
    import torch

    nx, ny, nz = shape
    idx = torch.nonzero(density)
    if idx.numel() == 0:
        return 0
    lo, hi = idx.min(0).values.tolist(), idx.max(0).values.tolist()
    groups = (nz // 2 + 3) // 4
    spans = [(lo[0] - 1, hi[0] - 1, nx - 3), (lo[1] - 1, hi[1] - 1, ny - 3), (lo[2] // 8, hi[2] // 8, groups - 1)]
    lanes = 1
    for a, b, top in spans:
        lanes *= max(0, min(b, top) - max(a, 0) + 1)
    return lanes


def sor_roofline(ox, geom, device, workload: str, timed_sweeps: int = 1000) -> dict:
    """Average duration of one red+black SOR sweep with the selected shape, measured with HIP
    events on its stream over `timed_sweeps` sweeps in one multi-sweep call.

    Algorithmic bytes of one red+black sweep, the `achieved` basis: 16 B per node (each colour
    pass reads the other colour and its own, writes its own and reads its density). The quad passes
    read the density only inside the box of nonzero density (16 B per lane inside it,
    `quad_box_lanes`), so the bytes they must move are 12 B per node plus those loads
    (`bytes_moved_per_sweep`); `achieved_moved` and `frac_moved` price the same time at those. A
    fused shape is one launch per sweep; the colour-packed shapes are two launches per sweep (one
    per colour), and their pack/unpack at the ends of the call are inside the timed region,
    amortised over the sweeps."""
    import torch

    from warpbiocell_amd import _lib
    from warpbiocell_amd.autotune import shape_name
    from warpbiocell_amd.fields.diffusion import solver_scalars

    kin = ox.species.kinetics
    dx2_over_D, _ = solver_scalars(ox.field, kin)
    stream = torch.cuda.current_stream(device)
    cfg = _lib.query("wbc_sor_fused_get_config")
    kind, tj, ti = (ctypes.c_int32() for _ in range(3))
    _lib.call("wbc_sor_fused_config_shape", cfg, ctypes.byref(kind), ctypes.byref(tj), ctypes.byref(ti))
    src = ox.field.values.clone()
    dst = torch.empty_like(src)
    ws = torch.empty(_lib.query("wbc_sor_sweeps_workspace_bytes", *geom.shape), dtype=torch.uint8, device=device)

    def sweeps(n):
        _lib.call("wbc_sor_sweeps_cfg", cfg, n, ox.settings.omega, dx2_over_D, kin.uptake_max, kin.michaelis_k, None,
                  ox.density.data_ptr(), None, 0, None, 1.0, src.data_ptr(), dst.data_ptr(), *geom.shape,
                  ws.data_ptr(), ws.numel(), stream.cuda_stream)

    sweeps(4)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    sweeps(timed_sweeps)
    end.record(stream)
    end.synchronize()
    per_sweep_s = start.elapsed_time(end) / 1e3 / timed_sweeps
    family = {0: "tile", 1: "march", 2: "pipe", 3: "copy", 4: "packed", 5: "quad"}[kind.value]
    shape = shape_name(kind.value, tj.value, ti.value)
    launches = 2 if family in ("packed", "quad") else 1
    box_lanes = quad_box_lanes(ox.density, geom.shape) if family == "quad" else None
    bytes_per_sweep = 16 * geom.node_count
    moved_per_sweep = bytes_per_sweep if box_lanes is None else 12 * geom.node_count + 2 * 16 * box_lanes
    achieved = bytes_per_sweep / per_sweep_s / 1e9
    achieved_moved = moved_per_sweep / per_sweep_s / 1e9
    sweep_traffic = measured_traffic(shape, workload, geom.shape)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": None if sweep_traffic is None else sweep_traffic // launches,
            "kernel": KERNEL_OF_KIND[family].format(tj=tj.value, ti=ti.value, q=max(ti.value, 1)), "shape": shape,
            "launches_per_sweep": launches,
            "timed_sweeps": timed_sweeps, "sweep_us": round(per_sweep_s * 1e6, 3),
            "launch_us": round(per_sweep_s * 1e6 / launches, 3),
            "algorithmic_bytes_per_launch": bytes_per_sweep // launches,
            "algorithmic_bytes_per_sweep": bytes_per_sweep,
            "density_box_lanes_per_pass": box_lanes,
            "bytes_moved_per_sweep": moved_per_sweep,
            "achieved_moved": round(achieved_moved, 1), "frac_moved": round(achieved_moved / HBM_PEAK_GBPS, 4),
            "traffic_per_sweep": sweep_traffic,
            "traffic_source": str(COUNTER_TABLE.relative_to(ROOT))}


def quiesce_gc() -> None:
    """Run Python's full garbage collection now and move every object that exists to the permanent
    generation, so that no full collection over the import- and setup-time objects (tens of ms with
    torch loaded) falls inside a timed region; which step it would hit depends on object counts
    that differ between a first run (bytecode compiled at import) and later ones."""
    import gc

    gc.collect()
    gc.freeze()


def adjoint_measurement(device) -> dict:
    """Pattern 3 on its own workload (100k cells, 129^3): one forward solve from a uniform start and
    one adjoint gradient (w.r.t. all per-cell rates, D and O_b) of sum (S - target)^2."""
    import torch

    from warpbiocell_amd import workloads as W
    from warpbiocell_amd.adjoint import OxygenAdjointProblem
    from warpbiocell_amd.fields.diffusion import SolverSettings
    from warpbiocell_amd.fields.scalar_field import GridGeometry

    w, rates, targets = W.adjoint_100k_129()
    geom = GridGeometry.centered_cube(w.box_um, w.dx_um)
    prob = OxygenAdjointProblem(geom, torch.from_numpy(w.positions), W.MICHAELIS_K, W.UPTAKE_MMHG_UM3_PER_H,
                                SolverSettings(omega=w.omega, tolerance=w.tolerance, max_sweeps=w.max_sweeps,
                                               check_every=w.check_every), device=device)
    q = torch.from_numpy(rates).to(device)
    T = torch.from_numpy(targets).to(device)
    S = prob.forward(q, W.D_UM2_PER_H, w.boundary_mmHg)
    prob.backward(2.0 * (S - T))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S = prob.forward(q, W.D_UM2_PER_H, w.boundary_mmHg)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gq, gD, gOb = prob.backward(2.0 * (S - T))
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    gradient = {"rates": gq.cpu().numpy(), "D": gD, "O_b": gOb}
    return {"workload": w.name, "forward_ms": round((t1 - t0) * 1e3, 3), "forward_sweeps": prob.last_forward.sweeps,
            "adjoint_ms": round((t2 - t1) * 1e3, 3), "adjoint_sweeps": prob.last_backward[0],
            "adjoint_over_forward": round((t2 - t1) / (t1 - t0), 3)}, gradient


def confinement_measurement(device, repeats: int = 25) -> dict:
    """The contact pass of the 100k-cell spheroid with and without a tissue region: a sphere just
    inside the spheroid's outer cells, on the oxygen grid, so the wall pushes the outer shell.
    Each call runs the step's mechanics (4 substeps of build -> gather (+ push) -> move)."""
    import numpy as np
    import torch

    from warpbiocell_amd import workloads as W
    from warpbiocell_amd.cells.mechanics import ContactParams, make_neighbor_grid, relax_contacts
    from warpbiocell_amd.cells.state import CellPopulation
    from warpbiocell_amd.fields.scalar_field import GridGeometry
    from warpbiocell_amd.geometry import Sphere, TissueRegion

    w = W.spheroid_100k_161()
    geom = GridGeometry.centered_cube(w.box_um, w.dx_um)
    r_cells = np.linalg.norm(w.positions, axis=1)
    region = TissueRegion.from_shape(Sphere(radius=float(np.percentile(r_cells, 95))), geom, device=device)
    params = ContactParams(stiffness=W.CONTACT_RATE, damping=1.0, max_radius=float(w.radii.max()))
    dt = W.DT_MECHANICS
    out = {"workload": w.name, "cells": int(w.n), "substeps_per_call": 4,
           "region": f"sphere r={float(np.percentile(r_cells, 95)):.1f} um"}
    for label, reg in (("plain", None), ("confined", region)):
        pop = CellPopulation.from_numpy(w.positions, w.radii, device=device)
        grid = make_neighbor_grid(params, device)
        relax_contacts(pop, grid, params, dt, 4, reg)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            relax_contacts(pop, grid, params, dt, 4, reg)
        torch.cuda.synchronize()
        out[f"{label}_us_per_substep"] = round((time.perf_counter() - t0) / (4 * repeats) * 1e6, 2)
        if reg is not None:
            inside = reg.contains(pop.positions_numpy(), margin=0.0)
            out["outside_after"] = int((~inside).sum())
    out["outside_before"] = int((~region.contains(w.positions)).sum())
    out["overhead"] = round(out["confined_us_per_substep"] / out["plain_us_per_substep"] - 1.0, 4)
    return out


def batched_measurement(device, steps: int) -> dict:
    """Batched sweep on the 15-replica workload: `steps` cell steps of every replica, once
    serially with the plain step (replica after replica, as the reference's sweep runs them, each
    with its autotuned SOR tile shape) and once batched (one hipGraph replay per step, the solves
    of all replicas in one persistent launch, autotuned schedule); the final states are compared
    bit for bit."""
    import numpy as np
    import torch

    from warpbiocell_amd import _lib
    from warpbiocell_amd import workloads as W
    from warpbiocell_amd.autotune import tune_batch, tune_sor
    from warpbiocell_amd.simulation import scenario
    from warpbiocell_amd.simulation.batched import CELL_FIELDS, BatchedSweep
    from warpbiocell_amd.simulation.simulator import cell_step

    ws = W.sweep_15()
    serial = [scenario.build(w, device) for w in ws]
    batched = [scenario.build(w, device) for w in ws]
    tuned = [tune_sor(sc.oxygen.field, sc.oxygen.density, sc.oxygen.species.kinetics, sc.settings.omega)
             for sc in serial]  # every distinct grid before the timed region (cached per shape)
    tuner = BatchedSweep([sc.replica() for sc in batched])
    tuner.step(1, use_graph=False)
    tune = tune_batch(tuner)
    tuner.close()
    del tuner
    batch = BatchedSweep([sc.replica() for sc in batched])
    tune_batch(batch)
    batch.capture()
    torch.cuda.synchronize()
    log(f"  batched sweep: {len(ws)} replicas, {sum(w.n for w in ws)} cells, "
        f"{sum(sc.geometry.node_count for sc in batched)} nodes; schedule {tune['schedule']}")

    quiesce_gc()
    t0 = time.perf_counter()
    serial_sweeps, per_replica = 0, []
    for sc, tune_r in zip(serial, tuned):
        _lib.call("wbc_sor_fused_set_config", tune_r["choice"])
        r0, sw = time.perf_counter(), 0
        for _ in range(steps):
            sw += cell_step(sc.population, sc.grid, sc.lifecycle, sc.contact, sc.stepping,
                            oxygen=sc.oxygen).field_sweeps
        torch.cuda.synchronize()
        per_replica.append({"grid": sc.geometry.shape[0], "cells": sc.workload.n, "sweeps": sw,
                            "sor_shape": tune_r["shape"],
                            "serial_s": round(time.perf_counter() - r0, 4)})
        serial_sweeps += sw
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    batch.step(steps, use_graph=True)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    batched_sweeps = sum(batch.total_sweeps())
    same = serial_sweeps == batched_sweeps
    for i, sc in enumerate(serial):
        n = sc.population.count
        for name in CELL_FIELDS:
            a = getattr(sc.population, name)[:n].cpu().numpy().view(np.uint8)
            same = same and np.array_equal(a, batch.population_view(i, name)[:n].cpu().numpy().view(np.uint8))
        same = same and np.array_equal(sc.oxygen.values.cpu().numpy().view(np.uint32),
                                       batch.field_view(i).cpu().numpy().view(np.uint32))
    cells = sum(w.n for w in ws)
    batch.close()
    # The colour passes' floor at the HBM peak over every replica's sweeps: 16 B per node and
    # sweep (two passes each reading both colours' field values once, one write, the density);
    # the quad passes load the density only inside its box, so the bytes they move are at least
    # 12 B per node and sweep (`sweep_floor_moved_s`).
    node_sweeps = sum(r["sweeps"] * r["grid"] ** 3 for r in per_replica)
    floor_s = 16 * node_sweeps / (HBM_PEAK_GBPS * 1e9)
    floor_moved_s = 12 * node_sweeps / (HBM_PEAK_GBPS * 1e9)
    return {"workload": "sweep_15", "replicas": len(ws), "cells": cells,
            "nodes": sum(sc.geometry.node_count for sc in batched), "steps": steps,
            "serial_s": round(t1 - t0, 4), "batched_s": round(t2 - t1, 4),
            "speedup": round((t1 - t0) / (t2 - t1), 3), "bit_exact": bool(same),
            "sweeps": batched_sweeps, "serial_cell_steps_per_s": round(cells * steps / (t1 - t0), 1),
            "batched_cell_steps_per_s": round(cells * steps / (t2 - t1), 1),
            "schedule": tune["schedule"], "schedule_ms_per_solve": tune["ms_per_solve"],
            "node_sweeps": node_sweeps, "sweep_floor_s": round(floor_s, 4),
            "speedup_bound_at_floor": round((t1 - t0) / floor_s, 2),
            "sweep_floor_moved_s": round(floor_moved_s, 4),
            "speedup_bound_at_moved_floor": round((t1 - t0) / floor_moved_s, 2), "per_replica": per_replica}


def snapshot(pop, ox):
    return {name: getattr(pop, name).cpu().numpy().copy() for name in
            ("position", "radius", "cell_state", "cell_type", "age", "oxygen_local", "glucose_local", "velocity",
             "neighbor_count", "divide_flag", "division_offset", "fate_flags")} | {
        "rng_state": pop.rng_numpy(), "count": pop.count, "field": ox.numpy()}


def host_threads() -> int:
    """Host cores this process may use: the affinity mask, capped by OMP_NUM_THREADS when set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() and int(env) > 0 else n


def device_copy(pop, ox):
    """A device copy of the population's arrays and count and of the field (restored by restore_copy)."""
    arrays = {name: getattr(pop, name).clone() for name in
              ("position", "radius", "cell_state", "cell_type", "age", "oxygen_local", "glucose_local", "velocity",
               "neighbor_count", "divide_flag", "division_offset", "fate_flags", "rng_state")}
    return arrays, pop.count, ox.values.clone()


def restore_copy(copy, pop, ox):
    arrays, count, field = copy
    for name, t in arrays.items():
        getattr(pop, name).copy_(t)
    pop.count = count
    pop.forget_deposit_order()
    ox.values.copy_(field)


def gpu_same_steps(copy, pop, ox, grid, lifecycle, contact, stepping, steps: int) -> dict:
    """The first `steps` timed steps again on the GPU, from the device copy of their start state."""
    import torch

    from warpbiocell_amd.simulation.simulator import cell_step

    restore_copy(copy, pop, ox)
    torch.cuda.synchronize()
    cells, sweeps = 0, []
    t0 = time.perf_counter()
    for _ in range(steps):
        cells += pop.count
        sweeps.append(cell_step(pop, grid, lifecycle, contact, stepping, oxygen=ox).field_sweeps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"value": round(cells / dt, 1), "seconds": round(dt, 4), "sweeps": sweeps}


def oracle_steps_seconds(snap, w, geom, params, settings, lifecycle, contact, stepping, threads: int, steps: int):
    """`steps` oracle steps from the snapshot on `threads` host threads: (seconds, cell-steps, sweeps per step)."""
    from oracle import oracle as O

    pop = O.OraclePopulation.from_numpy(snap["position"][: snap["count"]], snap["radius"][: snap["count"]],
                                        capacity=len(snap["radius"]), seed=w.seed)
    for name in ("position", "radius", "cell_state", "cell_type", "age", "oxygen_local", "glucose_local",
                 "velocity", "neighbor_count", "divide_flag", "division_offset", "fate_flags"):
        getattr(pop, name)[...] = snap[name]
    pop.rng_state[...] = snap["rng_state"]
    a = O.SolveArgs(dx=geom.dx, diffusion=params.diffusion_coefficient, uptake_max=params.uptake_max,
                    michaelis_k=params.michaelis_k, boundary_value=params.boundary_value)
    rox = O.OracleOxygen(shape=geom.shape, origin=geom.origin, args=a, omega=settings.omega,
                         max_sweeps=settings.max_sweeps, tolerance=settings.tolerance,
                         check_every=settings.check_every, values=snap["field"].copy())
    O.set_threads(threads)
    cells, sweeps = 0, []
    t0 = time.perf_counter()
    for _ in range(steps):
        cells += pop.count
        _, field = O.oracle_cell_step(pop, lifecycle, contact, stepping, oxygen=rox)
        sweeps.append(field[0])
    return time.perf_counter() - t0, cells, sweeps


def cpu_baseline(snap, w, geom, params, settings, lifecycle, contact, stepping, gpu_same: dict) -> dict:
    """The oracle (C restatement of the reference) on the host: the first CPU_STEPS timed steps from
    the snapshot they started at, on every host core the process may use, and the first two on one
    thread; beside it the GPU's time for the same steps (gpu_same)."""
    cores = host_threads()
    dt_all, cells, sweeps = oracle_steps_seconds(snap, w, geom, params, settings, lifecycle, contact, stepping,
                                                 cores, CPU_STEPS)
    dt_one, cells_one, sweeps_one = oracle_steps_seconds(snap, w, geom, params, settings, lifecycle, contact,
                                                         stepping, 1, 2)
    return {"value": round(cells / dt_all, 1), "unit": "cell-steps/s", "cores": cores, "kind": "port",
            "sample": f"the first {CPU_STEPS} timed steps of {w.name} ({cells} cell-steps, {sum(sweeps)} SOR sweeps) "
                      f"from the state they started at; oracle C restatement, -O2, OpenMP over {cores} host "
                      f"threads: {dt_all:.2f} s. The GPU ran the same steps (same sweeps) in "
                      f"{gpu_same['seconds']:.4f} s",
            "sweeps": sweeps, "seconds": round(dt_all, 3),
            "gpu_same_steps": gpu_same, "same_sweeps": sweeps == gpu_same["sweeps"],
            "single_thread": {"value": round(cells_one / dt_one, 1), "cores": 1, "seconds": round(dt_one, 3),
                              "steps": 2, "sweeps": sweeps_one}}


def field_accuracy(field32, density32, params, geom) -> dict:
    """The final headline field against float64: a float64 solve of the same discrete problem (the
    last step's density, the same boundary) started from the GPU's field and run to a residual of
    1e-13; also the float64 residual of the GPU's field."""
    import numpy as np

    from oracle import oracle as O

    a = O.SolveArgs(dx=geom.dx, diffusion=params.diffusion_coefficient, uptake_max=params.uptake_max,
                    michaelis_k=params.michaelis_k, boundary_value=params.boundary_value)
    f64 = field32.astype(np.float64)
    d64 = density32.astype(np.float64)
    res32 = O.residual(field32.astype(np.float64), d64, a)
    sweeps, res, conv = O.solve(f64, d64, a, O.optimal_omega(geom.shape), 20000, 1e-13, 10)
    err = np.abs(field32.astype(np.float64) - f64)
    return {"fp64_sweeps": sweeps, "fp64_residual": res, "fp64_converged": conv,
            "max_abs_err_over_boundary": float(err.max() / params.boundary_value),
            "rel_l2_err": float(np.linalg.norm(err) / np.linalg.norm(f64)),
            "fp64_residual_of_fp32_field": res32}


def mechanics_accuracy(pop, grid, contact) -> dict:
    """The contact gather at the final positions (one substep with dt = 0, which moves nothing)
    against a float64 restatement of kernels/mechanics_kernels.py:26-57 over every pair within the
    query radius (pairs found with scipy's cKDTree): velocities and neighbour counts."""
    import numpy as np
    import torch
    from scipy.spatial import cKDTree

    from warpbiocell_amd.cells.mechanics import contact_substep

    contact_substep(pop, grid, contact, 0.0)
    torch.cuda.synchronize()
    n = pop.count
    x = pop.position[:n].cpu().numpy().astype(np.float64)
    r = pop.radius[:n].cpu().numpy().astype(np.float64)
    v32 = pop.velocity[:n].cpu().numpy().astype(np.float64)
    c32 = pop.neighbor_count[:n].cpu().numpy()
    q, rate = float(contact.query_radius), float(contact.rate)
    pairs = cKDTree(x).query_pairs(q * (1.0 + 1e-5), output_type="ndarray")
    i, j = pairs[:, 0], pairs[:, 1]
    dv = x[i] - x[j]
    d = np.sqrt(np.sum(dv * dv, axis=1))
    near = d < q
    i, j, dv, d = i[near], j[near], dv[near], d[near]
    count = np.bincount(i, minlength=n) + np.bincount(j, minlength=n)
    push = (r[i] + r[j] - d > 0.0) & (d > 1e-6)
    coef = np.where(push, rate * (r[i] + r[j] - d) / np.where(d > 0.0, d, 1.0), 0.0)
    v64 = np.zeros((n, 3))
    np.add.at(v64, i, coef[:, None] * dv)
    np.add.at(v64, j, -coef[:, None] * dv)
    scale = float(np.abs(v64).max())
    return {"cells": n, "pairs_within_query_radius": int(near.sum()),
            "velocity_max_abs_err_over_max_speed": float(np.abs(v32 - v64).max() / scale) if scale > 0 else 0.0,
            "velocity_rel_l2_err": float(np.linalg.norm(v32 - v64) / np.linalg.norm(v64)) if scale > 0 else 0.0,
            "neighbor_count_mismatches": int(np.count_nonzero(c32 != count))}


def adjoint_accuracy(gradient: dict) -> dict:
    """The 129^3 adjoint workload's GPU gradient against the float64 adjoint (every rate, D, O_b) and
    central differences of float64 solves on 64 rates (oracle.adjoint_check)."""
    import numpy as np

    from oracle import oracle as O
    from warpbiocell_amd import workloads as W
    from warpbiocell_amd.fields.scalar_field import GridGeometry

    w, rates, targets = W.adjoint_100k_129()
    geom = GridGeometry.centered_cube(w.box_um, w.dx_um)
    O.set_threads(host_threads())
    r = O.adjoint_check(w.positions, rates, targets, W.D_UM2_PER_H, w.boundary_mmHg, W.MICHAELIS_K, geom.origin,
                        geom.dx, geom.shape, gradient["rates"], gradient["D"], gradient["O_b"], fd_cells=64)
    per_rate = np.asarray(r["gradient_vs_fd_rel_err"])
    r["gradient_vs_fd_rel_err_median"] = float(np.median(per_rate))
    r["gradient_vs_fd_rel_err_max"] = float(per_rate.max())
    r["fd_vs_fp64_rel_err_max"] = float(np.max(r["fd_vs_fp64_rel_err"]))
    r["fd_vs_fp64_rel_err"] = [float(f"{e:.3e}") for e in r["fd_vs_fp64_rel_err"]]
    r["gradient_vs_fd_rel_err"] = [float(f"{e:.3e}") for e in per_rate]
    return r


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="spheroid_100k_161")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-confinement", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--sor-shape", default=None,
                    help="run the SOR sweeps with this shape (autotune.config_shapes() name) instead of the "
                         "autotuner's choice; the autotuner still times every shape")
    ap.add_argument("--sweep-grid", type=int, default=None,
                    help="only time the SOR sweep (autotuned shape, 1000 sweeps) of the workload's cells on "
                         "an N^3 grid over its box, after one warmup step, and print the roofline as one JSON line")
    ap.add_argument("--batched-only", action="store_true",
                    help="only run the batched-sweep leg and print its result as one JSON line")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy 5 times (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "bench.py runs the hot path on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        return 2

    import torch

    from warpbiocell_amd import _lib
    from warpbiocell_amd.simulation.simulator import cell_step

    _lib.load()
    device = torch.device("cuda", 0)
    if args.copy_calibration:
        src = torch.ones(64 * 2 ** 20, dtype=torch.float32, device=device)
        dst = torch.empty_like(src)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(5):
            dst.copy_(src)
        end.record()
        torch.cuda.synchronize()
        moved = 5 * 2 * src.numel() * 4  # each copy reads and writes 256 MiB, past the 256 MB Infinity Cache
        log("copy calibration: 5 x 256 MiB device-to-device copies")
        print(json.dumps({"copy_bytes_moved": moved, "seconds": start.elapsed_time(end) / 1e3,
                          "GB_per_s": round(moved / (start.elapsed_time(end) / 1e3) / 1e9, 1)}), flush=True)
        return 0
    if args.batched_only:
        print(json.dumps({"batched": batched_measurement(device, args.batched_steps)}), flush=True)
        return 0
    if args.sweep_grid:
        from warpbiocell_amd.autotune import tune_sor

        w, geom, params, settings, lifecycle, contact, stepping, pop, ox, grid = setup(args.workload, device,
                                                                                        args.sweep_grid)
        cell_step(pop, grid, lifecycle, contact, stepping, oxygen=ox)
        tune = tune_sor(ox.field, ox.density, ox.species.kinetics, settings.omega)
        roof = sor_roofline(ox, geom, device, w.name)
        log(f"sweep only: {w.name} cells on {geom.shape}: {roof['shape']} {roof['sweep_us']} us per sweep, "
            f"frac {roof['frac']}")
        print(json.dumps({"workload": w.name, "grid": list(geom.shape), "roofline": roof, "sor_autotune": tune}),
              flush=True)
        return 0
    w, geom, params, settings, lifecycle, contact, stepping, pop, ox, grid = setup(args.workload, device)
    log(f"bench: {w.name}: {pop.count} cells, grid {geom.shape}, warmup {args.warmup}, steps {args.steps}")
    for s in range(args.warmup):
        r = cell_step(pop, grid, lifecycle, contact, stepping, oxygen=ox)
        log(f"  warmup {s}: sweeps {r.field_sweeps} residual {r.field_residual:.3e}")
    from warpbiocell_amd.autotune import tune_sor

    tune = tune_sor(ox.field, ox.density, ox.species.kinetics, settings.omega)
    log(f"  autotune: tile {tune['shape']} of {tune['us_per_sweep']} (us per sweep)")
    if args.sor_shape:
        from warpbiocell_amd.autotune import config_shapes

        _lib.call("wbc_sor_fused_set_config", config_shapes().index(args.sor_shape))
        tune = dict(tune, forced=args.sor_shape)
        log(f"  sor shape forced: {args.sor_shape}")
    headline_config = _lib.query("wbc_sor_fused_get_config")
    torch.cuda.synchronize()
    snap = None if args.no_cpu_baseline else snapshot(pop, ox)
    start_copy = None if args.no_cpu_baseline else device_copy(pop, ox)

    reports, cell_steps = [], 0
    quiesce_gc()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cell_steps += pop.count
        reports.append(cell_step(pop, grid, lifecycle, contact, stepping, oxygen=ox))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    sweeps = [r.field_sweeps for r in reports]
    log(f"  timed: {elapsed:.3f} s, sweeps/step mean {sum(sweeps) / len(sweeps):.1f} "
        f"(min {min(sweeps)}, max {max(sweeps)}), unconverged steps {sum(not r.field_converged for r in reports)}")

    final_field = ox.numpy() if snap is not None else None
    final_density = ox.density.cpu().numpy() if snap is not None else None
    mech_accuracy = mechanics_accuracy(pop, grid, contact) if snap is not None else None
    roof = None if args.no_roofline else sor_roofline(ox, geom, device, w.name)
    if roof:
        log(f"  sor sweep ({roof['shape']}): {roof['sweep_us']} us, {roof['achieved']} GB/s algorithmic")
    adj, gradient = adjoint_measurement(device) if not args.no_adjoint else (None, None)
    if adj:
        log(f"  adjoint: {adj}")
    conf = None if args.no_confinement else confinement_measurement(device)
    if conf:
        log(f"  confinement: {conf}")
    bat = None if args.no_batched else batched_measurement(device, args.batched_steps)
    if bat:
        log(f"  batched: {bat}")
    cpu, accuracy = None, None
    if snap is not None:
        # The batched leg selects each replica's tuned shape in turn; the headline's goes back.
        _lib.call("wbc_sor_fused_set_config", headline_config)
        same = gpu_same_steps(start_copy, pop, ox, grid, lifecycle, contact, stepping, CPU_STEPS)
        cpu = cpu_baseline(snap, w, geom, params, settings, lifecycle, contact, stepping, same)
        log(f"  cpu baseline: {cpu['value']} cell-steps/s over {CPU_STEPS} steps ({cpu['seconds']} s, "
            f"sweeps {sum(cpu['sweeps'])}); GPU same steps {same['value']} cell-steps/s ({same['seconds']} s)")
        accuracy = {"field": field_accuracy(final_field, final_density, params, geom), "mechanics": mech_accuracy}
        log(f"  accuracy, headline field vs float64: {accuracy['field']}")
        log(f"  accuracy, contact gather at the final positions vs float64: {mech_accuracy}")
        if gradient is not None:
            accuracy["adjoint"] = adjoint_accuracy(gradient)
            log("  accuracy, adjoint gradient: " + str({k: v for k, v in accuracy["adjoint"].items()
                                                       if not isinstance(v, list)}))
    result = {
        "metric": METRIC,
        "value": round(cell_steps / elapsed, 1),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "cells": w.n, "grid": list(geom.shape), "omega": settings.omega,
                   "tolerance": settings.tolerance, "check_every": settings.check_every,
                   "mechanics_substeps": stepping.mechanics_substeps,
                   "sweeps_per_step_mean": round(sum(sweeps) / len(sweeps), 1),
                   "unconverged_steps": sum(not r.field_converged for r in reports),
                   "final_residual": reports[-1].field_residual},
        "roofline": roof,
        "sor_autotune": tune,
        "adjoint": adj,
        "batched": bat,
        "confinement": conf,
        "cpu_baseline": cpu,
        "accuracy": accuracy,
    }
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
