"""Autotuners.

``tune_sor``: the fused SOR sweep. Times every tile shape of the library's table on the caller's
grid (HIP events on the launch stream) and selects the fastest for the solve loop. Cached per
(device, grid shape).

``tune_batch``: the persistent solve of a batched replica sweep. Times every (tile shape,
workgroups per CU) schedule on the batch's own replicas and selects the fastest. Cached per
(device, replica grids and cell counts).
"""

from __future__ import annotations

import ctypes

import torch

from warpbiocell_amd import _lib

_cache: dict = {}


def shape_name(kind: int, tj: int, ti: int) -> str:
    return {0: f"tile60x{tj}x{ti}", 1: f"march60x{tj}/{ti}", 2: f"pipe60x{tj}x{ti}", 3: f"copy60x{tj}x{ti}",
            4: f"packed/{tj}", 5: "packed/quad" if ti <= 1 else f"packed/quad{ti}"}[kind]


def config_shapes() -> list[str]:
    """Names of the shape table: 'tile60xTJxTI' (3-D tile), 'march60xTJ/CI' (marching i) or
    'pipe60xTJxTI' (3-D tile, pipelined persistent workgroups), 'copy60xTJxTI' (3-D tile,
    global->LDS copies), 'packed/RJ' (colour passes on colour-packed arrays, RJ rows per
    thread) or 'packed/quad', 'packed/quadQ' (colour passes on colour-packed arrays, four packed
    indices of one row per lane, Q rounds of 64 lanes per wave)."""
    out = []
    for c in range(_lib.query("wbc_sor_fused_config_count")):
        kind, tj, ti = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _lib.call("wbc_sor_fused_config_shape", c, ctypes.byref(kind), ctypes.byref(tj), ctypes.byref(ti))
        out.append(shape_name(kind.value, tj.value, ti.value))
    return out


def time_config(cfg: int, field, density: torch.Tensor, kinetics, omega: float, sweeps: int = 50) -> float:
    """Microseconds per sweep with shape ``cfg`` on ``field``'s grid (values untouched): ``sweeps``
    sweeps in one ``wbc_sor_sweeps_cfg`` call (for the colour-packed entry that includes packing
    once and unpacking once), after one warm-up call."""
    from warpbiocell_amd.fields.diffusion import solver_scalars

    dev = field.device
    stream = torch.cuda.current_stream(dev)
    dx2_over_D, _ = solver_scalars(field, kinetics)
    nx, ny, nz = field.geometry.shape
    src, dst = field.values.clone(), torch.empty_like(field.values)
    nbytes = _lib.query("wbc_sor_sweeps_workspace_bytes", nx, ny, nz)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)

    def run(n):
        _lib.call("wbc_sor_sweeps_cfg", cfg, n, omega, dx2_over_D, kinetics.uptake_max, kinetics.michaelis_k,
                  _lib.ptr(field.fixed_or_none), density.data_ptr(), None, 0, None, 1.0, src.data_ptr(),
                  dst.data_ptr(), nx, ny, nz, ws.data_ptr(), nbytes, stream.cuda_stream)

    run(4)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    run(sweeps)
    b.record(stream)
    b.synchronize()
    return a.elapsed_time(b) * 1e3 / sweeps


def supported_configs(field) -> list[int]:
    """Shape-table entries that can run on ``field`` (the colour-packed one needs Dirichlet faces
    and no pinned nodes; production terms are the caller's concern)."""
    out = []
    for c, name in enumerate(config_shapes()):
        if name.startswith("packed") and (field.fixed_or_none is not None or any(field.boundary_flags)):
            continue
        out.append(c)
    return out


def tune_sor(field, density: torch.Tensor, kinetics, omega: float, sweeps: int = 50) -> dict:
    """Time every tile shape, select the fastest; returns {shape: us} and the choice."""
    key = (str(field.device), tuple(field.geometry.shape), field.fixed_or_none is not None,
           tuple(field.boundary_flags))
    if key not in _cache:
        shapes = config_shapes()
        cfgs = supported_configs(field)
        times = {c: time_config(c, field, density, kinetics, omega, sweeps) for c in cfgs}
        best = min(times, key=times.get)
        _cache[key] = {"choice": best, "shape": shapes[best],
                       "us_per_sweep": {shapes[c]: round(t, 3) for c, t in times.items()}}
    _lib.call("wbc_sor_fused_set_config", _cache[key]["choice"])
    return _cache[key]


def batch_schedules() -> list[tuple[int, int, str]]:
    """(schedule, workgroups per CU, name) of every schedule of the batched solve: the persistent
    launch with each tile shape at 1, 2, ... up to the occupancy limit workgroups per CU
    ('copy' / 'pipe' 60 x TJ x TI @ n/CU), and the wide launches (one launch per colour pass over
    every replica: 'wide/RJ' with RJ rows per thread, 'wide/quad' with four packed indices per
    lane)."""
    count = _lib.query("wbc_batch_schedule_count")
    shapes = []
    for tile in range(count):
        tj, ti, most = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _lib.call("wbc_batch_schedule_shape", tile, ctypes.byref(tj), ctypes.byref(ti), ctypes.byref(most))
        shapes.append((tile, tj.value, ti.value, most.value))
    persistent = sum(1 for s in shapes if s[2] != 0)
    out = []
    for tile, tj, ti, most in shapes:
        if ti == 0:
            out.append((tile, 0, "wide/quad" if tj == 1 else f"wide/{tj}"))
            continue
        kind = "pipe" if tile >= persistent // 2 else "copy"
        for bpc in range(1, most + 1):
            out.append((tile, bpc, f"{kind}60x{tj}x{ti}@{bpc}/CU"))
    return out


def tune_batch(batch, repeats: int = 1) -> dict:
    """Time the batched solve of every replica of ``batch`` under every schedule, from the same
    start fields (restored after each run; the solve's results do not depend on the schedule),
    and select the fastest. Cached per (device, replica grids and cell counts)."""
    key = ("batch", str(batch.device), tuple((r.oxygen.geometry.shape, r.population.count) for r in batch.replicas))
    if key not in _cache:
        stream = torch.cuda.current_stream(batch.device)
        start = batch.field.clone()
        times = {}
        for tile, bpc, name in batch_schedules():
            batch.set_schedule(tile, bpc)
            best = float("inf")
            for _ in range(repeats + 1):  # the first run warms up
                batch.field.copy_(start)
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                batch.solve()
                b.record(stream)
                b.synchronize()
                best = min(best, a.elapsed_time(b))
            batch.report()  # raises if a barrier timed out
            times[(tile, bpc, name)] = best
        batch.field.copy_(start)
        choice = min(times, key=times.get)
        _cache[key] = {"tile": choice[0], "blocks_per_cu": choice[1], "schedule": choice[2],
                       "ms_per_solve": {n: round(t, 3) for (_, _, n), t in times.items()}}
    c = _cache[key]
    batch.set_schedule(c["tile"], c["blocks_per_cu"])
    return c
