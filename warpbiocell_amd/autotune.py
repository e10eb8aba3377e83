"""Autotuner for the fused SOR sweep: times every tile shape of the library's table on the
caller's grid (HIP events on the launch stream) and selects the fastest for the solve loop.

Results are cached per (device, grid shape); ``tune_sor`` is cheap to call again.
"""

from __future__ import annotations

import ctypes

import torch

from warpbiocell_amd import _lib

_cache: dict = {}


def config_shapes() -> list[str]:
    """Names of the shape table: 'tile60xTJxTI' (3-D tile) or 'march60xTJ/CI' (marching i)."""
    out = []
    for c in range(_lib.query("wbc_sor_fused_config_count")):
        kind, tj, ti = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _lib.call("wbc_sor_fused_config_shape", c, ctypes.byref(kind), ctypes.byref(tj), ctypes.byref(ti))
        out.append(f"tile60x{tj.value}x{ti.value}" if kind.value == 0 else f"march60x{tj.value}/{ti.value}")
    return out


def time_config(cfg: int, field, density: torch.Tensor, kinetics, omega: float, sweeps: int = 50) -> float:
    """Microseconds per fused sweep with tile shape ``cfg`` on ``field``'s grid (values untouched)."""
    from warpbiocell_amd.fields.diffusion import solver_scalars

    dev = field.device
    stream = torch.cuda.current_stream(dev)
    dx2_over_D, _ = solver_scalars(field, kinetics)
    bufs = [field.values.clone(), torch.empty_like(field.values)]

    def sweep(c):
        _lib.call("wbc_sor_sweep_fused_cfg", cfg, omega, dx2_over_D, kinetics.uptake_max, kinetics.michaelis_k,
                  _lib.ptr(field.fixed_or_none), density.data_ptr(), None, 0, None, 1.0, bufs[c & 1].data_ptr(),
                  bufs[(c + 1) & 1].data_ptr(), *field.geometry.shape, stream.cuda_stream)

    for c in range(4):
        sweep(c)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for c in range(sweeps):
        sweep(c)
    b.record(stream)
    b.synchronize()
    return a.elapsed_time(b) * 1e3 / sweeps


def tune_sor(field, density: torch.Tensor, kinetics, omega: float, sweeps: int = 50) -> dict:
    """Time every tile shape, select the fastest; returns {shape: us} and the choice."""
    key = (str(field.device), tuple(field.geometry.shape))
    if key not in _cache:
        shapes = config_shapes()
        times = [time_config(c, field, density, kinetics, omega, sweeps) for c in range(len(shapes))]
        best = min(range(len(shapes)), key=lambda c: times[c])
        _cache[key] = {"choice": best, "shape": shapes[best],
                       "us_per_sweep": {s: round(t, 3) for s, t in zip(shapes, times)}}
    _lib.call("wbc_sor_fused_set_config", _cache[key]["choice"])
    return _cache[key]
