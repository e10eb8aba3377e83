"""ctypes binding of ``libwbc_hip.so`` (the C ABI declared in ``include/wbc.h``).

The library must be built in-tree (``python -m warpbiocell_amd._build`` or
``__graft_entry__.build()``). There is no fallback: every hot-path operation raises if the
library or a ROCm GPU is missing.
"""

from __future__ import annotations

import ctypes
from ctypes import c_double, c_float, c_int32, c_int64, c_size_t, c_void_p
from pathlib import Path

import torch

LIBRARY_PATH = Path(__file__).resolve().parent / "libwbc_hip.so"

P = c_void_p  # device pointer / stream / workspace
I32 = c_int32
I64 = c_int64
F32 = c_float

# name -> (restype, argtypes); the same order as include/wbc.h.
SIGNATURES: dict[str, tuple] = {
    "wbc_init_rng": (I32, [P, I64, I32, P]),
    "wbc_scan_workspace_bytes": (c_size_t, [I64]),
    "wbc_scan_inclusive_i32": (I32, [P, P, I64, P, c_size_t, P]),
    "wbc_scan_exclusive_i32": (I32, [P, P, I64, P, c_size_t, P]),
    "wbc_scan_inclusive_i32_noalloc": (I32, [P, P, I64, P]),
    "wbc_lifecycle_decide": (I32, [F32, F32, F32, F32, I32, F32, F32, F32, F32, I32, I32, F32, F32,
                                   P, P, P, P, P, P, P, P, I64, P]),
    "wbc_place_daughters": (I32, [I32, F32, P, P, P, P, P, P, P, P, P, P, P, P, P, I64, P]),
    "wbc_commit_divisions": (I32, [I32, I64, F32, P, P, P, c_size_t, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "wbc_fill_zero_i64": (I32, [P, I64, P]),
    "wbc_deposit_trilinear_i64": (I32, [P, P, I64, F32, F32, F32, F32, I32, I32, I32, P, P]),
    "wbc_deposit_trilinear_i64_ordered": (I32, [P, P, P, I64, F32, F32, F32, F32, I32, I32, I32, P, P]),
    "wbc_deposit_accumulate_i64": (I32, [P, P, P, I64, F32, F32, F32, F32, I32, I32, I32, P, P, P]),
    "wbc_density_finalize": (I32, [P, I64, F32, F32, F32, F32, F32, P, P, P, P]),
    "wbc_combine_densities": (I32, [P, I64, F32, F32, F32, F32, P, P]),
    "wbc_sample_trilinear": (I32, [P, I32, I32, I32, P, I64, F32, F32, F32, F32, P, P]),
    "wbc_sor_sweep": (I32, [I32, F32, F32, F32, F32, I32, I32, I32, P, P, P, I32, P, F32, P, I32, I32, I32, P]),
    "wbc_steady_state_residual": (I32, [F32, F32, F32, F32, I32, I32, I32, P, P, P, I32, P, F32, P, I32, I32, I32,
                                        P, P]),
    "wbc_sor_sweep_fused": (I32, [F32, F32, F32, F32, P, P, P, I32, P, F32, P, P, I32, I32, I32, P]),
    "wbc_sor_fused_config_count": (I32, []),
    "wbc_sor_fused_config_shape": (I32, [I32, P, P, P]),
    "wbc_sor_fused_set_config": (I32, [I32]),
    "wbc_solve_set_lookahead": (I32, [I32, F32]),
    "wbc_sor_fused_get_config": (I32, []),
    "wbc_sor_sweep_fused_cfg": (I32, [I32, F32, F32, F32, F32, P, P, P, I32, P, F32, P, P, I32, I32, I32, P]),
    "wbc_sor_sweeps_workspace_bytes": (c_size_t, [I32, I32, I32]),
    "wbc_sor_sweeps_cfg": (I32, [I32, I32, F32, F32, F32, F32, P, P, P, I32, P, F32, P, P, I32, I32, I32, P, c_size_t,
                                 P]),
    "wbc_solve_workspace_bytes": (c_size_t, [I32, I32, I32, I32, I32]),
    "wbc_solve_steady_state": (I32, [F32, I32, F32, I32, F32, F32, F32, F32, I32, I32, I32, P, P, P, I32, P, F32,
                                     P, I32, I32, I32, P, c_size_t, P, P, P, P, P]),
    "wbc_hashgrid_workspace_bytes": (c_size_t, [I32, I64]),
    "wbc_hashgrid_init": (I32, [P, c_size_t, I32, I64, P]),
    "wbc_hashgrid_build": (I32, [P, P, I32, F32, I32, I64, P, c_size_t, P]),
    "wbc_hashgrid_point_ids": (I32, [P, c_size_t, I32, I64, P, P]),
    "wbc_contact_velocities": (I32, [P, c_size_t, I32, I64, I32, F32, F32, F32, P, P, P]),
    "wbc_integrate_positions": (I32, [P, P, I64, F32, P]),
    "wbc_relax_contacts": (I32, [P, P, I32, F32, I32, F32, F32, F32, I32, P, P, I64, P, c_size_t, P]),
    "wbc_confinement_velocities": (I32, [P, F32, P, P, I64, P, P]),
    "wbc_relax_contacts_confined": (I32, [P, P, I32, F32, I32, F32, F32, F32, I32, P, F32, P, P, I64, P, c_size_t,
                                          P]),
    "wbc_sample_sdf": (I32, [P, P, I64, P, P]),
    "wbc_region_gradient": (I32, [P, I32, I32, I32, c_double, P, P]),
    "wbc_deposit_rates_i64": (I32, [P, P, I64, F32, F32, F32, F32, F32, I32, I32, I32, P, P]),
    "wbc_fixed_to_float": (I32, [P, I64, F32, P, P]),
    "wbc_sample_adjoint": (I32, [P, P, I64, F32, F32, F32, F32, I32, I32, I32, P, P]),
    "wbc_adjoint_workspace_bytes": (c_size_t, [I32, I32, I32]),
    "wbc_adjoint_set_packed": (I32, [I32]),
    "wbc_adjoint_solve": (I32, [F32, I32, F32, I32, F32, F32, F32, P, P, P, P, P, P, P, I64, F32, F32, F32, F32,
                                I32, I32, I32, P, P, P, P, P, c_size_t, P]),
    "wbc_metrics_workspace_bytes": (c_size_t, [I32, I32]),
    "wbc_cell_moments": (I32, [P, P, P, P, I32, P, c_size_t, P, P]),
    "wbc_centre_distances": (I32, [P, I32, P, P, P, P, c_size_t, P]),
    "wbc_select_rank": (I32, [P, I32, I32, P, P, c_size_t, P]),
    "wbc_grid_moments": (I32, [P, P, I32, I32, I32, I32, I32, I32, P, c_size_t, P, P]),
    "wbc_radial_profile": (I32, [P, P, P, P, P, I32, P, I32, P, P, P, c_size_t, P]),
    "wbc_batch_struct_size": (c_size_t, [I32]),
    "wbc_batch_workspace_bytes": (c_size_t, [P, I32, I32]),
    "wbc_batch_create": (I32, [P, I32, I32, P, P, P, c_size_t, P, P]),
    "wbc_batch_schedule_count": (I32, []),
    "wbc_batch_schedule_shape": (I32, [I32, P, P, P]),
    "wbc_batch_set_schedule": (I32, [P, I32, I32]),
    "wbc_batch_step": (I32, [P, I32, I32, P]),
    "wbc_batch_solve": (I32, [P, P]),
    "wbc_batch_report": (I32, [P, P, P, P, P, P, P]),
    "wbc_batch_cell_counts": (I32, [P, P, P, P, P]),
    "wbc_batch_destroy": (I32, [P]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the library once (after torch, so it binds to the HIP runtime torch loaded)."""
    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            raise RuntimeError(f"{LIBRARY_PATH.name} is not built: run `python -m warpbiocell_amd._build`")
        lib = ctypes.CDLL(str(LIBRARY_PATH))
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = lib
    return _lib


def require_gpu(device) -> torch.device:
    """The HIP path runs only on a ROCm GPU; raise (never fall back) otherwise."""
    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    if dev.type != "cuda":
        raise RuntimeError(f"warpbiocell_amd runs on a ROCm GPU (cuda:N); got device {dev}")
    if not torch.cuda.is_available():
        raise RuntimeError("no ROCm GPU is visible: the HIP hot path cannot run")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


def stream_of(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def call(name: str, *args) -> None:
    """Call an entry point; a nonzero hipError_t becomes RuntimeError."""
    rc = getattr(load(), name)(*args)
    if rc != 0:
        raise RuntimeError(f"{name} failed with hipError_t {rc}")


def query(name: str, *args) -> int:
    return int(getattr(load(), name)(*args))


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("device arrays passed to the HIP library must be contiguous")
    return t.data_ptr()
