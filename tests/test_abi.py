"""The C-ABI library loads without a GPU and exports every entry point include/wbc.h declares;
the host-side size queries answer on the CPU."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "wbc.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|size_t)\s+(wbc_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_hot_path():
    names = declared_symbols()
    for required in ("wbc_deposit_trilinear_i64", "wbc_sor_sweep", "wbc_solve_steady_state", "wbc_sample_trilinear",
                     "wbc_hashgrid_build", "wbc_contact_velocities", "wbc_scan_inclusive_i32",
                     "wbc_place_daughters", "wbc_lifecycle_decide", "wbc_init_rng"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from warpbiocell_amd import _lib

    lib = _lib.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_covers_the_header():
    from warpbiocell_amd import _lib

    assert set(_lib.SIGNATURES) == set(declared_symbols())


def test_workspace_queries_run_on_the_host():
    from warpbiocell_amd import _lib

    assert _lib.query("wbc_scan_workspace_bytes", 4096) == 256
    assert _lib.query("wbc_scan_workspace_bytes", 10 ** 6) >= (10 ** 6 // 4096 + 1) * 4
    hg = _lib.query("wbc_hashgrid_workspace_bytes", 128, 100_000)
    assert hg >= 2 * 4 * 128 ** 3 + 28 * 100_000
    assert _lib.query("wbc_hashgrid_workspace_bytes", 2, 10) == 0  # rejected
    assert _lib.query("wbc_solve_workspace_bytes", 2000, 10, 41, 41, 41) >= 4 * (204 + 41 ** 3)


def test_library_links_the_hip_runtime_by_soname():
    import subprocess

    out = subprocess.run(["readelf", "-d", str(ROOT / "warpbiocell_amd" / "libwbc_hip.so")], capture_output=True,
                         text=True).stdout
    assert "libamdhip64.so.7" in out


def test_no_gpu_means_a_loud_error():
    import torch

    from warpbiocell_amd import _lib

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        _lib.require_gpu("cuda:0")
    with pytest.raises(RuntimeError):
        _lib.require_gpu("cpu")


def test_noalloc_scan_rejects_bad_sizes_before_any_device_work():
    """The no-allocation scan checks n against its fixed tile-sum array (2^28 values) and its
    pointers before touching the device, so these calls return on a machine without a GPU."""
    from warpbiocell_amd import _lib

    fn = _lib.load().wbc_scan_inclusive_i32_noalloc
    invalid = 1  # hipErrorInvalidValue
    assert fn(None, None, 2 ** 28 + 1, None) == invalid
    assert fn(None, None, -1, None) == invalid
    assert fn(None, None, 1000, None) == invalid  # null pointers
    assert fn(None, None, 0, None) == 0  # nothing to do
