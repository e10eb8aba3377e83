"""The scan entry point that allocates nothing (wbc_scan_inclusive_i32_noalloc) against the scan with
a caller workspace (wbc_scan_inclusive_i32), through the C ABI: equal bit for bit on two calls in a
row, the second smaller, so the library's fixed tile-sum array is reused with a larger call's sums
still in it."""

import numpy as np
import pytest
import torch

from warpbiocell_amd import _lib

pytestmark = pytest.mark.gpu


def _scan_with_workspace(x: torch.Tensor) -> torch.Tensor:
    n = x.numel()
    ws = torch.empty(_lib.query("wbc_scan_workspace_bytes", n), dtype=torch.uint8, device=x.device)
    out = torch.empty_like(x)
    _lib.call("wbc_scan_inclusive_i32", x.data_ptr(), out.data_ptr(), n, ws.data_ptr(), ws.numel(),
              _lib.stream_of(x.device))
    return out


def _scan_noalloc(x: torch.Tensor) -> torch.Tensor:
    out = torch.empty_like(x)
    _lib.call("wbc_scan_inclusive_i32_noalloc", x.data_ptr(), out.data_ptr(), x.numel(), _lib.stream_of(x.device))
    return out


def test_noalloc_scan_equals_the_workspace_scan_on_two_calls(device):
    rng = np.random.default_rng(7)
    # 6,000,000 values: 1,465 tiles, so the tile sums are scanned by their own launch; then
    # 100,003 values: 25 tiles, each adding up the earlier tiles' sums itself.
    for n in (6_000_000, 100_003):
        host = rng.integers(0, 3, n, dtype=np.int32)
        x = torch.from_numpy(host).to(device)
        a = _scan_noalloc(x)
        b = _scan_with_workspace(x)
        torch.cuda.synchronize()
        assert torch.equal(a, b), n
        np.testing.assert_array_equal(a.cpu().numpy(), np.cumsum(host, dtype=np.int64).astype(np.int32))


def test_noalloc_scan_of_the_division_flags_matches_commit_divisions_offsets(device):
    """The scan commit_divisions runs (divide_flag -> division_offset, cells/lifecycle.py:146), on the
    flags of a population where about 5 % of cells divide."""
    rng = np.random.default_rng(3)
    flags = (rng.random(250_000) < 0.05).astype(np.int32)
    x = torch.from_numpy(flags).to(device)
    a = _scan_noalloc(x)
    torch.cuda.synchronize()
    assert int(a[-1]) == int(flags.sum())
    assert torch.equal(a, _scan_with_workspace(x))
