"""Compile the HIP kernels for gfx950 into ``warpbiocell_amd/libwbc_hip.so`` (in-tree).

The library is a plain C-ABI shared object loaded with ctypes, not a torch extension, so it
does not depend on torch's C++ ABI. It links the HIP runtime by soname (``libamdhip64.so.7``),
which resolves to the runtime torch has already loaded into the process.

Flags that matter for results:
  -ffp-contract=off                         no fused a*b+c: one rounding per operation, the
                                            order the reference writes its expressions in
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE FP32 division and square root
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                            no packed FP32 instructions
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
CSRC = PACKAGE_DIR / "csrc"
INCLUDE = PACKAGE_DIR.parent / "include"
LIBRARY = PACKAGE_DIR / "libwbc_hip.so"
SOURCES = ("wbc_scan.hip", "wbc_cells.hip", "wbc_field.hip", "wbc_sor_fused.hip", "wbc_sor_march.hip",
           "wbc_sor_packed.hip", "wbc_mechanics.hip", "wbc_batched.hip", "wbc_adjoint.hip", "wbc_metrics.hip")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _torch_lib_dir() -> str | None:
    try:
        import torch  # noqa: F401

        return str(Path(torch.__file__).resolve().parent / "lib")
    except Exception:
        return None


def build_library(verbose: bool = False) -> Path:
    """Compile every source (always, even when a library exists) and link the shared object."""
    hipcc = _hipcc()
    sources = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    with tempfile.TemporaryDirectory(prefix="wbc_build_") as tmp:
        objects = []

        def compile_one(src: Path) -> Path:
            obj = Path(tmp) / (src.stem + ".o")
            cmd = [hipcc, *HIP_FLAGS, "-I", str(CSRC), "-I", str(INCLUDE), "-c", str(src), "-o", str(obj)]
            res = subprocess.run(cmd, capture_output=True, text=True)
            out = "\n".join(line for line in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in line)
            if res.returncode != 0:
                raise RuntimeError(f"hipcc failed on {src.name}:\n{out}")
            if verbose and out.strip():
                print(out, file=sys.stderr)
            return obj

        with ThreadPoolExecutor(max_workers=min(8, len(sources))) as pool:
            objects = list(pool.map(compile_one, sources))
        tmp_lib = Path(tmp) / LIBRARY.name
        link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp_lib), *map(str, objects)]
        torch_lib = _torch_lib_dir()
        if torch_lib:
            link.append(f"-Wl,-rpath,{torch_lib}")
        res = subprocess.run(link, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"link failed:\n{res.stdout}{res.stderr}")
        staged = LIBRARY.with_name(LIBRARY.name + ".tmp")
        shutil.copy2(tmp_lib, staged)
        os.replace(staged, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build_library(verbose=True))
