"""Build the gfx950 HIP library in-tree with hipcc (no JIT cache, no torch
extension machinery): the shared object lands next to this file so it
travels with the source tree."""

from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
LIB_NAME = "libptyrodactyl_amd.so"
LIB_PATH = PKG_DIR / LIB_NAME
SOURCES = ("forward.hip", "adjoint.hip", "aux_kernels.hip", "optim.hip", "potential.hip", "xl2.hip", "grid.hip", "jvp.hip", "capi.hip")
ARCH = os.environ.get("PTY_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # compiler-generated packed FP32 (v_pk_*_f32) is disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc_path() -> str:
    cand = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(cand).exists():
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return cand


def build_library(verbose: bool = False) -> Path:
    """Compile every HIP source into ``libptyrodactyl_amd.so`` (always rebuilds)."""
    srcs = [str(CSRC / s) for s in SOURCES if (CSRC / s).exists()]
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc_path(), *HIPCC_FLAGS, *srcs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{proc.stderr[-8000:]}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
