"""Build the gfx950 HIP library in-tree with hipcc (no JIT cache, no torch
extension machinery): the shared object lands next to this file so it
travels with the source tree.  Each source compiles to its own object in
parallel; the objects are then linked into one shared library."""

from __future__ import annotations

import os
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
LIB_NAME = "libptyrodactyl_amd.so"
LIB_PATH = PKG_DIR / LIB_NAME
SOURCES = ("forward.hip", "adjoint.hip", "adjoint_stream.hip", "aux_kernels.hip", "optim.hip", "potential.hip", "xl2.hip", "grid.hip", "jvp.hip", "capi.hip")
ARCH = os.environ.get("PTY_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # compiler-generated packed FP32 (v_pk_*_f32) is disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc_path() -> str:
    cand = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(cand).exists():
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return cand


def _run(cmd: list) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stderr[-8000:]}")


def build_library(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every HIP source into ``libptyrodactyl_amd.so`` (always rebuilds)."""
    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    hipcc = hipcc_path()
    jobs = jobs or min(len(srcs), max(1, min(os.cpu_count() or 1, 16)))
    with tempfile.TemporaryDirectory(prefix="pty_build_") as tmpdir:
        objs = [Path(tmpdir) / (s.stem + ".o") for s in srcs]
        cmds = [[hipcc, *HIPCC_FLAGS, "-c", str(s), "-o", str(o)] for s, o in zip(srcs, objs)]
        if verbose:
            for c in cmds:
                print(" ".join(c), flush=True)
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            list(pool.map(_run, cmds))
        tmp = LIB_PATH.with_suffix(".so.tmp")
        link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
        if verbose:
            print(" ".join(link), flush=True)
        _run(link)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
