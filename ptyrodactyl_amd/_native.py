"""ctypes binding of the C-ABI in ``include/ptyrodactyl_amd.h``.

The HIP library is the only compute path for the multislice kernels: if it
is missing this module raises on first use (there is no CPU fallback).
torch is imported first so that the library binds to the HIP runtime torch
already loaded (both carry the soname ``libamdhip64.so.7``).
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import torch  # noqa: F401  (must precede loading the HIP library)

from .build import LIB_PATH

PTY_OK = 0
PTY_ERR_INVALID = 1
PTY_ERR_DEVICE = 2
ABI_VERSION = 8

_lock = threading.Lock()
_lib = None

c_int = ctypes.c_int
c_ll = ctypes.c_longlong
c_size = ctypes.c_size_t
c_vp = ctypes.c_void_p
c_dbl = ctypes.c_double

# name -> (restype, argtypes); mirrors include/ptyrodactyl_amd.h
SIGNATURES = {
    "pty_abi_version": (c_int, []),
    "pty_last_error": (ctypes.c_char_p, []),
    "pty_forward_workspace_bytes": (c_size, [c_int, c_ll, c_int, c_ll]),
    "pty_forward": (c_int, [c_int, c_int, c_ll, c_int, c_vp, c_ll, c_ll, c_vp, c_vp, c_vp, c_vp,
                            c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "pty_forward_noalloc": (c_int, [c_int, c_int, c_ll, c_int, c_vp, c_ll, c_ll, c_vp, c_vp, c_vp, c_vp,
                                    c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pty_jvp_workspace_bytes": (c_size, [c_int, c_ll, c_int, c_ll]),
    "pty_jvp": (c_int, [c_int, c_int, c_ll, c_int, c_vp, c_ll, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                        c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "pty_forward_checkpointed": (c_int, [c_int, c_int, c_ll, c_int, c_vp, c_ll, c_ll, c_vp, c_vp, c_vp, c_vp,
                                         c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_size, c_vp]),
    "pty_adjoint_workspace_bytes": (c_size, [c_int, c_int, c_int, c_ll, c_int]),
    "pty_adjoint": (c_int, [c_int, c_int, c_ll, c_int, c_vp, c_ll, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                            c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "pty_adjoint_amplitude": (c_int, [c_int, c_int, c_ll, c_int, c_vp, c_ll, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp,
                                      c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
    "pty_transmission": (c_int, [c_vp, c_dbl, c_vp, c_ll, c_vp]),
    "pty_sqrt_intensity_loss_workspace_bytes": (c_size, []),
    "pty_sqrt_intensity_loss": (c_int, [c_ll, c_vp, c_vp, ctypes.c_float, c_vp, c_vp, c_vp, c_size, c_vp]),
    "pty_normal_equations_workspace_bytes": (c_size, [c_ll, c_int, c_int]),
    "pty_normal_equations": (c_int, [c_vp, c_vp, c_ll, c_int, c_int, c_vp, c_size, c_vp, c_vp, c_vp, c_vp]),
    "pty_damped_solve": (c_int, [c_vp, c_vp, c_dbl, c_int, c_vp, c_vp, c_vp, c_vp]),
    "pty_forward_batched": (c_int, [c_int, c_int, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_size,
                                    c_vp]),
    "pty_adjoint_batched": (c_int, [c_int, c_int, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp,
                                    c_size, c_vp]),
    "pty_annular_integrate": (c_int, [c_vp, c_ll, c_ll, c_vp, c_vp, c_vp]),
    "pty_forward_detectors": (c_int, [c_int, c_int, c_ll, c_int, c_vp, c_ll, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                      c_vp, c_int, c_vp, c_vp, c_vp, c_size, c_vp]),
    "pty_random_normal": (c_int, [ctypes.c_uint, ctypes.c_uint, c_ll, c_vp, c_vp]),
    "pty_projected_potential": (c_int, [c_int, c_vp, c_ll, c_vp, c_vp, c_vp]),
    "pty_atom_potential_grid": (c_int, [c_int, c_int, c_vp, c_dbl, c_int, c_int, c_int, c_vp, c_vp, c_vp]),
    "pty_potential_slices_fourier": (c_int, [c_int, c_int, c_int, c_int, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "pty_atom_center_grad_workspace_bytes": (c_size, [c_int]),
    "pty_potential_shift_grad_workspace_bytes": (c_size, [c_int, c_int, c_ll]),
    "pty_potential_slices_shift_grad": (c_int, [c_int, c_int, c_int, c_int, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                                c_size, c_vp]),
    "pty_atom_potential_grid_center_grad": (c_int, [c_int, c_int, c_vp, c_dbl, c_int, c_int, c_int, c_vp, c_vp, c_vp,
                                                    c_vp, c_size, c_vp]),
    "pty_optimizer_update": (c_int, [c_int, c_int, c_int, c_ll, c_vp, c_vp, c_vp, c_vp, c_ll, c_dbl, c_vp]),
    "pty_persistent_state_bytes": (c_size, []),
    "pty_persistent_workspace_bytes": (c_size, [c_int, c_int]),
    "pty_forward_persistent": (c_int, [c_int, c_int, c_ll, c_vp, c_ll, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                       c_vp, c_int, c_int, c_vp, c_size, c_vp, c_vp]),
    "pty_grid_workspace_bytes": (c_size, [c_int, c_int, c_int, c_int, c_ll, c_int]),
    "pty_forward_grid": (c_int, [c_int, c_int, c_int, c_ll, c_int, c_vp, c_ll, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp,
                                 c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_size, c_vp]),
    "pty_adjoint_grid": (c_int, [c_int, c_int, c_int, c_ll, c_int, c_vp, c_ll, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp,
                                 c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_size, c_vp]),
}

# numpy view of pty_position_desc (include/ptyrodactyl_amd.h); 96 bytes
POSITION_DESC_FIELDS = [("slab_offset", "<i8"), ("slab_h", "<i8"), ("slab_w", "<i8"), ("pattern_offset", "<i8"),
                        ("probe_offset", "<i8"), ("prop_offset", "<i8"), ("grad_offset", "<i8"),
                        ("gprobe_offset", "<i8"), ("origin_y", "<i4"), ("origin_x", "<i4"), ("shift_y", "<f8"),
                        ("shift_x", "<f8")]


class NativeLibraryMissing(RuntimeError):
    """The HIP library has not been built (run ``__graft_entry__.build()``)."""


def load(path: Path | None = None):
    """Load (once) and return the ctypes handle of the HIP library."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = Path(path) if path is not None else LIB_PATH
        if not p.exists():
            raise NativeLibraryMissing(
                f"{p} not found: build it with `python -m ptyrodactyl_amd.build`")
        lib = ctypes.CDLL(str(p), mode=ctypes.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            if not hasattr(lib, name):
                continue
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.pty_abi_version() != ABI_VERSION:
            raise NativeLibraryMissing("HIP library ABI version mismatch; rebuild it")
        _lib = lib
        return lib


def check(status: int, what: str) -> None:
    """Map a C-ABI status to the reference's error conventions."""
    if status == PTY_OK:
        return
    msg = load().pty_last_error().decode(errors="replace")
    if status == PTY_ERR_INVALID:
        raise ValueError(f"{what}: {msg}")
    raise RuntimeError(f"{what}: {msg}")


def ptr(t) -> int:
    """Device address of a tensor (0 for None)."""
    return 0 if t is None else int(t.data_ptr())


def stream_handle(device=None) -> int:
    """hipStream_t of torch's current stream on ``device``."""
    return int(torch.cuda.current_stream(device).cuda_stream)
