"""JAX-compatible counter-based random numbers for the helpers that take a
``jax.random`` key (``decompose_beam_to_modes``).

Keys are raw ``uint32[2]`` pairs as ``jax.random.PRNGKey(seed)`` makes them
(seed >> 32, seed & 0xffffffff); ``split`` follows jax's partitionable
Threefry layout (key i = threefry(key, (0, i))); ``normal`` runs on the GPU
(``pty_random_normal``).  Threefry-2x32 is the published Random123 generator
(Salmon et al., SC'11) that JAX uses.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native

_ROTATIONS = ((13, 15, 26, 6), (17, 29, 16, 24))


def PRNGKey(seed: int) -> np.ndarray:  # noqa: N802  (the reference's name)
    """Raw key of ``jax.random.PRNGKey(seed)``."""
    s = int(seed)
    return np.array([(s >> 32) & 0xFFFFFFFF, s & 0xFFFFFFFF], dtype=np.uint32)


def _as_key(key) -> np.ndarray:
    if isinstance(key, (int, np.integer)):
        return PRNGKey(int(key))
    k = key.detach().cpu().numpy() if isinstance(key, torch.Tensor) else np.asarray(key)
    k = k.astype(np.uint64).reshape(-1)
    if k.shape != (2,):
        raise ValueError("a key is two uint32 words (jax.random.PRNGKey layout) or an integer seed")
    return k.astype(np.uint32)


def _threefry(key: np.ndarray, x0: np.ndarray, x1: np.ndarray):
    k0, k1 = np.uint32(key[0]), np.uint32(key[1])
    ks = (k0, k1, np.uint32(k0 ^ k1 ^ np.uint32(0x1BD11BDA)))
    with np.errstate(over="ignore"):
        x0 = (x0.astype(np.uint32) + ks[0]).astype(np.uint32)
        x1 = (x1.astype(np.uint32) + ks[1]).astype(np.uint32)
        for block in range(5):
            for r in _ROTATIONS[block % 2]:
                x0 = (x0 + x1).astype(np.uint32)
                x1 = ((x1 << np.uint32(r)) | (x1 >> np.uint32(32 - r))).astype(np.uint32)
                x1 = (x1 ^ x0).astype(np.uint32)
            x0 = (x0 + ks[(block + 1) % 3]).astype(np.uint32)
            x1 = (x1 + ks[(block + 2) % 3] + np.uint32(block + 1)).astype(np.uint32)
    return x0, x1


def split(key, num: int = 2):
    """``jax.random.split(key, num)`` (partitionable layout): a list of raw keys."""
    k = _as_key(key)
    b0, b1 = _threefry(k, np.zeros(num, np.uint32), np.arange(num, dtype=np.uint32))
    return [np.array([b0[i], b1[i]], dtype=np.uint32) for i in range(num)]


def normal(key, shape, device=None) -> torch.Tensor:
    """``jax.random.normal(key, shape, float64)`` generated on the GPU."""
    k = _as_key(key)
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    if dev.type != "cuda":
        raise ValueError("random.normal runs on the GPU (no CPU fallback)")
    shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
    out = torch.empty(shape, dtype=torch.float64, device=dev)
    lib = _native.load()
    with torch.cuda.device(dev):
        status = lib.pty_random_normal(int(k[0]), int(k[1]), out.numel(), _native.ptr(out),
                                       _native.stream_handle(dev))
    _native.check(status, "pty_random_normal")
    return out
