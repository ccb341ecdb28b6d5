"""Device-side multislice engine: validates operands and calls the C-ABI.

This is the single entry to the HIP forward kernels.  Operands are torch
tensors that already live on the GPU; the engine checks dtype, layout,
shape and device on the host before any launch, allocates outputs and the
streaming workspace with torch's caching allocator, and launches on torch's
current stream.
"""

from __future__ import annotations

import os
from typing import Optional, Tuple

import torch

from . import _native

SUPPORTED_N = (32, 64, 128, 256, 512, 1024, 2048, 4096)
# windows of the checkpointed forward and the adjoint
GRADIENT_N = SUPPORTED_N
# windows of the fused virtual-detector epilogue
DETECTOR_N = (32, 64, 128, 256, 512, 1024)
# from 512 up the default streaming batch is sized by bytes: 192 MiB of waves measured fastest at 512^2
# (96 waves: 64.4 ms vs 71.8 ms with 512) and 1024^2 (24 waves: 17.2 ms vs 19.6 ms with 32)
LARGE_BATCH_BYTES = 192 << 20
# Waves (positions x modes) per streaming batch for n >= 256; the batch's
# waves live in a device workspace that should stay cache-resident.
DEFAULT_BATCH_WAVES = int(os.environ.get("PTY_MAX_BATCH_WAVES", "512"))


def _require(cond: bool, msg: str) -> None:
    if not cond:
        raise ValueError(msg)


def _check_cuda(t: torch.Tensor, name: str, dtype, ndim: int, device) -> None:
    _require(isinstance(t, torch.Tensor), f"{name} must be a torch tensor")
    _require(t.is_cuda, f"{name} must be a GPU tensor (the multislice path has no CPU fallback)")
    _require(t.dtype == dtype, f"{name} must be {dtype}, got {t.dtype}")
    _require(t.ndim == ndim, f"{name} must be {ndim}-D, got shape {tuple(t.shape)}")
    _require(t.is_contiguous(), f"{name} must be contiguous")
    _require(t.device == device, f"{name} is on {t.device}, expected {device}")


def workspace_bytes(n: int, num_positions: int, num_modes: int, max_batch_waves: Optional[int] = None) -> int:
    lib = _native.load()
    if max_batch_waves is not None:
        cap = int(max_batch_waves)
    elif n <= 256:
        cap = DEFAULT_BATCH_WAVES
    else:
        cap = max(int(num_modes), LARGE_BATCH_BYTES // (int(n) * int(n) * 8))
    return int(lib.pty_forward_workspace_bytes(int(n), int(num_positions), int(num_modes), cap))


def multislice_forward(slab_t: torch.Tensor, probe_k: torch.Tensor, weights: torch.Tensor,
                       origins: torch.Tensor, shifts_px: torch.Tensor, prop: torch.Tensor, *,
                       want_patterns: bool = True, want_amplitude: bool = False,
                       patterns_out: Optional[torch.Tensor] = None,
                       max_batch_waves: Optional[int] = None) -> Tuple[Optional[torch.Tensor], Optional[torch.Tensor]]:
    """Windowed multislice forward on the GPU.

    slab_t (S, Hs, Ws) complex64 transmission; probe_k (M, n, n) complex64
    fft2 of the probe modes; weights (M,) float32; origins (P, 2) int32;
    shifts_px (P, 2) float64; prop (n,) complex64 separable propagator
    factor including 1/n.  Returns (patterns (P, n, n) float32 fftshifted,
    amplitude (P, M, n, n) complex64 fftshifted) with None for outputs not
    requested.
    """
    _require(isinstance(slab_t, torch.Tensor) and slab_t.is_cuda,
             "slab must be a GPU tensor (the multislice path has no CPU fallback)")
    dev = slab_t.device
    _check_cuda(slab_t, "slab", torch.complex64, 3, dev)
    _check_cuda(probe_k, "probe_k", torch.complex64, 3, dev)
    m_count, n, n2 = probe_k.shape
    _require(n == n2, "probe window must be square")
    _require(n in SUPPORTED_N, f"window size {n} is not one of {SUPPORTED_N}")
    s_count, hs, ws = slab_t.shape
    _require(s_count >= 1, "at least one slice is required")
    _require(hs >= n and ws >= n, f"slab {hs}x{ws} is smaller than the {n}x{n} window")
    _check_cuda(weights, "weights", torch.float32, 1, dev)
    _require(weights.shape[0] == m_count, "weights must have shape (M,)")
    _check_cuda(origins, "origins", torch.int32, 2, dev)
    _check_cuda(shifts_px, "shifts_px", torch.float64, 2, dev)
    p_count = origins.shape[0]
    _require(origins.shape == (p_count, 2) and shifts_px.shape == (p_count, 2),
             "origins and shifts must both have shape (P, 2)")
    _require(p_count >= 1, "at least one scan position is required")
    _check_cuda(prop, "prop", torch.complex64, 1, dev)
    _require(prop.shape[0] == n, "propagator factor must have n entries")
    _require(want_patterns or want_amplitude, "nothing requested")

    patterns = None
    if want_patterns:
        if patterns_out is not None:
            _check_cuda(patterns_out, "patterns_out", torch.float32, 3, dev)
            _require(tuple(patterns_out.shape) == (p_count, n, n), "patterns_out has the wrong shape")
            patterns = patterns_out
        else:
            patterns = torch.empty((p_count, n, n), dtype=torch.float32, device=dev)
    amplitude = torch.empty((p_count, m_count, n, n), dtype=torch.complex64, device=dev) if want_amplitude else None
    ws_bytes = workspace_bytes(n, p_count, m_count, max_batch_waves)
    workspace = torch.empty((max(ws_bytes, 1),), dtype=torch.uint8, device=dev) if ws_bytes else None

    lib = _native.load()
    with torch.cuda.device(dev):
        status = lib.pty_forward(
            n, s_count, p_count, m_count, _native.ptr(slab_t), hs, ws, _native.ptr(probe_k),
            _native.ptr(weights), _native.ptr(origins), _native.ptr(shifts_px), _native.ptr(prop),
            _native.ptr(prop), _native.ptr(patterns), _native.ptr(amplitude), _native.ptr(workspace),
            ws_bytes, _native.stream_handle(dev))
    _native.check(status, "pty_forward")
    return patterns, amplitude


class CapturedForward:
    """A forward over fixed operands captured once in a HIP graph and replayed.

    The launch sequence of :func:`multislice_forward` (one launch per position
    group for n <= 128, 2S + 2 launches per batch for n >= 256) is recorded on
    a capture stream after one eager warm-up call, together with its
    workspace; :meth:`replay` re-runs it on the current stream with whatever
    the operand tensors hold at that time (update them in place between
    replays, e.g. the object in a reconstruction loop).  The captured
    sequence contains kernels only (no memset or copy nodes).
    """

    def __init__(self, slab_t, probe_k, weights, origins, shifts_px, prop, max_batch_waves: Optional[int] = None):
        _check_forward_operands(slab_t, probe_k, weights, origins, shifts_px, prop)
        self.operands = (slab_t, probe_k, weights, origins, shifts_px, prop)
        p_count, n = origins.shape[0], probe_k.shape[-1]
        self.patterns = torch.empty((p_count, n, n), dtype=torch.float32, device=slab_t.device)
        self.max_batch_waves = max_batch_waves
        multislice_forward(*self.operands, patterns_out=self.patterns, max_batch_waves=max_batch_waves)  # warm-up
        torch.cuda.synchronize(slab_t.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            multislice_forward(*self.operands, patterns_out=self.patterns, max_batch_waves=max_batch_waves)

    def replay(self) -> torch.Tensor:
        self.graph.replay()
        return self.patterns


MAX_DETECTORS = 4


def multislice_forward_detectors(slab_t: torch.Tensor, probe_k: torch.Tensor, weights: torch.Tensor,
                                 origins: torch.Tensor, shifts_px: torch.Tensor, prop: torch.Tensor,
                                 masks: torch.Tensor, *, want_patterns: bool = False,
                                 max_batch_waves: Optional[int] = None
                                 ) -> Tuple[Optional[torch.Tensor], torch.Tensor]:
    """Forward with virtual detectors fused into the |psi|^2 epilogue.

    masks: (D, n, n) uint8 0/1 in the fftshifted pattern layout, D <= 4.  Returns (patterns or None,
    signals (P, D) float64) where signals[p, d] is the sum of pattern p over mask d.  Without
    ``want_patterns`` the pattern stack is never written.
    """
    dev, n, s_count, hs, ws, m_count, p_count = _check_forward_operands(slab_t, probe_k, weights, origins,
                                                                        shifts_px, prop, DETECTOR_N,
                                                                        "virtual detectors")
    _check_cuda(masks, "masks", torch.uint8, 3, dev)
    d_count = masks.shape[0]
    _require(1 <= d_count <= MAX_DETECTORS and tuple(masks.shape[1:]) == (n, n),
             f"masks must have shape (D, {n}, {n}) with 1 <= D <= {MAX_DETECTORS}")
    patterns = torch.empty((p_count, n, n), dtype=torch.float32, device=dev) if want_patterns else None
    signals = torch.empty((p_count, d_count), dtype=torch.float64, device=dev)
    ws_bytes = workspace_bytes(n, p_count, m_count, max_batch_waves)
    workspace = torch.empty((max(ws_bytes, 1),), dtype=torch.uint8, device=dev) if ws_bytes else None
    lib = _native.load()
    with torch.cuda.device(dev):
        status = lib.pty_forward_detectors(
            n, s_count, p_count, m_count, _native.ptr(slab_t), hs, ws, _native.ptr(probe_k), _native.ptr(weights),
            _native.ptr(origins), _native.ptr(shifts_px), _native.ptr(prop), _native.ptr(prop),
            _native.ptr(patterns), d_count, _native.ptr(masks), _native.ptr(signals), _native.ptr(workspace),
            ws_bytes, _native.stream_handle(dev))
    _native.check(status, "pty_forward_detectors")
    return patterns, signals


PERSISTENT_GROUP_WAVES = 2
PERSISTENT_BLOCKS_PER_CU = 3
_STATE_ERROR_INDEX = 0


def multislice_forward_persistent(slab_t: torch.Tensor, probe_k: torch.Tensor, weights: torch.Tensor,
                                  origins: torch.Tensor, shifts_px: torch.Tensor, prop: torch.Tensor, *,
                                  patterns_out: Optional[torch.Tensor] = None,
                                  group_waves: int = PERSISTENT_GROUP_WAVES, grid: Optional[int] = None,
                                  state: Optional[torch.Tensor] = None, workspace: Optional[torch.Tensor] = None,
                                  check: bool = True) -> torch.Tensor:
    """Single-launch forward (n = 256, one probe mode) that keeps wave groups in one XCD's L2.

    Same inputs and bitwise the same patterns as :func:`multislice_forward`.
    ``state`` and ``workspace`` may be passed in to reuse them across calls;
    with ``check`` the call waits for the launch and raises if a wait timed
    out or an XCD left work undone.
    """
    dev, n, s_count, hs, ws, m_count, p_count = _check_forward_operands(slab_t, probe_k, weights, origins,
                                                                        shifts_px, prop)
    _require(n == 256 and m_count == 1, "the persistent forward supports n = 256 with one probe mode")
    g = int(group_waves)
    lib = _native.load()
    if patterns_out is None:
        patterns_out = torch.empty((p_count, n, n), dtype=torch.float32, device=dev)
    _check_cuda(patterns_out, "patterns_out", torch.float32, 3, dev)
    _require(tuple(patterns_out.shape) == (p_count, n, n), "patterns_out has the wrong shape")
    state_bytes = int(lib.pty_persistent_state_bytes())
    if state is None:
        state = torch.empty((state_bytes // 4,), dtype=torch.int32, device=dev)
    _require(state.dtype == torch.int32 and state.numel() * 4 >= state_bytes, "state buffer too small")
    ws_bytes = int(lib.pty_persistent_workspace_bytes(n, g))
    _require(ws_bytes > 0, "group_waves must be >= 1")
    if workspace is None:
        workspace = torch.empty((ws_bytes,), dtype=torch.uint8, device=dev)
    _require(workspace.numel() * workspace.element_size() >= ws_bytes, "workspace too small")
    if grid is None:
        grid = torch.cuda.get_device_properties(dev).multi_processor_count * PERSISTENT_BLOCKS_PER_CU
    with torch.cuda.device(dev):
        status = lib.pty_forward_persistent(
            n, s_count, p_count, _native.ptr(slab_t), hs, ws, _native.ptr(probe_k), _native.ptr(weights),
            _native.ptr(origins), _native.ptr(shifts_px), _native.ptr(prop), _native.ptr(prop),
            _native.ptr(patterns_out), g, int(grid), _native.ptr(workspace), ws_bytes, _native.ptr(state),
            _native.stream_handle(dev))
    _native.check(status, "pty_forward_persistent")
    if check:
        flag = int(state[_STATE_ERROR_INDEX].item())
        if flag & 1:
            raise RuntimeError("pty_forward_persistent: a wait timed out; the patterns are incomplete")
        if flag & 2:
            raise RuntimeError("pty_forward_persistent: an XCD left tasks undone (unexpected XCD numbering)")
    return patterns_out


def _check_forward_operands(slab_t, probe_k, weights, origins, shifts_px, prop, allowed=SUPPORTED_N,
                            what: str = "the forward"):
    _require(isinstance(slab_t, torch.Tensor) and slab_t.is_cuda,
             "slab must be a GPU tensor (the multislice path has no CPU fallback)")
    dev = slab_t.device
    _check_cuda(slab_t, "slab", torch.complex64, 3, dev)
    _check_cuda(probe_k, "probe_k", torch.complex64, 3, dev)
    m_count, n, n2 = probe_k.shape
    _require(n == n2, "probe window must be square")
    _require(n in allowed, f"window size {n} is not one of {allowed} ({what})")
    s_count, hs, ws = slab_t.shape
    _require(hs >= n and ws >= n, f"slab {hs}x{ws} is smaller than the {n}x{n} window")
    _check_cuda(weights, "weights", torch.float32, 1, dev)
    _require(weights.shape[0] == m_count, "weights must have shape (M,)")
    _check_cuda(origins, "origins", torch.int32, 2, dev)
    _check_cuda(shifts_px, "shifts_px", torch.float64, 2, dev)
    p_count = origins.shape[0]
    _require(p_count >= 1 and origins.shape == (p_count, 2) and shifts_px.shape == (p_count, 2),
             "origins and shifts must both have shape (P, 2) with P >= 1")
    _check_cuda(prop, "prop", torch.complex64, 1, dev)
    _require(prop.shape[0] == n, "propagator factor must have n entries")
    return dev, n, s_count, hs, ws, m_count, p_count


def num_checkpoints(num_slices: int, ckpt_every: int) -> int:
    return (num_slices - 1) // ckpt_every


def multislice_forward_checkpointed(slab_t, probe_k, weights, origins, shifts_px, prop, ckpt_every: int,
                                    max_batch_waves: Optional[int] = None):
    """Forward that also returns the wave entering slices k, 2k, ... for the adjoint.

    Returns (patterns (P, n, n) float32, checkpoints (P*M, (S-1)//k, n, n) complex64).
    """
    dev, n, s_count, hs, ws, m_count, p_count = _check_forward_operands(slab_t, probe_k, weights, origins,
                                                                        shifts_px, prop, GRADIENT_N,
                                                                        "checkpoints")
    _require(int(ckpt_every) >= 1, "ckpt_every must be >= 1")
    nc = num_checkpoints(s_count, int(ckpt_every))
    patterns = torch.empty((p_count, n, n), dtype=torch.float32, device=dev)
    ckpt = torch.empty((p_count * m_count, nc, n, n), dtype=torch.complex64, device=dev)
    ws_bytes = workspace_bytes(n, p_count, m_count, max_batch_waves)
    workspace = torch.empty((max(ws_bytes, 1),), dtype=torch.uint8, device=dev) if ws_bytes else None
    lib = _native.load()
    with torch.cuda.device(dev):
        status = lib.pty_forward_checkpointed(
            n, s_count, p_count, m_count, _native.ptr(slab_t), hs, ws, _native.ptr(probe_k), _native.ptr(weights),
            _native.ptr(origins), _native.ptr(shifts_px), _native.ptr(prop), _native.ptr(prop),
            _native.ptr(patterns), _native.ptr(ckpt) if nc else 0, int(ckpt_every), _native.ptr(workspace),
            ws_bytes, _native.stream_handle(dev))
    _native.check(status, "pty_forward_checkpointed")
    return patterns, ckpt


def multislice_adjoint(slab_t, probe_k, weights, origins, shifts_px, prop, grad_patterns, ckpt, ckpt_every: int,
                       grad_slab: Optional[torch.Tensor] = None, grad_probe_k: Optional[torch.Tensor] = None,
                       want_probe: bool = True, grad_amplitude: Optional[torch.Tensor] = None,
                       grad_shifts: Optional[torch.Tensor] = None):
    """Reverse-mode multislice: accumulate dL/dT into grad_slab and dL/dprobe_k into grad_probe_k.

    The seed is dL/dI (``grad_patterns``, (P, n, n) float32), or, when ``grad_amplitude`` is given,
    dL/dA of the fftshifted per-mode amplitudes ((P, M, n, n) complex64; ``grad_patterns`` is then
    ignored).  Gradients are dL/dRe + i dL/dIm (torch's convention for a real loss).
    ``grad_shifts`` ((P, 2) float64, may be None) accumulates dL/d(shifts_px).
    Returns (grad_slab, grad_probe_k or None).
    """
    dev, n, s_count, hs, ws, m_count, p_count = _check_forward_operands(slab_t, probe_k, weights, origins,
                                                                        shifts_px, prop, GRADIENT_N,
                                                                        "the adjoint")
    k = int(ckpt_every)
    _require(k >= 1, "ckpt_every must be >= 1")
    nc = num_checkpoints(s_count, k)
    if grad_amplitude is not None:
        _check_cuda(grad_amplitude, "grad_amplitude", torch.complex64, 4, dev)
        _require(tuple(grad_amplitude.shape) == (p_count, m_count, n, n), "grad_amplitude must have shape (P, M, n, n)")
    else:
        _check_cuda(grad_patterns, "grad_patterns", torch.float32, 3, dev)
        _require(tuple(grad_patterns.shape) == (p_count, n, n), "grad_patterns must have shape (P, n, n)")
    if nc:
        _check_cuda(ckpt, "checkpoints", torch.complex64, 4, dev)
        _require(tuple(ckpt.shape) == (p_count * m_count, nc, n, n), "checkpoint buffer has the wrong shape")
    if grad_slab is None:
        grad_slab = torch.zeros_like(slab_t)
    _check_cuda(grad_slab, "grad_slab", torch.complex64, 3, dev)
    _require(grad_slab.shape == slab_t.shape, "grad_slab must match the slab")
    if want_probe and grad_probe_k is None:
        grad_probe_k = torch.zeros_like(probe_k)
    if grad_probe_k is not None:
        _check_cuda(grad_probe_k, "grad_probe_k", torch.complex64, 3, dev)
        _require(grad_probe_k.shape == probe_k.shape, "grad_probe_k must match probe_k")
    if grad_shifts is not None:
        _check_cuda(grad_shifts, "grad_shifts", torch.float64, 2, dev)
        _require(tuple(grad_shifts.shape) == (p_count, 2), "grad_shifts must have shape (P, 2)")
    lib = _native.load()
    ws_bytes = int(lib.pty_adjoint_workspace_bytes(n, s_count, k, p_count, m_count))
    workspace = torch.empty((max(ws_bytes, 1),), dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        fn = lib.pty_adjoint_amplitude if grad_amplitude is not None else lib.pty_adjoint
        seed = grad_amplitude if grad_amplitude is not None else grad_patterns
        status = fn(
            n, s_count, p_count, m_count, _native.ptr(slab_t), hs, ws, _native.ptr(probe_k), _native.ptr(weights),
            _native.ptr(origins), _native.ptr(shifts_px), _native.ptr(prop), _native.ptr(prop),
            _native.ptr(seed), _native.ptr(ckpt) if nc else 0, k, _native.ptr(grad_slab),
            _native.ptr(grad_probe_k), _native.ptr(grad_shifts), _native.ptr(workspace), ws_bytes,
            _native.stream_handle(dev))
    _native.check(status, "pty_adjoint_amplitude" if grad_amplitude is not None else "pty_adjoint")
    return grad_slab, grad_probe_k


# ---------------------------------------------------------------------------
# Batched launches over positions of several requests (continuous batching)
# ---------------------------------------------------------------------------

_BASES = {}


def arena_base(device) -> torch.Tensor:
    """A small per-device allocation used as the common base of every arena.

    Descriptor offsets are (tensor address - base address) / element size, so
    one launch (or one captured graph) can address tensors that were
    allocated independently by different requests.
    """
    key = str(device)
    if key not in _BASES:
        _BASES[key] = torch.zeros(64, dtype=torch.float64, device=device)
    return _BASES[key]


def element_offset(t: torch.Tensor, base: torch.Tensor) -> int:
    diff = int(t.data_ptr()) - int(base.data_ptr())
    size = t.element_size()
    _require(diff % size == 0, "tensor is not aligned to its element size relative to the arena base")
    return diff // size


def forward_batched(n: int, num_slices: int, descs: torch.Tensor, base: torch.Tensor,
                    ckpt: Optional[torch.Tensor] = None, ckpt_every: int = 0,
                    max_batch_waves: Optional[int] = None, amplitudes: Optional[torch.Tensor] = None) -> None:
    """Launch the forward over ``descs`` (device uint8 view of pty_position_desc records).

    ``amplitudes`` (optional, (P, n, n) complex64, call-local position order) receives the fftshifted exit-plane
    amplitudes."""
    dev = base.device
    p_count = descs.numel() // 88
    _require(p_count >= 1 and descs.numel() == p_count * 88, "descriptor buffer must hold whole records")
    ws_bytes = workspace_bytes(n, p_count, 1, max_batch_waves)
    workspace = torch.empty((max(ws_bytes, 1),), dtype=torch.uint8, device=dev) if ws_bytes else None
    lib = _native.load()
    if amplitudes is not None:
        _check_cuda(amplitudes, "amplitudes", torch.complex64, 3, dev)
        _require(tuple(amplitudes.shape) == (p_count, n, n), "amplitudes must have shape (P, n, n)")
    b = _native.ptr(base)
    status = lib.pty_forward_batched(int(n), int(num_slices), p_count, _native.ptr(descs), b, b, b, b,
                                     _native.ptr(amplitudes), _native.ptr(ckpt), int(ckpt_every),
                                     _native.ptr(workspace), ws_bytes, _native.stream_handle(dev))
    _native.check(status, "pty_forward_batched")


def adjoint_batched(n: int, num_slices: int, descs: torch.Tensor, base: torch.Tensor, ckpt: Optional[torch.Tensor],
                    ckpt_every: int) -> None:
    """Adjoint over ``descs``: dL/dI at pattern offsets, gradients into each request's buffers."""
    dev = base.device
    p_count = descs.numel() // 88
    lib = _native.load()
    ws_bytes = int(lib.pty_adjoint_workspace_bytes(n, int(num_slices), int(ckpt_every), p_count, 1))
    workspace = torch.empty((max(ws_bytes, 1),), dtype=torch.uint8, device=dev)
    b = _native.ptr(base)
    status = lib.pty_adjoint_batched(int(n), int(num_slices), p_count, _native.ptr(descs), b, b, b, b,
                                     _native.ptr(ckpt), int(ckpt_every), b, b, _native.ptr(workspace), ws_bytes,
                                     _native.stream_handle(dev))
    _native.check(status, "pty_adjoint_batched")
