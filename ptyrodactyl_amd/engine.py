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
# Waves (positions x modes) per streaming batch at n = 256; the batch's waves live in a device workspace that
# should stay cache-resident.  The library runs it as three lanes of 128 waves on their own streams.
DEFAULT_BATCH_WAVES = int(os.environ.get("PTY_MAX_BATCH_WAVES", "384"))
# largest edge of the runtime-size kernels (any h x w that is not a square power of two)
GRID_MAX_EDGE = 4096


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


def is_fast_grid(h: int, w: int) -> bool:
    """Square power-of-two windows run on the dedicated kernels; every other h x w on the runtime-size ones."""
    return h == w and h in SUPPORTED_N


def _prop_pair(prop, h: int, w: int, dev):
    """(prop_y, prop_x) from one factor (square grids) or a pair; each includes 1/h or 1/w."""
    prop_y, prop_x = prop if isinstance(prop, (tuple, list)) else (prop, prop)
    _check_cuda(prop_y, "prop_y", torch.complex64, 1, dev)
    _check_cuda(prop_x, "prop_x", torch.complex64, 1, dev)
    _require(prop_y.shape[0] == h and prop_x.shape[0] == w,
             f"propagator factors must have {h} and {w} entries, got {prop_y.shape[0]} and {prop_x.shape[0]}")
    return prop_y, prop_x


def _grid_operands(slab_t, probe_k, weights, origins, shifts_px, prop):
    _require(isinstance(slab_t, torch.Tensor) and slab_t.is_cuda,
             "slab must be a GPU tensor (the multislice path has no CPU fallback)")
    dev = slab_t.device
    _check_cuda(slab_t, "slab", torch.complex64, 3, dev)
    _check_cuda(probe_k, "probe_k", torch.complex64, 3, dev)
    m_count, h, w = probe_k.shape
    _require(1 <= h <= GRID_MAX_EDGE and 1 <= w <= GRID_MAX_EDGE,
             f"grid {h}x{w}: each edge must be in [1, {GRID_MAX_EDGE}]")
    s_count, hs, ws = slab_t.shape
    _require(s_count >= 1, "at least one slice is required")
    _require(hs >= h and ws >= w, f"slab {hs}x{ws} is smaller than the {h}x{w} window")
    _check_cuda(weights, "weights", torch.float32, 1, dev)
    _require(weights.shape[0] == m_count, "weights must have shape (M,)")
    _check_cuda(origins, "origins", torch.int32, 2, dev)
    _check_cuda(shifts_px, "shifts_px", torch.float64, 2, dev)
    p_count = origins.shape[0]
    _require(p_count >= 1 and origins.shape == (p_count, 2) and shifts_px.shape == (p_count, 2),
             "origins and shifts must both have shape (P, 2) with P >= 1")
    prop_y, prop_x = _prop_pair(prop, h, w, dev)
    return dev, h, w, s_count, hs, ws, m_count, p_count, prop_y, prop_x


def grid_workspace_bytes(h: int, w: int, num_slices: int, ckpt_every: int, num_positions: int, num_modes: int,
                         adjoint: bool) -> int:
    """Workspace of ``pty_forward_grid`` / ``pty_adjoint_grid`` for every (position, mode) wave in flight."""
    lib = _native.load()
    return int(lib.pty_grid_workspace_bytes(int(h), int(w), int(num_slices), int(ckpt_every),
                                            int(num_positions) * int(num_modes), 1 if adjoint else 0))


def _grid_forward(slab_t, probe_k, weights, origins, shifts_px, prop, want_patterns, want_amplitude, patterns_out,
                  ckpt_every):
    """Forward on any h x w grid (``pty_forward_grid``); checkpoints when ckpt_every >= 1."""
    dev, h, w, s_count, hs, ws, m_count, p_count, prop_y, prop_x = _grid_operands(slab_t, probe_k, weights, origins,
                                                                                shifts_px, prop)
    patterns = None
    if want_patterns:
        if patterns_out is not None:
            _check_cuda(patterns_out, "patterns_out", torch.float32, 3, dev)
            _require(tuple(patterns_out.shape) == (p_count, h, w), "patterns_out has the wrong shape")
            patterns = patterns_out
        else:
            patterns = torch.empty((p_count, h, w), dtype=torch.float32, device=dev)
    amplitude = torch.empty((p_count, m_count, h, w), dtype=torch.complex64, device=dev) if want_amplitude else None
    ckpt = None
    nc = num_checkpoints(s_count, ckpt_every) if ckpt_every else 0
    if ckpt_every:
        ckpt = torch.empty((p_count * m_count, nc, h, w), dtype=torch.complex64, device=dev)
    lib = _native.load()
    # size for every (position, mode) wave: amplitude-only forwards run each wave as its own task, and the
    # streaming kernels (grids of 2^14 pixels or more) hold all M waves of a position at once, so a scan with
    # fewer positions than modes would otherwise fall back to one workgroup per position
    ws_bytes = grid_workspace_bytes(h, w, s_count, 1, p_count, m_count, adjoint=False)
    workspace = torch.empty((ws_bytes,), dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        status = lib.pty_forward_grid(
            h, w, s_count, p_count, m_count, _native.ptr(slab_t), hs, ws, _native.ptr(probe_k), _native.ptr(weights),
            _native.ptr(origins), _native.ptr(shifts_px), _native.ptr(prop_y), _native.ptr(prop_x),
            _native.ptr(patterns), _native.ptr(amplitude), _native.ptr(ckpt) if nc else 0, max(int(ckpt_every), 1),
            _native.ptr(workspace), ws_bytes, _native.stream_handle(dev))
    _native.check(status, "pty_forward_grid")
    return patterns, amplitude, ckpt


def _grid_adjoint(slab_t, probe_k, weights, origins, shifts_px, prop, grad_patterns, ckpt, ckpt_every, grad_slab,
                  grad_probe_k, want_probe, grad_amplitude, grad_shifts):
    dev, h, w, s_count, hs, ws, m_count, p_count, prop_y, prop_x = _grid_operands(slab_t, probe_k, weights, origins,
                                                                                shifts_px, prop)
    k = int(ckpt_every)
    _require(k >= 1, "ckpt_every must be >= 1")
    nc = num_checkpoints(s_count, k)
    if grad_amplitude is not None:
        _check_cuda(grad_amplitude, "grad_amplitude", torch.complex64, 4, dev)
        _require(tuple(grad_amplitude.shape) == (p_count, m_count, h, w), "grad_amplitude must have shape (P, M, h, w)")
        grad_patterns = None
    else:
        _check_cuda(grad_patterns, "grad_patterns", torch.float32, 3, dev)
        _require(tuple(grad_patterns.shape) == (p_count, h, w), "grad_patterns must have shape (P, h, w)")
    if nc:
        _check_cuda(ckpt, "checkpoints", torch.complex64, 4, dev)
        _require(tuple(ckpt.shape) == (p_count * m_count, nc, h, w), "checkpoint buffer has the wrong shape")
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
    # the streaming adjoint holds the M waves of each position in flight: size for P * M waves
    ws_bytes = grid_workspace_bytes(h, w, s_count, k, p_count, m_count, adjoint=True)
    workspace = torch.empty((ws_bytes,), dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        status = lib.pty_adjoint_grid(
            h, w, s_count, p_count, m_count, _native.ptr(slab_t), hs, ws, _native.ptr(probe_k), _native.ptr(weights),
            _native.ptr(origins), _native.ptr(shifts_px), _native.ptr(prop_y), _native.ptr(prop_x),
            _native.ptr(grad_patterns), _native.ptr(grad_amplitude), _native.ptr(ckpt) if nc else 0, k,
            _native.ptr(grad_slab), _native.ptr(grad_probe_k), _native.ptr(grad_shifts), _native.ptr(workspace),
            ws_bytes, _native.stream_handle(dev))
    _native.check(status, "pty_adjoint_grid")
    return grad_slab, grad_probe_k


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

    slab_t (S, Hs, Ws) complex64 transmission; probe_k (M, h, w) complex64
    fft2 of the probe modes; weights (M,) float32; origins (P, 2) int32;
    shifts_px (P, 2) float64; prop the separable propagator factor including
    1/n ((n,) complex64 on a square grid) or a pair (prop_y (h,), prop_x (w,)).
    Returns (patterns (P, h, w) float32 fftshifted, amplitude (P, M, h, w)
    complex64 fftshifted) with None for outputs not requested.  Square
    power-of-two windows run on ``pty_forward`` (``max_batch_waves`` caps its
    streaming batch); any other h x w (edges up to 4096) on ``pty_forward_grid``.
    """
    _require(isinstance(slab_t, torch.Tensor) and slab_t.is_cuda,
             "slab must be a GPU tensor (the multislice path has no CPU fallback)")
    dev = slab_t.device
    _check_cuda(slab_t, "slab", torch.complex64, 3, dev)
    _check_cuda(probe_k, "probe_k", torch.complex64, 3, dev)
    if not is_fast_grid(probe_k.shape[1], probe_k.shape[2]):
        _require(want_patterns or want_amplitude, "nothing requested")
        patterns, amplitude, _ = _grid_forward(slab_t, probe_k, weights, origins, shifts_px, prop,
                                               want_patterns, want_amplitude, patterns_out, 0)
        return patterns, amplitude
    m_count, n, _ = probe_k.shape
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
    prop_y, prop_x = _prop_pair(prop, n, n, dev)
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
            _native.ptr(weights), _native.ptr(origins), _native.ptr(shifts_px), _native.ptr(prop_y),
            _native.ptr(prop_x), _native.ptr(patterns), _native.ptr(amplitude), _native.ptr(workspace),
            ws_bytes, _native.stream_handle(dev))
    _native.check(status, "pty_forward")
    return patterns, amplitude


class CapturedForward:
    """A forward over fixed operands captured once in a HIP graph and replayed.

    The launch sequence of :func:`multislice_forward` (one launch per position
    group for n <= 128, 2S + 2 launches per batch for n >= 256; one launch, or
    2S + 2 per batch when streaming, on the runtime-size grids) is recorded on
    a capture stream after one eager warm-up call, together with its
    workspace; :meth:`replay` re-runs it on the current stream with whatever
    the operand tensors hold at that time (update them in place between
    replays, e.g. the object in a reconstruction loop).  The captured
    sequence contains kernels only (no memset or copy nodes).
    """

    def __init__(self, slab_t, probe_k, weights, origins, shifts_px, prop, max_batch_waves: Optional[int] = None):
        if is_fast_grid(probe_k.shape[-2], probe_k.shape[-1]):
            _check_forward_operands(slab_t, probe_k, weights, origins, shifts_px, prop)
        else:
            _grid_operands(slab_t, probe_k, weights, origins, shifts_px, prop)
        self.operands = (slab_t, probe_k, weights, origins, shifts_px, prop)
        p_count, h, w = origins.shape[0], probe_k.shape[-2], probe_k.shape[-1]
        self.patterns = torch.empty((p_count, h, w), dtype=torch.float32, device=slab_t.device)
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
            _native.ptr(origins), _native.ptr(shifts_px), *_prop_ptrs(prop),
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
            _native.ptr(origins), _native.ptr(shifts_px), *_prop_ptrs(prop),
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
    _prop_pair(prop, n, n, dev)
    return dev, n, s_count, hs, ws, m_count, p_count


def _prop_ptrs(prop):
    prop_y, prop_x = prop if isinstance(prop, (tuple, list)) else (prop, prop)
    return _native.ptr(prop_y), _native.ptr(prop_x)


def num_checkpoints(num_slices: int, ckpt_every: int) -> int:
    return (num_slices - 1) // ckpt_every


def multislice_forward_checkpointed(slab_t, probe_k, weights, origins, shifts_px, prop, ckpt_every: int,
                                    max_batch_waves: Optional[int] = None):
    """Forward that also returns the wave entering slices k, 2k, ... for the adjoint.

    Returns (patterns (P, n, n) float32, checkpoints (P*M, (S-1)//k, n, n) complex64).
    """
    patterns, _, ckpt = _forward_with_checkpoints(slab_t, probe_k, weights, origins, shifts_px, prop, ckpt_every,
                                                  True, False, max_batch_waves)
    return patterns, ckpt


def multislice_amplitude_checkpointed(slab_t, probe_k, weights, origins, shifts_px, prop, ckpt_every: int,
                                      max_batch_waves: Optional[int] = None):
    """The per-mode fftshifted amplitudes (P, M, n, n) complex64 and the adjoint's checkpoints from one forward."""
    _, amplitude, ckpt = _forward_with_checkpoints(slab_t, probe_k, weights, origins, shifts_px, prop, ckpt_every,
                                                   False, True, max_batch_waves)
    return amplitude, ckpt


def _forward_with_checkpoints(slab_t, probe_k, weights, origins, shifts_px, prop, ckpt_every, want_patterns,
                              want_amplitude, max_batch_waves):
    if isinstance(probe_k, torch.Tensor) and probe_k.ndim == 3 and not is_fast_grid(probe_k.shape[1],
                                                                                     probe_k.shape[2]):
        _require(int(ckpt_every) >= 1, "ckpt_every must be >= 1")
        return _grid_forward(slab_t, probe_k, weights, origins, shifts_px, prop, want_patterns, want_amplitude, None,
                             int(ckpt_every))
    dev, n, s_count, hs, ws, m_count, p_count = _check_forward_operands(slab_t, probe_k, weights, origins,
                                                                        shifts_px, prop, GRADIENT_N,
                                                                        "checkpoints")
    _require(int(ckpt_every) >= 1, "ckpt_every must be >= 1")
    nc = num_checkpoints(s_count, int(ckpt_every))
    patterns = torch.empty((p_count, n, n), dtype=torch.float32, device=dev) if want_patterns else None
    amplitude = torch.empty((p_count, m_count, n, n), dtype=torch.complex64, device=dev) if want_amplitude else None
    ckpt = torch.empty((p_count * m_count, nc, n, n), dtype=torch.complex64, device=dev)
    ws_bytes = workspace_bytes(n, p_count, m_count, max_batch_waves)
    workspace = torch.empty((max(ws_bytes, 1),), dtype=torch.uint8, device=dev) if ws_bytes else None
    lib = _native.load()
    with torch.cuda.device(dev):
        status = lib.pty_forward_checkpointed(
            n, s_count, p_count, m_count, _native.ptr(slab_t), hs, ws, _native.ptr(probe_k), _native.ptr(weights),
            _native.ptr(origins), _native.ptr(shifts_px), *_prop_ptrs(prop),
            _native.ptr(patterns), _native.ptr(amplitude), _native.ptr(ckpt) if nc else 0, int(ckpt_every),
            _native.ptr(workspace), ws_bytes, _native.stream_handle(dev))
    _native.check(status, "pty_forward_checkpointed")
    return patterns, amplitude, ckpt


JVP_N = (32, 64, 128, 256, 512)
# default JVP batch: 256 MiB of (primal, tangent) wave pairs
JVP_BATCH_BYTES = 256 << 20


def multislice_jvp(slab_t, probe_k, weights, origins, shifts_px, prop, d_slab: Optional[torch.Tensor] = None,
                   d_probe_k: Optional[torch.Tensor] = None, d_shifts: Optional[torch.Tensor] = None, *,
                   want_patterns: bool = False, want_amplitude: bool = False, want_d_patterns: bool = True,
                   want_d_amplitude: bool = False, max_batch_waves: Optional[int] = None):
    """Forward-mode multislice (``pty_jvp``): tangents of the patterns / amplitudes for tangents of the slab,
    probe_k and shifts (each None = zero).

    Operands as for :func:`multislice_forward` (square power-of-two windows up to 512).  Returns
    (patterns, amplitude, d_patterns, d_amplitude), None for outputs not requested; the primal outputs equal
    :func:`multislice_forward`'s to rounding, the tangents are complex64 / float32 in the fftshifted layout.
    """
    dev, n, s_count, hs, ws, m_count, p_count = _check_forward_operands(slab_t, probe_k, weights, origins,
                                                                        shifts_px, prop, JVP_N, "the JVP")
    _require(want_d_patterns or want_d_amplitude, "no tangent output requested")
    if d_slab is not None:
        _check_cuda(d_slab, "d_slab", torch.complex64, 3, dev)
        _require(d_slab.shape == slab_t.shape, "d_slab must match the slab")
    if d_probe_k is not None:
        _check_cuda(d_probe_k, "d_probe_k", torch.complex64, 3, dev)
        _require(d_probe_k.shape == probe_k.shape, "d_probe_k must match probe_k")
    if d_shifts is not None:
        _check_cuda(d_shifts, "d_shifts", torch.float64, 2, dev)
        _require(tuple(d_shifts.shape) == (p_count, 2), "d_shifts must have shape (P, 2)")
    patterns = torch.empty((p_count, n, n), dtype=torch.float32, device=dev) if want_patterns else None
    amplitude = torch.empty((p_count, m_count, n, n), dtype=torch.complex64, device=dev) if want_amplitude else None
    d_patterns = torch.empty((p_count, n, n), dtype=torch.float32, device=dev) if want_d_patterns else None
    d_amplitude = (torch.empty((p_count, m_count, n, n), dtype=torch.complex64, device=dev)
                   if want_d_amplitude else None)
    cap = int(max_batch_waves) if max_batch_waves is not None else max(m_count, JVP_BATCH_BYTES // (16 * n * n))
    lib = _native.load()
    ws_bytes = int(lib.pty_jvp_workspace_bytes(n, p_count, m_count, cap))
    workspace = torch.empty((ws_bytes,), dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        status = lib.pty_jvp(
            n, s_count, p_count, m_count, _native.ptr(slab_t), hs, ws, _native.ptr(probe_k), _native.ptr(weights),
            _native.ptr(origins), _native.ptr(shifts_px), *_prop_ptrs(prop), _native.ptr(d_slab),
            _native.ptr(d_probe_k), _native.ptr(d_shifts), _native.ptr(patterns), _native.ptr(amplitude),
            _native.ptr(d_patterns), _native.ptr(d_amplitude), _native.ptr(workspace), ws_bytes,
            _native.stream_handle(dev))
    _native.check(status, "pty_jvp")
    return patterns, amplitude, d_patterns, d_amplitude


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
    if isinstance(probe_k, torch.Tensor) and probe_k.ndim == 3 and not is_fast_grid(probe_k.shape[1],
                                                                                     probe_k.shape[2]):
        return _grid_adjoint(slab_t, probe_k, weights, origins, shifts_px, prop, grad_patterns, ckpt, ckpt_every,
                             grad_slab, grad_probe_k, want_probe, grad_amplitude, grad_shifts)
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
            _native.ptr(origins), _native.ptr(shifts_px), *_prop_ptrs(prop),
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
