"""torch.autograd.Function around the HIP forward, adjoint and JVP kernels.

``multislice_patterns(slab_t, probe_k, weights, ...)`` is differentiable
w.r.t. the complex transmission slab, probe_k (fft2 of the probe), the
mode weights and the probe shifts (pixels); ``multislice_amplitudes``
(per-mode complex detector amplitudes) w.r.t. the slab, probe_k and the
shifts.  Both run in reverse mode (``backward``, the HIP adjoint) and in
forward mode (``jvp``, the HIP tangent recurrence ``pty_jvp``), and both
work under ``torch.func`` (``vjp``, ``jvp``, ``grad``), which the
Gauss-Newton and Levenberg-Marquardt solvers use the way the reference
uses ``jax.jvp`` / ``jax.vjp`` (``src/ptyrodactyl/jacobian/operators.py:30-185``).
Gradients follow torch's convention for complex leaves (dL/dRe + i dL/dIm),
which is the complex conjugate of what ``jax.grad`` returns for the
reference (``src/ptyrodactyl/multislice/multislice_recon.py:195-207``
conjugates it before Optax for that reason).  The integer window origins
are piecewise constant and get no gradient.
"""

from __future__ import annotations

import contextlib
import os
from typing import Optional

import torch

from . import engine

DEFAULT_CKPT_EVERY = 4
# Device bytes a checkpoint at every slice may take before the differentiable forward falls back to every 4th.
CHECKPOINT_STORE_BYTES = int(os.environ.get("PTY_CHECKPOINT_STORE_BYTES", str(64 << 30)))


def checkpoint_interval(slab_t: torch.Tensor, probe_k: torch.Tensor, origins: torch.Tensor) -> int:
    """Checkpoint interval of the differentiable forward when the caller gives none.  Square power-of-two windows
    of 256 and more (the streaming kernels) checkpoint every slice when the P·M·(S−1) stored waves fit
    ``CHECKPOINT_STORE_BYTES``: their column pass then writes each wave straight into its slot, and the adjoint
    recomputes nothing (256² × 32 slices × 4096 positions: 163 ms against 190 ms at 4, for 62 GiB of store).
    Otherwise, and for the LDS-resident windows (where every 4th costs the same time), every 4th slice."""
    h, w = int(probe_k.shape[-2]), int(probe_k.shape[-1])
    s_count = int(slab_t.shape[0])
    waves = int(origins.shape[0]) * (int(probe_k.shape[0]) if probe_k.dim() == 3 else 1)
    if h == w and h >= 256 and engine.is_fast_grid(h, w) and s_count > 1:
        if waves * (s_count - 1) * h * w * 8 <= CHECKPOINT_STORE_BYTES:
            return 1
    return DEFAULT_CKPT_EVERY


def _pair(prop):
    """(prop_y, prop_x): one factor serves both axes of a square grid."""
    return tuple(prop) if isinstance(prop, (tuple, list)) else (prop, prop)


def _empty_like_ckpt(ref: torch.Tensor) -> torch.Tensor:
    return torch.empty((0,), dtype=torch.complex64, device=ref.device)


def _transformed(t) -> bool:
    """True for a tensor that a torch.func transform (vjp, jvp, grad) wraps, or that carries a forward-mode
    tangent: such tensors have no storage the kernels could read, so they go through the Functions below."""
    if not isinstance(t, torch.Tensor):
        return False
    try:
        if torch._C._functorch.is_functorch_wrapped_tensor(t):
            return True
    except (AttributeError, RuntimeError):
        pass
    try:
        return torch.autograd.forward_ad.unpack_dual(t).tangent is not None
    except RuntimeError:
        return False


def _plain(t):
    """The tensor a torch.func transform wraps (its innermost value).  The kernels read raw memory, so the
    backward and jvp rules below hand them plain tensors; the results are first-order derivatives (the
    kernels' derivatives are not differentiated again)."""
    if not isinstance(t, torch.Tensor):
        return t
    try:
        while torch._C._functorch.is_functorch_wrapped_tensor(t):
            t = torch._C._functorch.get_unwrapped(t)
    except AttributeError:
        pass
    return t


def _kernel_scope():
    """A scope in which torch ops on plain tensors stay plain: inside a torch.func transform every op would
    otherwise wrap its result again at the transform's level, and the kernels cannot read wrapped tensors."""
    disable = getattr(torch._C, "_DisableFuncTorch", None)
    return disable() if disable is not None else contextlib.nullcontext()


def differentiable(*tensors) -> bool:
    """Whether a multislice over these operands must run through the differentiable Functions: grad mode with
    an operand that requires grad, or an operand under a torch.func transform or with a forward-mode tangent."""
    if torch.is_grad_enabled() and any(isinstance(t, torch.Tensor) and t.requires_grad for t in tensors):
        return True
    return any(_transformed(t) for t in tensors)


def _jvp_supported(probe_k: torch.Tensor) -> bool:
    m_count, h, w = probe_k.shape
    return h == w and h in engine.JVP_N


class MultisliceFunction(torch.autograd.Function):
    """(P, n, n) float32 patterns; the second output holds the adjoint's checkpoints (not differentiable)."""

    @staticmethod
    def forward(slab_t, probe_k, weights, origins, shifts_px, prop_y, prop_x, ckpt_every, want_ckpt):
        slab_c, probe_c = slab_t.detach().contiguous(), probe_k.detach().contiguous()
        w_c = weights.detach().to(torch.float32).contiguous()
        prop = (prop_y.detach(), prop_x.detach())
        if want_ckpt:
            patterns, ckpt = engine.multislice_forward_checkpointed(slab_c, probe_c, w_c, origins, shifts_px.detach(),
                                                                    prop, ckpt_every)
            if ckpt is None:
                ckpt = _empty_like_ckpt(slab_c)
        else:
            patterns, _ = engine.multislice_forward(slab_c, probe_c, w_c, origins, shifts_px.detach(), prop)
            ckpt = _empty_like_ckpt(slab_c)
        return patterns, ckpt

    @staticmethod
    def setup_context(ctx, inputs, output):
        slab_t, probe_k, weights, origins, shifts_px, prop_y, prop_x, ckpt_every, want_ckpt = inputs
        _, ckpt = output
        ctx.mark_non_differentiable(ckpt)
        # no zero tensor for the checkpoints' (absent) gradient: at 256^2 x 32 slices x 4096 positions the
        # checkpoint store is 62 GiB, and materialising its zero gradient cost 10 ms of fills per step
        ctx.set_materialize_grads(False)
        slab_c, probe_c = slab_t.detach().contiguous(), probe_k.detach().contiguous()
        w_c = weights.detach().to(torch.float32).contiguous()
        ctx.save_for_backward(slab_c, probe_c, w_c, origins, shifts_px.detach(), prop_y.detach(), prop_x.detach(),
                              ckpt)
        ctx.save_for_forward(slab_c, probe_c, w_c, origins, shifts_px.detach(), prop_y.detach(), prop_x.detach())
        ctx.ckpt_every = int(ckpt_every)
        ctx.has_ckpt = bool(want_ckpt)
        ctx.weights_dtype = weights.dtype

    @staticmethod
    def backward(ctx, grad_patterns, _grad_ckpt):
        with _kernel_scope():
            saved = (_plain(t) for t in ctx.saved_tensors)
            slab_t, probe_k, weights, origins, shifts_px, prop_y, prop_x, ckpt = saved
            if not ctx.has_ckpt:
                raise RuntimeError("multislice_patterns ran without checkpoints (grad mode was off): no reverse mode")
            if grad_patterns is None:  # the patterns do not reach the loss
                return (None,) * 9
            prop = (prop_y, prop_x)
            g = _plain(grad_patterns).to(torch.float32).contiguous()
            need_slab, need_probe, need_w = ctx.needs_input_grad[0], ctx.needs_input_grad[1], ctx.needs_input_grad[2]
            need_shift = ctx.needs_input_grad[4]
            g_slab = g_probe = g_w = g_shift = None
            if need_shift:
                g_shift = torch.zeros(shifts_px.shape, dtype=torch.float64, device=shifts_px.device)
            if need_slab or need_probe or need_shift:
                ck = ckpt if ckpt.numel() else None
                g_slab, g_probe = engine.multislice_adjoint(slab_t, probe_k, weights, origins, shifts_px, prop, g, ck,
                                                            ctx.ckpt_every, want_probe=need_probe, grad_shifts=g_shift)
            if need_w:
                # I = sum_m w_m |A_m|^2  =>  dL/dw_m = sum over positions and pixels of g |A_m|^2
                _, amp = engine.multislice_forward(slab_t, probe_k, weights, origins, shifts_px, prop,
                                                   want_patterns=False, want_amplitude=True)
                inten = amp.real.double() ** 2 + amp.imag.double() ** 2
                g_w = (g.double()[:, None] * inten).sum(dim=(0, 2, 3)).to(ctx.weights_dtype)
            return (g_slab if need_slab else None, g_probe if need_probe else None, g_w, None, g_shift, None, None,
                    None, None)

    @staticmethod
    def jvp(ctx, d_slab, d_probe_k, d_weights, _d_origins, d_shifts, _d_prop_y, _d_prop_x, _d_k, _d_want):
        with _kernel_scope():
            slab_t, probe_k, weights, origins, shifts_px, prop_y, prop_x = (_plain(t) for t in ctx.saved_tensors)
            d_slab, d_probe_k, d_weights, d_shifts = (_plain(t) for t in (d_slab, d_probe_k, d_weights, d_shifts))
            if not _jvp_supported(probe_k):
                raise ValueError("forward-mode multislice (pty_jvp) takes square windows n = 32 ... 512")
            prop = (prop_y, prop_x)
            if d_slab is None and d_probe_k is None and d_weights is None and d_shifts is None:
                n = probe_k.shape[-1]
                return torch.zeros((origins.shape[0], n, n), dtype=torch.float32, device=probe_k.device), None
            want_amp = d_weights is not None
            _, amp, d_pat, _ = engine.multislice_jvp(
                slab_t, probe_k, weights, origins, shifts_px, prop,
                d_slab.to(torch.complex64).contiguous() if d_slab is not None else None,
                d_probe_k.to(torch.complex64).contiguous() if d_probe_k is not None else None,
                d_shifts.to(torch.float64).contiguous() if d_shifts is not None else None,
                want_amplitude=want_amp)
            if want_amp:
                inten = amp.real ** 2 + amp.imag ** 2
                d_pat = d_pat + (d_weights.to(torch.float32)[None, :, None, None] * inten).sum(dim=1)
            return d_pat, None


class MultisliceAmplitudeFunction(torch.autograd.Function):
    """Per-mode fftshifted complex amplitudes (P, M, n, n), differentiable w.r.t. the slab, probe_k and shifts."""

    @staticmethod
    def forward(slab_t, probe_k, origins, shifts_px, prop_y, prop_x, ckpt_every, want_ckpt):
        slab_c, probe_c = slab_t.detach().contiguous(), probe_k.detach().contiguous()
        ones = torch.ones(probe_c.shape[0], dtype=torch.float32, device=probe_c.device)
        prop = (prop_y.detach(), prop_x.detach())
        ckpt = _empty_like_ckpt(slab_c)
        if want_ckpt:
            amp, ck = engine.multislice_amplitude_checkpointed(slab_c, probe_c, ones, origins, shifts_px.detach(),
                                                               prop, ckpt_every)
            if ck is not None:
                ckpt = ck
        else:
            _, amp = engine.multislice_forward(slab_c, probe_c, ones, origins, shifts_px.detach(), prop,
                                               want_patterns=False, want_amplitude=True)
        return amp, ckpt

    @staticmethod
    def setup_context(ctx, inputs, output):
        slab_t, probe_k, origins, shifts_px, prop_y, prop_x, ckpt_every, want_ckpt = inputs
        _, ckpt = output
        ctx.mark_non_differentiable(ckpt)
        ctx.set_materialize_grads(False)  # no zero gradient for the checkpoint store
        slab_c, probe_c = slab_t.detach().contiguous(), probe_k.detach().contiguous()
        ones = torch.ones(probe_c.shape[0], dtype=torch.float32, device=probe_c.device)
        ctx.save_for_backward(slab_c, probe_c, ones, origins, shifts_px.detach(), prop_y.detach(), prop_x.detach(),
                              ckpt)
        ctx.save_for_forward(slab_c, probe_c, ones, origins, shifts_px.detach(), prop_y.detach(), prop_x.detach())
        ctx.ckpt_every = int(ckpt_every)
        ctx.has_ckpt = bool(want_ckpt)

    @staticmethod
    def backward(ctx, grad_amp, _grad_ckpt):
        with _kernel_scope():
            slab_t, probe_k, ones, origins, shifts_px, prop_y, prop_x, ckpt = (_plain(t) for t in ctx.saved_tensors)
            grad_amp = _plain(grad_amp)
            if not ctx.has_ckpt:
                raise RuntimeError("multislice_amplitudes ran without checkpoints (grad mode was off): no reverse mode")
            if grad_amp is None:  # the amplitudes do not reach the loss
                return (None,) * 8
            prop = (prop_y, prop_x)
            g_shift = None
            if ctx.needs_input_grad[3]:
                g_shift = torch.zeros(shifts_px.shape, dtype=torch.float64, device=shifts_px.device)
            ck = ckpt if ckpt.numel() else None
            g_slab, g_probe = engine.multislice_adjoint(slab_t, probe_k, ones, origins, shifts_px, prop, None, ck,
                                                        ctx.ckpt_every, want_probe=ctx.needs_input_grad[1],
                                                        grad_amplitude=grad_amp.to(torch.complex64).contiguous(),
                                                        grad_shifts=g_shift)
            return (g_slab if ctx.needs_input_grad[0] else None, g_probe if ctx.needs_input_grad[1] else None, None,
                    g_shift, None, None, None, None)

    @staticmethod
    def jvp(ctx, d_slab, d_probe_k, _d_origins, d_shifts, _d_prop_y, _d_prop_x, _d_k, _d_want):
        with _kernel_scope():
            slab_t, probe_k, ones, origins, shifts_px, prop_y, prop_x = (_plain(t) for t in ctx.saved_tensors)
            d_slab, d_probe_k, d_shifts = (_plain(t) for t in (d_slab, d_probe_k, d_shifts))
            if not _jvp_supported(probe_k):
                raise ValueError("forward-mode multislice (pty_jvp) takes square windows n = 32 ... 512")
            if d_slab is None and d_probe_k is None and d_shifts is None:
                return torch.zeros((origins.shape[0],) + tuple(probe_k.shape), dtype=torch.complex64,
                                   device=probe_k.device), None
            _, _, _, d_amp = engine.multislice_jvp(
                slab_t, probe_k, ones, origins, shifts_px, (prop_y, prop_x),
                d_slab.to(torch.complex64).contiguous() if d_slab is not None else None,
                d_probe_k.to(torch.complex64).contiguous() if d_probe_k is not None else None,
                d_shifts.to(torch.float64).contiguous() if d_shifts is not None else None,
                want_d_patterns=False, want_d_amplitude=True)
            return d_amp, None


def multislice_patterns(slab_t: torch.Tensor, probe_k: torch.Tensor, weights: torch.Tensor, origins: torch.Tensor,
                        shifts_px: torch.Tensor, prop, ckpt_every: Optional[int] = None):
    """Differentiable (P, n, n) float32 patterns of the windowed multislice (reverse and forward mode); checkpoints
    every ``ckpt_every`` slices (default ``checkpoint_interval``)."""
    prop_y, prop_x = _pair(prop)
    if ckpt_every is None:
        ckpt_every = checkpoint_interval(slab_t, probe_k, origins)
    patterns, _ = MultisliceFunction.apply(slab_t, probe_k, weights, origins, shifts_px, prop_y, prop_x,
                                           int(ckpt_every), torch.is_grad_enabled())
    return patterns


def multislice_amplitudes(slab_t: torch.Tensor, probe_k: torch.Tensor, origins: torch.Tensor,
                          shifts_px: torch.Tensor, prop, ckpt_every: Optional[int] = None):
    """Differentiable (P, M, n, n) complex64 fftshifted amplitudes of the windowed multislice (checkpoint
    interval as for ``multislice_patterns``)."""
    prop_y, prop_x = _pair(prop)
    if ckpt_every is None:
        ckpt_every = checkpoint_interval(slab_t, probe_k, origins)
    amp, _ = MultisliceAmplitudeFunction.apply(slab_t, probe_k, origins, shifts_px, prop_y, prop_x,
                                               int(ckpt_every), torch.is_grad_enabled())
    return amp


# ---------------------------------------------------------------------------
# Batched waves: one descriptor launch over many (wave, slab, propagator) triples
# ---------------------------------------------------------------------------


def _wave_descs(n: int, slabs: torch.Tensor, probe_k: torch.Tensor, props: torch.Tensor, group,
                patterns: torch.Tensor, base: torch.Tensor, grad_slabs=None, grad_probe=None) -> torch.Tensor:
    """Descriptor records for waves i = 0 .. W-1: slab ``group[i]`` of ``slabs`` (E, S, n, n), probe_k[i],
    propagator row ``group[i]`` of ``props`` (E, 2n), pattern (or dL/dI) i of ``patterns`` (W, n, n); origin 0,
    no shift.  For the adjoint, the slab gradient of wave i goes to grad_slabs[group[i]] and its probe gradient
    to grad_probe[i]."""
    import numpy as np
    from . import _native
    count = probe_k.shape[0]
    s_count = slabs.shape[1]
    rec = np.zeros(count, dtype=np.dtype(_native.POSITION_DESC_FIELDS))
    grp = np.asarray(group, dtype=np.int64)
    idx = np.arange(count, dtype=np.int64)
    rec["slab_offset"] = engine.element_offset(slabs, base) + grp * (s_count * n * n)
    rec["slab_h"] = n
    rec["slab_w"] = n
    rec["pattern_offset"] = engine.element_offset(patterns, base) + idx * n * n
    rec["probe_offset"] = engine.element_offset(probe_k, base) + idx * n * n
    rec["prop_offset"] = engine.element_offset(props, base) + grp * (2 * n)
    if grad_slabs is not None:
        rec["grad_offset"] = engine.element_offset(grad_slabs, base) + grp * (s_count * n * n)
        rec["gprobe_offset"] = engine.element_offset(grad_probe, base) + idx * n * n
    return torch.from_numpy(rec.view(np.uint8)).to(slabs.device)


class BatchedWavesFunction(torch.autograd.Function):
    """Intensities (W, n, n) float32 (fftshifted) of W waves, each with its own probe_k and one of E (slab,
    propagator) pairs, in one ``pty_forward_batched`` launch with checkpoints; differentiable w.r.t. the slabs
    and every probe_k through one ``pty_adjoint_batched`` launch (float32 atomics add the waves that share a
    slab).  The ensemble axes use it for every (position, sample) wave: the reference maps cbed_image over
    them (src/ptyrodactyl/multislice/simulations.py:769-887)."""

    @staticmethod
    def forward(slabs, probe_k, props, group, ckpt_every):
        n, s_count = probe_k.shape[-1], slabs.shape[1]
        base = engine.arena_base(slabs.device)
        slabs_c, pk_c, props_c = slabs.detach().contiguous(), probe_k.detach().contiguous(), props.contiguous()
        patterns = torch.empty(probe_k.shape, dtype=torch.float32, device=slabs.device)
        nc = (s_count - 1) // ckpt_every  # checkpoints for the backward's segment recompute
        ckpt = torch.empty((probe_k.shape[0], nc, n, n), dtype=torch.complex64, device=slabs.device) if nc else None
        descs = _wave_descs(n, slabs_c, pk_c, props_c, group, patterns, base)
        engine.forward_batched(n, s_count, descs, base, ckpt, ckpt_every if nc else 0)
        return patterns, (ckpt if ckpt is not None else _empty_like_ckpt(slabs))

    @staticmethod
    def setup_context(ctx, inputs, output):
        slabs, probe_k, props, group, ckpt_every = inputs
        ctx.save_for_backward(slabs.detach().contiguous(), probe_k.detach().contiguous(), props.contiguous(),
                              output[1])
        ctx.group, ctx.ckpt_every = group, ckpt_every
        ctx.mark_non_differentiable(output[1])
        ctx.set_materialize_grads(False)  # no zero gradient for the checkpoints

    @staticmethod
    def backward(ctx, grad_patterns, _grad_ckpt):
        if grad_patterns is None:  # the patterns do not reach the loss
            return None, None, None, None, None
        slabs, probe_k, props, ckpt = ctx.saved_tensors
        n, s_count = probe_k.shape[-1], slabs.shape[1]
        base = engine.arena_base(slabs.device)
        gi = grad_patterns.to(torch.float32).contiguous()
        grad_slabs = torch.zeros_like(slabs)
        grad_probe = torch.zeros_like(probe_k)
        descs = _wave_descs(n, slabs, probe_k, props, ctx.group, gi, base, grad_slabs, grad_probe)
        nc = ckpt.shape[1] if ckpt.ndim == 4 else 0
        engine.adjoint_batched(n, s_count, descs, base, ckpt if nc else None, ctx.ckpt_every)
        return grad_slabs, grad_probe, None, None, None


def batched_wave_patterns(slabs: torch.Tensor, probe_k: torch.Tensor, props: torch.Tensor, group,
                          ckpt_every: int = DEFAULT_CKPT_EVERY) -> torch.Tensor:
    """(W, n, n) float32 intensities of the waves (see ``BatchedWavesFunction``)."""
    return BatchedWavesFunction.apply(slabs, probe_k, props, group, int(ckpt_every))[0]
