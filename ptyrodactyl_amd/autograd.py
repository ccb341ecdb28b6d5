"""torch.autograd.Function around the HIP forward and adjoint kernels.

``multislice_patterns(slab_t, probe_k, weights, ...)`` is differentiable
w.r.t. the complex transmission slab, probe_k (fft2 of the probe), the
mode weights and the probe shifts (pixels); ``multislice_amplitudes``
(per-mode complex detector amplitudes) w.r.t. the slab, probe_k and the
shifts.  Gradients follow torch's convention for complex leaves
(dL/dRe + i dL/dIm), which is the complex conjugate of what ``jax.grad``
returns for the reference (``src/ptyrodactyl/multislice/multislice_recon.py:195-207``
conjugates it before Optax for that reason).  The integer window origins
are piecewise constant and get no gradient.
"""

from __future__ import annotations

import torch

from . import engine

DEFAULT_CKPT_EVERY = 4


def _pair(prop):
    """(prop_y, prop_x): one factor serves both axes of a square grid."""
    return tuple(prop) if isinstance(prop, (tuple, list)) else (prop, prop)


class MultisliceFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, slab_t, probe_k, weights, origins, shifts_px, prop, ckpt_every):
        need_field = ctx.needs_input_grad[0] or ctx.needs_input_grad[1] or ctx.needs_input_grad[4]
        slab_c, probe_c = slab_t.detach().contiguous(), probe_k.detach().contiguous()
        w_c = weights.detach().to(torch.float32).contiguous()
        if not need_field:
            patterns, _ = engine.multislice_forward(slab_c, probe_c, w_c, origins, shifts_px, prop)
            ckpt = None
        else:
            patterns, ckpt = engine.multislice_forward_checkpointed(slab_c, probe_c, w_c, origins, shifts_px, prop,
                                                                    ckpt_every)
        ctx.save_for_backward(slab_c, probe_c, w_c, origins, shifts_px.detach(), *_pair(prop), ckpt)
        ctx.ckpt_every = ckpt_every
        ctx.weights_dtype = weights.dtype
        return patterns

    @staticmethod
    def backward(ctx, grad_patterns):
        slab_t, probe_k, weights, origins, shifts_px, prop_y, prop_x, ckpt = ctx.saved_tensors
        prop = (prop_y, prop_x)
        g = grad_patterns.to(torch.float32).contiguous()
        g_slab = g_probe = g_w = g_shift = None
        if ctx.needs_input_grad[4]:
            g_shift = torch.zeros(shifts_px.shape, dtype=torch.float64, device=shifts_px.device)
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1] or ctx.needs_input_grad[4]:
            g_slab, g_probe = engine.multislice_adjoint(slab_t, probe_k, weights, origins, shifts_px, prop, g, ckpt,
                                                        ctx.ckpt_every, want_probe=ctx.needs_input_grad[1],
                                                        grad_shifts=g_shift)
        if ctx.needs_input_grad[2]:
            # I = sum_m w_m |A_m|^2  =>  dL/dw_m = sum over positions and pixels of g |A_m|^2
            _, amp = engine.multislice_forward(slab_t, probe_k, weights, origins, shifts_px, prop,
                                               want_patterns=False, want_amplitude=True)
            inten = amp.real.double() ** 2 + amp.imag.double() ** 2
            g_w = (g.double()[:, None] * inten).sum(dim=(0, 2, 3)).to(ctx.weights_dtype)
        return (g_slab if ctx.needs_input_grad[0] else None, g_probe, g_w, None, g_shift, None, None)


class MultisliceAmplitudeFunction(torch.autograd.Function):
    """Per-mode fftshifted complex amplitudes (P, M, n, n), differentiable w.r.t. the slab, probe_k and shifts."""

    @staticmethod
    def forward(ctx, slab_t, probe_k, origins, shifts_px, prop, ckpt_every):
        slab_c, probe_c = slab_t.detach().contiguous(), probe_k.detach().contiguous()
        ones = torch.ones(probe_c.shape[0], dtype=torch.float32, device=probe_c.device)
        _, amp = engine.multislice_forward(slab_c, probe_c, ones, origins, shifts_px, prop, want_patterns=False,
                                           want_amplitude=True)
        ckpt = None
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1] or ctx.needs_input_grad[3]:
            _, ckpt = engine.multislice_forward_checkpointed(slab_c, probe_c, ones, origins, shifts_px, prop,
                                                             ckpt_every)
        ctx.save_for_backward(slab_c, probe_c, ones, origins, shifts_px.detach(), *_pair(prop), ckpt)
        ctx.ckpt_every = ckpt_every
        return amp

    @staticmethod
    def backward(ctx, grad_amp):
        slab_t, probe_k, ones, origins, shifts_px, prop_y, prop_x, ckpt = ctx.saved_tensors
        prop = (prop_y, prop_x)
        g_shift = None
        if ctx.needs_input_grad[3]:
            g_shift = torch.zeros(shifts_px.shape, dtype=torch.float64, device=shifts_px.device)
        g_slab, g_probe = engine.multislice_adjoint(slab_t, probe_k, ones, origins, shifts_px, prop, None, ckpt,
                                                    ctx.ckpt_every, want_probe=ctx.needs_input_grad[1],
                                                    grad_amplitude=grad_amp.to(torch.complex64).contiguous(),
                                                    grad_shifts=g_shift)
        return (g_slab if ctx.needs_input_grad[0] else None, g_probe, None, g_shift, None, None)


def multislice_patterns(slab_t: torch.Tensor, probe_k: torch.Tensor, weights: torch.Tensor, origins: torch.Tensor,
                        shifts_px: torch.Tensor, prop: torch.Tensor, ckpt_every: int = DEFAULT_CKPT_EVERY):
    """Differentiable (P, n, n) float32 patterns of the windowed multislice."""
    return MultisliceFunction.apply(slab_t, probe_k, weights, origins, shifts_px, prop, int(ckpt_every))


def multislice_amplitudes(slab_t: torch.Tensor, probe_k: torch.Tensor, origins: torch.Tensor,
                          shifts_px: torch.Tensor, prop: torch.Tensor, ckpt_every: int = DEFAULT_CKPT_EVERY):
    """Differentiable (P, M, n, n) complex64 fftshifted amplitudes of the windowed multislice."""
    return MultisliceAmplitudeFunction.apply(slab_t, probe_k, origins, shifts_px, prop, int(ckpt_every))
