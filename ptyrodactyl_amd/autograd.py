"""torch.autograd.Function around the HIP forward and adjoint kernels.

``multislice_patterns(slab_t, probe_k, ...)`` is differentiable w.r.t. the
complex transmission slab and probe_k (fft2 of the probe); gradients follow
torch's convention for complex leaves (dL/dRe + i dL/dIm), which is the
complex conjugate of what ``jax.grad`` returns for the reference
(``src/ptyrodactyl/multislice/multislice_recon.py:195-207`` conjugates it
before Optax for that reason).
"""

from __future__ import annotations

import torch

from . import engine

DEFAULT_CKPT_EVERY = 4


class MultisliceFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, slab_t, probe_k, weights, origins, shifts_px, prop, ckpt_every):
        need = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
        slab_c, probe_c = slab_t.detach().contiguous(), probe_k.detach().contiguous()
        if not need:
            patterns, _ = engine.multislice_forward(slab_c, probe_c, weights, origins, shifts_px, prop)
            return patterns
        patterns, ckpt = engine.multislice_forward_checkpointed(slab_c, probe_c, weights, origins, shifts_px, prop,
                                                                ckpt_every)
        ctx.save_for_backward(slab_c, probe_c, weights, origins, shifts_px, prop, ckpt)
        ctx.ckpt_every = ckpt_every
        return patterns

    @staticmethod
    def backward(ctx, grad_patterns):
        slab_t, probe_k, weights, origins, shifts_px, prop, ckpt = ctx.saved_tensors
        g = grad_patterns.to(torch.float32).contiguous()
        g_slab, g_probe = engine.multislice_adjoint(slab_t, probe_k, weights, origins, shifts_px, prop, g, ckpt,
                                                    ctx.ckpt_every, want_probe=ctx.needs_input_grad[1])
        return (g_slab if ctx.needs_input_grad[0] else None, g_probe, None, None, None, None, None)


def multislice_patterns(slab_t: torch.Tensor, probe_k: torch.Tensor, weights: torch.Tensor, origins: torch.Tensor,
                        shifts_px: torch.Tensor, prop: torch.Tensor, ckpt_every: int = DEFAULT_CKPT_EVERY):
    """Differentiable (P, n, n) float32 patterns of the windowed multislice."""
    return MultisliceFunction.apply(slab_t, probe_k, weights, origins, shifts_px, prop, int(ckpt_every))
