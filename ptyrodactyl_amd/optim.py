"""Optax update rules of the reconstructors, applied in place on the GPU.

``OPTIMIZERS`` mirrors the reference registry
(``src/ptyrodactyl/multislice/multislice_recon.py:70-83``): optax 0.2.3
``adam(lr)``, ``adagrad(lr, initial_accumulator_value=0.0, eps=1e-8)`` and
``rmsprop(lr, decay=0.9, eps=1e-8, initial_scale=0.0)``.  Each step is one
launch of the fused HIP kernel behind ``pty_optimizer_update`` (parameter
and both moments read and written once).  The rules receive torch's
gradient dL/dRe + i dL/dIm, which is the ``conj(jax.grad)`` the reference
hands to Optax (``_apply_optimizer_step`` ``:166-210``), so no conjugation
happens here.  Unlike Optax the update is in place: the reconstructors own
their parameter tensors.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Dict

import torch

from . import _native

_RULES = {"adam": 0, "adagrad": 1, "rmsprop": 2}
_DTYPES = {torch.float32: (0, False), torch.float64: (1, False), torch.complex64: (0, True),
           torch.complex128: (1, True)}


@dataclass
class OptaxState:
    """Moments of one parameter tensor: ``mu`` (adam's first moment, parameter dtype), ``nu`` (real
    second moment or sum of squares) and the number of updates applied."""
    mu: torch.Tensor
    nu: torch.Tensor
    count: int = 0


class OptaxRule:
    """One named rule at one learning rate (``init`` / ``update_``)."""

    def __init__(self, name: str, learning_rate: float):
        if name not in _RULES:
            raise ValueError(f"Unknown optimizer: {name}")
        self.name = name
        self.rule = _RULES[name]
        self.learning_rate = float(learning_rate)

    def init(self, params: torch.Tensor) -> OptaxState:
        if params.dtype not in _DTYPES:
            raise ValueError(f"parameters must be float32/64 or complex64/128, got {params.dtype}")
        real = params.real.dtype if params.is_complex() else params.dtype
        mu = torch.zeros_like(params) if self.rule == 0 else torch.zeros(1, dtype=params.dtype,
                                                                            device=params.device)
        return OptaxState(mu, torch.zeros(params.shape, dtype=real, device=params.device), 0)

    def update_(self, params: torch.Tensor, grads: torch.Tensor, state: OptaxState) -> None:
        """p <- p + (-lr) u(g) in place, advancing ``state``."""
        if not (params.is_cuda and params.is_contiguous()):
            raise ValueError("parameters must be a contiguous GPU tensor (the update rules have no CPU fallback)")
        fp64, cplx = _DTYPES[params.dtype]
        g = grads.detach().to(dtype=params.dtype, device=params.device).contiguous()
        if g.shape != params.shape:
            raise ValueError(f"gradient shape {tuple(g.shape)} does not match parameters {tuple(params.shape)}")
        state.count += 1
        lib = _native.load()
        with torch.cuda.device(params.device):
            status = lib.pty_optimizer_update(
                self.rule, fp64, int(cplx), params.numel(), _native.ptr(params), _native.ptr(g),
                _native.ptr(state.mu) if self.rule == 0 else 0, _native.ptr(state.nu), state.count,
                self.learning_rate, _native.stream_handle(params.device))
        _native.check(status, "pty_optimizer_update")


OPTIMIZERS: Dict[str, Callable[[float], OptaxRule]] = {
    "adam": lambda learning_rate: OptaxRule("adam", learning_rate),
    "adagrad": lambda learning_rate: OptaxRule("adagrad", learning_rate),
    "rmsprop": lambda learning_rate: OptaxRule("rmsprop", learning_rate),
}


def get_optimizer(optimizer_name: str, learning_rate: float) -> OptaxRule:
    """``_get_optimizer`` (``multislice_recon.py:95-124``): ``ValueError`` for an unknown name."""
    if optimizer_name not in OPTIMIZERS:
        raise ValueError(f"Unknown optimizer: {optimizer_name}")
    return OPTIMIZERS[optimizer_name](learning_rate)
