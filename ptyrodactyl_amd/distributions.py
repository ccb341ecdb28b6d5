"""Ensemble-axis carriers: distributions over probe modes, position jitter and
partial coherence, and the per-sample axis update.

Mirrors ``src/ptyrodactyl/types/distributions.py:38-215`` (``ReductionMode``,
``Distribution``, ``create_distribution``, ``create_trivial_distribution``,
``TRIVIAL``) and the ``AxisUpdate`` carrier of
``src/ptyrodactyl/types/electron_types.py:227-486`` with torch tensors and
the same two validation tiers as :mod:`.types` (structure: ``ValueError``;
values: :class:`RuntimeCheckError` with the reference's messages).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from enum import Enum
from typing import Callable, Optional, Tuple

import torch

from .types import error_if


class ReductionMode(str, Enum):
    """How an axis is summed: amplitudes before |.|^2 (coherent) or intensities after (incoherent)."""

    COHERENT = "coherent"
    INCOHERENT = "incoherent"


@dataclass(frozen=True)
class Distribution:
    """Weighted samples of one ensemble axis: samples (N, D) float64, weights (N,) normalised to sum 1."""

    samples: torch.Tensor
    weights: torch.Tensor
    reduction: ReductionMode
    axis_id: Optional[str] = None

    def bind(self, binder: Callable[["Distribution"], Callable]) -> Callable:
        return binder(self)


def _normalize_probability_weights(weights: torch.Tensor) -> torch.Tensor:
    clipped = torch.clamp(weights.to(torch.float64), min=0.0)
    total = clipped.sum()
    if float(total) > 0.0:
        return clipped / total
    return torch.ones_like(clipped) / clipped.shape[0]


def create_distribution(samples, weights, reduction=ReductionMode.INCOHERENT,
                        axis_id: Optional[str] = None) -> Distribution:
    """``distributions.py:105-181``: (N, D) samples, (N,) non-negative weights with positive total."""
    s = samples.to(torch.float64) if isinstance(samples, torch.Tensor) else torch.as_tensor(samples,
                                                                                            dtype=torch.float64)
    w = weights.to(torch.float64) if isinstance(weights, torch.Tensor) else torch.as_tensor(weights,
                                                                                           dtype=torch.float64)
    if s.ndim != 2:
        raise ValueError("samples must have shape (N, D)")
    if w.ndim != 1:
        raise ValueError("weights must have shape (N,)")
    if s.shape[0] <= 0:
        raise ValueError("samples must contain at least one row")
    if s.shape[0] != w.shape[0]:
        raise ValueError("samples and weights must share leading dimension")
    mode = ReductionMode(reduction)
    error_if(~torch.isfinite(s).all(), "samples must be finite")
    error_if(~torch.isfinite(w).all(), "weights must be finite")
    error_if((w < 0.0).any(), "weights must be non-negative")
    error_if(w.sum() <= 0.0, "weights must have positive total probability")
    return Distribution(s, _normalize_probability_weights(w), mode, axis_id)


def create_trivial_distribution(sample_dim: int = 1, reduction=ReductionMode.INCOHERENT,
                                axis_id: Optional[str] = "trivial") -> Distribution:
    """One zero sample of width ``sample_dim`` with weight 1 (``distributions.py:184-209``)."""
    dim = int(sample_dim)
    if dim <= 0:
        raise ValueError("sample_dim must be positive")
    return create_distribution(torch.zeros((1, dim), dtype=torch.float64), torch.ones(1, dtype=torch.float64),
                               reduction, axis_id)


TRIVIAL_DISTRIBUTION = create_trivial_distribution()
TRIVIAL = TRIVIAL_DISTRIBUTION


def _zero_scalar() -> torch.Tensor:
    return torch.zeros((), dtype=torch.float64)


def _zero_pair() -> torch.Tensor:
    return torch.zeros(2, dtype=torch.float64)


@dataclass(frozen=True)
class AxisUpdate:
    """Per-sample change of the nominal beam: energy (eV), position (Angstrom), tilt (mrad, (x, y))."""

    energy_delta_ev: torch.Tensor = field(default_factory=_zero_scalar)
    position_delta_ang: torch.Tensor = field(default_factory=_zero_pair)
    tilt_delta_mrad: torch.Tensor = field(default_factory=_zero_pair)


def _vector2(value, name: str) -> torch.Tensor:
    v = _zero_pair() if value is None else (value.to(torch.float64) if isinstance(value, torch.Tensor)
                                            else torch.as_tensor(value, dtype=torch.float64))
    if tuple(v.shape) != (2,):
        raise ValueError(f"{name} must have shape (2,)")
    return v


def _scalar(value, name: str) -> torch.Tensor:
    v = _zero_scalar() if value is None else (value.to(torch.float64) if isinstance(value, torch.Tensor)
                                              else torch.as_tensor(value, dtype=torch.float64))
    if v.ndim != 0:
        raise ValueError(f"{name} must be a scalar")
    return v


def create_axis_update(position_delta_ang=None, energy_delta_ev=None, tilt_delta_mrad=None) -> AxisUpdate:
    """``electron_types.py:417-486``."""
    pos = _vector2(position_delta_ang, "position_delta_ang")
    energy = _scalar(energy_delta_ev, "energy_delta_ev")
    tilt = _vector2(tilt_delta_mrad, "tilt_delta_mrad")
    error_if(~torch.isfinite(pos).all(), "position_delta_ang must be finite")
    error_if(~torch.isfinite(energy), "energy_delta_ev must be finite")
    error_if(~torch.isfinite(tilt).all(), "tilt_delta_mrad must be finite")
    return AxisUpdate(energy, pos, tilt)


def combine_axis_updates(updates: Tuple[AxisUpdate, ...]) -> AxisUpdate:
    """Sum of the updates' fields (``electron_types.py:377-414``)."""
    if len(updates) == 0:
        return create_axis_update()
    return create_axis_update(position_delta_ang=torch.stack([u.position_delta_ang for u in updates]).sum(0),
                              energy_delta_ev=torch.stack([u.energy_delta_ev for u in updates]).sum(),
                              tilt_delta_mrad=torch.stack([u.tilt_delta_mrad for u in updates]).sum(0))


__all__ = ["AxisUpdate", "Distribution", "ReductionMode", "TRIVIAL", "TRIVIAL_DISTRIBUTION", "combine_axis_updates",
           "create_axis_update", "create_distribution", "create_trivial_distribution"]
