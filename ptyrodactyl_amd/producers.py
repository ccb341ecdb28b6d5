"""Gauss-Hermite ensemble distributions for position jitter and partial
coherence (``src/ptyrodactyl/multislice/producers.py:46-301``).

Both use the normal-density Gauss-Hermite rule: nodes sqrt(2) x_i and
weights w_i / sqrt(pi) of ``numpy.polynomial.hermite.hermgauss(n_quad)`` (the
routine the reference calls), scaled by the width; a zero width collapses
every node to 0.
"""

from __future__ import annotations

import numpy as np
import torch

from .distributions import Distribution, ReductionMode, create_distribution
from .types import error_if

AXIS_PROBE_MODES = "probe_modes"
AXIS_POSITION_JITTER = "position_jitter"
AXIS_COHERENCE = "coherence"
PROBE_MODE_DIM = 1
POSITION_DIM = 2
COHERENCE_DIM = 3


def _normal_gauss_hermite_rule(n_quad):
    q = int(n_quad)
    if q <= 0:
        raise ValueError("n_quad must be positive")
    x, w = np.polynomial.hermite.hermgauss(q)
    return (torch.as_tensor(np.sqrt(2.0) * x, dtype=torch.float64),
            torch.as_tensor(w / np.sqrt(np.pi), dtype=torch.float64))


def _collapse_width(width, nodes: torch.Tensor, name: str) -> torch.Tensor:
    w = width.to(torch.float64) if isinstance(width, torch.Tensor) else torch.as_tensor(width, dtype=torch.float64)
    if w.ndim != 0:
        raise ValueError(f"{name} must be a scalar")
    error_if(~torch.isfinite(w), f"{name} must be finite")
    error_if(w < 0.0, f"{name} must be non-negative")
    return torch.where(w > 0.0, w * nodes, torch.zeros_like(nodes))


def coherence_to_distribution(energy_spread_ev, angular_divergence_mrad, n_quad) -> Distribution:
    """Q^3 samples (energy eV, tilt_x mrad, tilt_y mrad) on the product rule (``producers.py:46-128``)."""
    nodes, weights = _normal_gauss_hermite_rule(n_quad)
    e = _collapse_width(energy_spread_ev, nodes, "energy_spread_ev")
    t = _collapse_width(angular_divergence_mrad, nodes, "angular_divergence_mrad")
    eg, txg, tyg = torch.meshgrid(e, t, t, indexing="ij")
    we, wx, wy = torch.meshgrid(weights, weights, weights, indexing="ij")
    samples = torch.stack((eg.reshape(-1), txg.reshape(-1), tyg.reshape(-1)), dim=1)
    return create_distribution(samples, (we * wx * wy).reshape(-1), ReductionMode.INCOHERENT, AXIS_COHERENCE)


def position_jitter_to_distribution(sigma_ang, n_quad) -> Distribution:
    """Q^2 samples (dy, dx) in Angstrom on the product rule (``producers.py:131-200``)."""
    nodes, weights = _normal_gauss_hermite_rule(n_quad)
    d = _collapse_width(sigma_ang, nodes, "sigma_ang")
    dy, dx = torch.meshgrid(d, d, indexing="ij")
    wy, wx = torch.meshgrid(weights, weights, indexing="ij")
    samples = torch.stack((dy.reshape(-1), dx.reshape(-1)), dim=1)
    return create_distribution(samples, (wy * wx).reshape(-1), ReductionMode.INCOHERENT, AXIS_POSITION_JITTER)


__all__ = ["coherence_to_distribution", "position_jitter_to_distribution"]
