"""Late reduction of ensemble amplitudes (``src/ptyrodactyl/multislice/reduce.py:29-201``).

``bound_amplitude_fn`` maps one sample row to an (H, W) complex amplitude.
Coherent axes are summed in amplitude (complex weights) before |.|^2,
incoherent axes in intensity afterwards; several axes form a Cartesian
product in tuple order (``indexing="ij"``) with each sample's columns
concatenated in that order.  These host reductions serve arbitrary callables;
the multislice simulators evaluate all samples of an ensemble in one batched
HIP launch and reduce the same way.
"""

from __future__ import annotations

from typing import Callable, Sequence, Tuple

import torch

from .distributions import Distribution, ReductionMode


def apply_distribution(distribution: Distribution, bound_amplitude_fn: Callable) -> torch.Tensor:
    """One axis: |sum w A|^2 (coherent) or sum w |A|^2 (incoherent)."""
    amps = torch.stack([bound_amplitude_fn(s) for s in distribution.samples])
    w = distribution.weights.to(amps.device)
    if distribution.reduction is ReductionMode.COHERENT:
        field = torch.einsum("n,nhw->hw", w.to(amps.dtype), amps)
        return torch.abs(field) ** 2
    if distribution.reduction is ReductionMode.INCOHERENT:
        return torch.einsum("n,nhw->hw", w.to(amps.real.dtype), torch.abs(amps) ** 2)
    raise ValueError("unknown distribution reduction mode")


def cartesian_samples(distributions: Sequence[Distribution]) -> torch.Tensor:
    """(prod N_i, sum D_i) sample rows of the axes' Cartesian product, first axis slowest."""
    counts = [d.samples.shape[0] for d in distributions]
    grids = torch.meshgrid(*[torch.arange(c) for c in counts], indexing="ij")
    return torch.cat([d.samples[g.reshape(-1)] for d, g in zip(distributions, grids)], dim=1)


def reduce_amplitudes(distributions: Sequence[Distribution], amplitudes: torch.Tensor) -> torch.Tensor:
    """Reduce amplitudes shaped (N_1, ..., N_k, H, W): coherent axes first (last axis first), then |.|^2, then the
    incoherent axes (last axis first)."""
    labels = list(range(len(distributions)))
    field = amplitudes
    for axis in sorted((i for i, d in enumerate(distributions) if d.reduction is ReductionMode.COHERENT),
                       reverse=True):
        cur = labels.index(axis)
        field = _weighted_sum_axis(field, distributions[axis].weights.to(field.device, field.dtype), cur)
        labels.pop(cur)
    intensity = torch.abs(field) ** 2
    for axis in sorted((i for i, d in enumerate(distributions) if d.reduction is ReductionMode.INCOHERENT),
                       reverse=True):
        cur = labels.index(axis)
        intensity = _weighted_sum_axis(intensity, distributions[axis].weights.to(intensity.device,
                                                                                 intensity.dtype), cur)
        labels.pop(cur)
    return intensity


def apply_distributions(distributions: Tuple[Distribution, ...], bound_amplitude_fn: Callable) -> torch.Tensor:
    """Several axes: evaluate every Cartesian sample, then :func:`reduce_amplitudes`."""
    if len(distributions) == 0:
        raise ValueError("distributions must contain at least one axis")
    counts = [d.samples.shape[0] for d in distributions]
    flat = torch.stack([bound_amplitude_fn(s) for s in cartesian_samples(distributions)])
    return reduce_amplitudes(distributions, flat.reshape(*counts, *flat.shape[-2:]))


def _weighted_sum_axis(values: torch.Tensor, weights: torch.Tensor, axis: int) -> torch.Tensor:
    shape = (1,) * axis + (weights.shape[0],) + (1,) * (values.ndim - axis - 1)
    return torch.sum(values * weights.reshape(shape), dim=axis)


__all__ = ["apply_distribution", "apply_distributions"]
