"""Atom-based slice provider and ``stem4d_sharded`` with the reference's signatures.

Follows ``src/ptyrodactyl/multislice/parallelized.py``: slices are built from
atoms (per-type delta grids FFT-convolved with precomputed per-type kernels,
``_compute_slice_potential`` ``:54-169``) and fed to the same multislice as
``stem_4d`` with positions in Angstrom (``stem4d_sharded`` ``:348-536``).  The
reference rebuilds every slice for every position; here each slice is built
once on the device (float64 torch set-up) and the positions run through the
HIP forward kernels.  ``mesh`` is accepted for signature compatibility: this
build runs on one GPU.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import torch

from . import engine, physics
from .simulations import _device_for, _grid_shape, _probe_k
from .types import (STEM4D, DetectorConfig, MicroscopeConfig, ProbeModes, _as_tensor, create_stem4d, error_if)


@dataclass(frozen=True)
class AtomicSliceData:
    """``types/electron_types.py:192-225``: coordinates (x, y, z) in Angstrom."""

    atom_coords: torch.Tensor
    atom_types: torch.Tensor
    slice_z_bounds: torch.Tensor
    atom_potentials: torch.Tensor
    atom_mask: Optional[torch.Tensor] = None


def create_atomic_slice_data(atom_coords, atom_types, slice_z_bounds, atom_potentials,
                             atom_mask=None) -> AtomicSliceData:
    """``types/electron_types.py:735-851`` (same structural and value checks)."""
    coords = _as_tensor(atom_coords, torch.float64)
    types = _as_tensor(atom_types, torch.int64)
    bounds = _as_tensor(slice_z_bounds, torch.float64)
    kern = _as_tensor(atom_potentials, torch.float64)
    mask = None if atom_mask is None else _as_tensor(atom_mask, torch.float64)
    if coords.ndim != 2 or coords.shape[1] != 3:
        raise ValueError("atom_coords must have shape (N, 3)")
    if types.shape != (coords.shape[0],):
        raise ValueError("atom_types must have shape (N,)")
    if bounds.ndim != 2 or bounds.shape[1] != 2:
        raise ValueError("slice_z_bounds must have shape (S, 2)")
    if kern.ndim != 3:
        raise ValueError("atom_potentials must have shape (T, H, W)")
    if mask is not None and mask.shape != (coords.shape[0],):
        raise ValueError("atom_mask must have shape (N,)")
    error_if(~torch.isfinite(coords).all(), "atom_coords contain non-finite values")
    error_if((types < 0).any(), "atom_types must be non-negative")
    error_if((types >= kern.shape[0]).any(), "atom_types must index atom_potentials")
    error_if(~torch.isfinite(bounds).all(), "slice_z_bounds contain non-finite values")
    error_if((bounds[:, 1] - bounds[:, 0] <= 0).any(), "slice_z_bounds thicknesses must be positive")
    error_if(~torch.isfinite(kern).all(), "atom_potentials contain non-finite values")
    if mask is not None:
        error_if(~torch.isfinite(mask).all(), "atom_mask contains non-finite values")
    return AtomicSliceData(coords, types, bounds, kern, mask)


def slice_potentials(sample: AtomicSliceData, grid_shape, calib_ang: float, device) -> torch.Tensor:
    """(S, H, W) float64 projected potentials (``_compute_slice_potential`` for every slice)."""
    h, w = grid_shape
    coords = sample.atom_coords.to(device)
    types = sample.atom_types.to(device)
    kern_k = torch.fft.fft2(sample.atom_potentials.to(device))
    x_idx = torch.remainder(torch.floor(coords[:, 0] / calib_ang).to(torch.int64), w)
    y_idx = torch.remainder(torch.floor(coords[:, 1] / calib_ang).to(torch.int64), h)
    flat = y_idx * w + x_idx
    mask = sample.atom_mask.to(device) if sample.atom_mask is not None else torch.ones_like(coords[:, 0])
    out = []
    for b in sample.slice_z_bounds.tolist():
        in_slice = ((coords[:, 2] >= b[0]) & (coords[:, 2] < b[1])).to(torch.float64) * mask
        total = torch.zeros((h, w), dtype=torch.float64, device=device)
        for t in range(kern_k.shape[0]):
            grid = torch.zeros(h * w, dtype=torch.float64, device=device)
            grid.index_add_(0, flat, (types == t).to(torch.float64) * in_slice)
            total = total + torch.real(torch.fft.ifft2(torch.fft.fft2(grid.reshape(h, w)) * kern_k[t]))
        out.append(total)
    return torch.stack(out)


def _transmission_slab(pot: torch.Tensor, voltage_kv: float) -> torch.Tensor:
    sigma = physics.interaction_sigma(voltage_kv)
    return torch.polar(torch.ones_like(pot), pot * sigma).to(torch.complex64).contiguous()


def cbed_amplitude_from_atoms(beam: torch.Tensor, atom_coords, atom_types, slice_z_bounds, atom_potentials,
                              voltage_kv, calib_ang, atom_mask=None) -> torch.Tensor:
    """Per-mode exit amplitudes (H, W, M) complex64 of an unshifted beam through slices built from atoms
    (``parallelized.py:172-259``): the slice thickness is that of the first slice's z bounds, the propagator and
    the atom grid use ``calib_ang``."""
    sample = create_atomic_slice_data(atom_coords, atom_types, slice_z_bounds, atom_potentials, atom_mask)
    modes = beam if beam.ndim == 3 else beam[..., None]
    dev = _device_for(modes, sample.atom_coords)
    h, w = _grid_shape(modes.shape[0], modes.shape[1])
    calib = float(calib_ang)
    slab = _transmission_slab(slice_potentials(sample, (h, w), calib, dev), float(voltage_kv))
    dz = float(sample.slice_z_bounds[0, 1] - sample.slice_z_bounds[0, 0])
    prop = physics.propagator_pair(h, w, dz, float(voltage_kv), calib, device=dev)
    pk = _probe_k(modes, dev)
    ones = torch.ones(pk.shape[0], dtype=torch.float32, device=dev)
    zeros = torch.zeros((1, 2), dtype=torch.int32, device=dev)
    _, amp = engine.multislice_forward(slab, pk, ones, zeros, torch.zeros((1, 2), dtype=torch.float64, device=dev),
                                       prop, want_patterns=False, want_amplitude=True)
    return amp[0].permute(1, 2, 0)


def cbed_image_from_atoms(beam: torch.Tensor, mode_distribution, atom_coords, atom_types, slice_z_bounds,
                          atom_potentials, voltage_kv, calib_ang, atom_mask=None) -> torch.Tensor:
    """CBED intensity (H, W) of :func:`cbed_amplitude_from_atoms` reduced over ``mode_distribution``
    (sample = mode index; ``parallelized.py:262-345``), float64."""
    from .reduce import apply_distribution
    amps = cbed_amplitude_from_atoms(beam, atom_coords, atom_types, slice_z_bounds, atom_potentials, voltage_kv,
                                     calib_ang, atom_mask).to(torch.complex128)
    return apply_distribution(mode_distribution, lambda sample: amps[..., int(sample[0])])


def stem4d_sharded(probe_modes: ProbeModes, sample: AtomicSliceData, microscope: MicroscopeConfig,
                   detector: DetectorConfig, mesh=None) -> STEM4D:
    """4D-STEM over ``detector.scan_positions_ang`` with slices built from atoms (``parallelized.py:348-536``)."""
    if detector.scan_positions_ang is None:
        raise ValueError("detector.scan_positions_ang is required")
    dev = _device_for(probe_modes.modes, sample.atom_coords)
    h, w = _grid_shape(probe_modes.modes.shape[0], probe_modes.modes.shape[1])
    calib = float(detector.real_space_calib_ang)
    if detector.scan_positions_ang.shape[0] == 0:  # empty scan: empty stack
        pos0 = detector.scan_positions_ang.to(device=dev, dtype=torch.float64)
        return create_stem4d(torch.empty((0, h, w), dtype=torch.float32, device=dev), calib, 1.0 / (h * calib), pos0,
                             microscope.voltage_kv)
    pot = slice_potentials(sample, (h, w), calib, dev)
    slab = _transmission_slab(pot, microscope.voltage_kv)
    dz = float(sample.slice_z_bounds[0, 1] - sample.slice_z_bounds[0, 0])
    pos = detector.scan_positions_ang.to(device=dev, dtype=torch.float64)
    shifts = physics.shift_to_pixels(pos, calib)
    origins = torch.zeros((pos.shape[0], 2), dtype=torch.int32, device=dev)
    prop = physics.propagator_pair(h, w, dz, microscope.voltage_kv, calib, device=dev)
    weights = probe_modes.weights.to(device=dev, dtype=torch.float32).contiguous()
    pk = _probe_k(probe_modes.modes, dev)
    from . import autograd
    if autograd.differentiable(slab, pk, weights, shifts):
        # differentiable w.r.t. the modes, their weights, the atom kernels and the positions (HIP adjoint)
        patterns = autograd.multislice_patterns(slab, pk, weights, origins, shifts, prop)
    else:
        patterns, _ = engine.multislice_forward(slab, pk, weights, origins, shifts, prop)
    return create_stem4d(patterns, calib, 1.0 / (h * calib), pos, microscope.voltage_kv)
