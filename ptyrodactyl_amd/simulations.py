"""Drop-in multislice simulators with the reference's Python signatures.

Same names, parameter order and defaults as
``src/ptyrodactyl/multislice/simulations.py``; torch tensors replace JAX
arrays.  The per-position multislice (the hot path) always runs in the HIP
kernels through :mod:`ptyrodactyl_amd.engine`; the small set-up helpers
(probe, propagator, phase ramps) are float64 torch code.

Precision: the multislice runs in complex64 with FP32 accumulation; the
reference runs complex128.  Patterns are returned as float32.
"""

from __future__ import annotations

import math
from typing import Optional

import torch

from dataclasses import replace as _replace

import numpy as np

from . import _native, engine, physics
from .distributions import AxisUpdate, Distribution, ReductionMode, combine_axis_updates, create_axis_update, \
    create_distribution
from .producers import AXIS_COHERENCE, AXIS_POSITION_JITTER, AXIS_PROBE_MODES, COHERENCE_DIM, POSITION_DIM, \
    PROBE_MODE_DIM
from .reduce import apply_distributions, cartesian_samples
from .types import (STEM4D, CalibratedArray, DetectorConfig, MicroscopeConfig, PotentialSlices, ProbeModes,
                    create_calibrated_array, create_probe_modes, create_stem4d)


def _device_for(*tensors) -> torch.device:
    for t in tensors:
        if isinstance(t, torch.Tensor) and t.is_cuda:
            return t.device
    if not torch.cuda.is_available():
        raise RuntimeError("the multislice kernels need a GPU; none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def transmission_func(pot_slice, voltage_kv) -> torch.Tensor:
    """T = exp(i sigma V) (``simulations.py:105-159``); complex128 for float64 input."""
    v = pot_slice if isinstance(pot_slice, torch.Tensor) else torch.as_tensor(pot_slice, dtype=torch.float64)
    sigma = physics.interaction_sigma(float(voltage_kv))
    phase = v.to(torch.float64) * sigma
    out = torch.polar(torch.ones_like(phase), phase)
    return out if v.dtype == torch.float64 else out.to(torch.complex64)


def propagation_func(imsize_y: int, imsize_x: int, thickness_ang, voltage_kv, calib_ang) -> torch.Tensor:
    """exp(-i pi lambda dz |q|^2) on the unshifted fftfreq grid (``simulations.py:162-225``)."""
    qy = physics.fftfreq(imsize_y, float(calib_ang))
    qx = physics.fftfreq(imsize_x, float(calib_ang))
    lya, lxa = torch.meshgrid(qy, qx, indexing="ij")
    lam = physics.relativistic_wavelength_ang(float(voltage_kv))
    return torch.polar(torch.ones_like(lya), -math.pi * lam * float(thickness_ang) * (lxa**2 + lya**2))


def aberration(fourier_coord, lambda_angstrom, defocus=0.0, c3=0.0, c5=0.0) -> torch.Tensor:
    """chi = 2pi/lambda (C1 p^2/2 + C3 1e7 p^4/4 + C5 1e7 p^6/6), p = lambda q (``simulations.py:434-498``)."""
    q = fourier_coord if isinstance(fourier_coord, torch.Tensor) else torch.as_tensor(fourier_coord,
                                                                                      dtype=torch.float64)
    lam = float(lambda_angstrom)
    p = lam * q.to(torch.float64)

    def coeff(v):  # tensors (which may require grad) are kept; numbers become floats
        return v.to(device=p.device, dtype=torch.float64) if isinstance(v, torch.Tensor) else float(v)

    chi = (coeff(defocus) * p**2) / 2 + (coeff(c3) * 1e7 * p**4) / 4 + (coeff(c5) * 1e7 * p**6) / 6
    return (2 * math.pi * chi) / lam


def make_probe(microscope: MicroscopeConfig, detector: DetectorConfig) -> torch.Tensor:
    """Real-space probe ifftshift(ifft2(A exp(-i chi))) (``simulations.py:334-431``).

    Hard aperture |q| <= alpha/lambda on the grid q = roll(arange(-n/2, n/2)/fov, n//2),
    fov = n * 0.01 * probe_calibration_pm.  Returns (H, W) complex128.
    """
    if microscope.probe_shape is None:
        raise ValueError("microscope.probe_shape is required")
    image_y, image_x = microscope.probe_shape
    lam = physics.relativistic_wavelength_ang(microscope.voltage_kv)
    l_max = (microscope.aperture_mrad / 1000.0) / lam
    x_fov = image_x * 0.01 * detector.probe_calibration_pm
    y_fov = image_y * 0.01 * detector.probe_calibration_pm
    qx = torch.arange(-image_x / 2, image_x / 2, 1.0, dtype=torch.float64) / x_fov
    qy = torch.arange(-image_y / 2, image_y / 2, 1.0, dtype=torch.float64) / y_fov
    lx = torch.roll(qx, image_x // 2)
    ly = torch.roll(qy, image_y // 2)
    lya, lxa = torch.meshgrid(lx, ly, indexing="xy")
    q = torch.sqrt(lxa * lxa + lya * lya)
    chi = aberration(q, lam, microscope.defocus_ang, microscope.c3_ang, microscope.c5_ang)
    a_dist = (q <= l_max).to(torch.float64)
    a_dist = torch.polar(a_dist, -chi)
    return torch.fft.ifftshift(torch.fft.ifft2(a_dist))


def shift_beam_fourier(beam, pos, calib_ang) -> torch.Tensor:
    """Shift beam modes by ``pos`` Angstrom with a Fourier phase ramp (``simulations.py:1236-1327``).

    Materialises (P, H, W, M) complex128 like the reference; the HIP forward
    never calls this (it applies the same ramp per position on the fly).
    """
    b = beam if isinstance(beam, torch.Tensor) else torch.as_tensor(beam)
    b = b.to(torch.complex128)
    if b.ndim == 2:
        b = b[..., None]
    p = torch.as_tensor(pos, dtype=torch.float64, device=b.device)
    p = torch.atleast_2d(p)
    hh, ww = b.shape[0], b.shape[1]
    qy = physics.fftfreq(hh, float(calib_ang), device=b.device)
    qx = physics.fftfreq(ww, float(calib_ang), device=b.device)
    qya, qxa = torch.meshgrid(qy, qx, indexing="ij")
    beam_k = torch.fft.fft2(b, dim=(0, 1))
    phase = -2.0 * math.pi * (qya[None] * p[:, 0, None, None] + qxa[None] * p[:, 1, None, None])
    ramp = torch.polar(torch.ones_like(phase), phase)
    return torch.fft.ifft2(beam_k[None] * ramp[..., None], dim=(1, 2))


# ---------------------------------------------------------------------------
# HIP-backed simulators
# ---------------------------------------------------------------------------

def _check_no_ensemble(microscope: MicroscopeConfig) -> None:
    if not microscope.ensemble.is_empty():
        raise ValueError("ensemble axes (position jitter / coherence) are not supported by this build")


def _slab_from_potential(pot_slices: PotentialSlices, voltage_kv: float, device) -> torch.Tensor:
    """(H, W, S) potential -> (S, H, W) complex64 transmission on ``device``."""
    v = pot_slices.slices
    if v.ndim == 2:
        v = v[..., None]
    v = v.to(device=device, dtype=torch.float64)
    sigma = physics.interaction_sigma(voltage_kv)
    t = torch.polar(torch.ones_like(v), v * sigma).to(torch.complex64)
    return t.permute(2, 0, 1).contiguous()


def _probe_k(modes: torch.Tensor, device) -> torch.Tensor:
    """(H, W, M) real-space modes -> (M, H, W) complex64 fft2 on ``device``."""
    m = modes if modes.ndim == 3 else modes[..., None]
    m = m.to(device=device, dtype=torch.complex128).permute(2, 0, 1)
    return torch.fft.fft2(m).to(torch.complex64).contiguous()


def _grid_shape(h: int, w: int):
    """(h, w) of a probe grid the kernels take: any edge from 1 to 4096 (square powers of two from 32 run on
    the dedicated kernels, every other shape on the runtime-size ones)."""
    h, w = int(h), int(w)
    if not (1 <= h <= engine.GRID_MAX_EDGE and 1 <= w <= engine.GRID_MAX_EDGE):
        raise ValueError(f"grid {h}x{w}: each edge must be in [1, {engine.GRID_MAX_EDGE}]")
    return h, w


def _run_positions(pot_slices: PotentialSlices, beam: ProbeModes, microscope: MicroscopeConfig,
                   shifts_px: torch.Tensor, want_amplitude: bool, device):
    h, w = _grid_shape(beam.modes.shape[0], beam.modes.shape[1])
    if tuple(pot_slices.slices.shape[:2]) != (h, w):
        raise ValueError("beam.modes spatial shape must match pot_slices.slices")
    calib = min(float(pot_slices.calib), float(beam.calib))
    slab = _slab_from_potential(pot_slices, microscope.voltage_kv, device)
    pk = _probe_k(beam.modes, device)
    weights = beam.weights.to(device=device, dtype=torch.float32).contiguous()
    p_count = shifts_px.shape[0]
    origins = torch.zeros((p_count, 2), dtype=torch.int32, device=device)
    prop = physics.propagator_pair(h, w, pot_slices.slice_thickness, microscope.voltage_kv, calib, device=device)
    shifts = shifts_px.to(device).contiguous()
    from . import autograd
    if autograd.differentiable(slab, pk, weights, shifts):
        # differentiable path: the potential, probe modes and positions reach the kernels through torch
        # (exp(i sigma V), fft2, positions / calibration); the HIP adjoint supplies the gradient and the HIP
        # tangent recurrence the forward-mode derivative w.r.t. the transmission, probe_k and the shifts
        if want_amplitude:
            return (None, autograd.multislice_amplitudes(slab, pk, origins, shifts, prop)), calib, h
        return (autograd.multislice_patterns(slab, pk, weights, origins, shifts, prop), None), calib, h
    return engine.multislice_forward(slab, pk, weights, origins, shifts, prop,
                                     want_patterns=True, want_amplitude=want_amplitude), calib, h


def cbed_amplitude(pot_slices: PotentialSlices, beam: ProbeModes, microscope: MicroscopeConfig) -> torch.Tensor:
    """Complex detector amplitude per mode, (H, W, M) complex64 (``simulations.py:606-679``).  Like the
    reference it uses the beam's modes as given and ignores ``microscope.ensemble``."""
    dev = _device_for(pot_slices.slices, beam.modes)
    shifts = torch.zeros((1, 2), dtype=torch.float64, device=dev)
    (_, amp), _, _ = _run_positions(pot_slices, beam, microscope, shifts, True, dev)
    return amp[0].permute(1, 2, 0)


def cbed_image(pot_slices: PotentialSlices, beam: ProbeModes, microscope: MicroscopeConfig) -> CalibratedArray:
    """CBED intensity sum_m w_m |A_m|^2 with Fourier calibration (``simulations.py:1130-1233``).

    With ensemble axes on ``microscope.ensemble`` the pattern is the reduction of every sample's amplitude
    (``bind_cbed_axes`` + ``apply_distributions``, float64 result): all samples run in one batched HIP
    launch.  When an input requires grad, the batched launch keeps the checkpoints and one batched adjoint
    returns the gradients (square power-of-two grids up to 512, incoherent axes); other grids and coherent axes
    go sample by sample through the differentiable :func:`bind_cbed_axes`."""
    dev = _device_for(pot_slices.slices, beam.modes)
    if not microscope.ensemble.is_empty():
        calib = min(float(pot_slices.calib), float(beam.calib))
        n = beam.modes.shape[0]
        needs_grad = torch.is_grad_enabled() and (pot_slices.slices.requires_grad or beam.modes.requires_grad
                                                  or _ensemble_requires_grad(microscope.ensemble))
        if needs_grad and _batched_gradient_ok(pot_slices, beam, microscope):
            pattern = _ensemble_patterns(pot_slices, beam, microscope, None, calib, dev, grad=True)[0]
        elif needs_grad:
            from .types import create_detector_config
            axes = _ensemble_axes_for_cbed(microscope.ensemble, beam)
            det = create_detector_config(real_space_calib_ang=calib, probe_calibration_pm=calib * 100.0)
            pattern = apply_distributions(axes, bind_cbed_axes(pot_slices, beam, microscope, det, axes))
        else:
            pattern = _ensemble_patterns(pot_slices, beam, microscope, None, calib, dev)[0]
        inv = 1.0 / (n * calib)
        return create_calibrated_array(pattern, inv, inv, False)
    shifts = torch.zeros((1, 2), dtype=torch.float64, device=dev)
    (patterns, _), calib, n = _run_positions(pot_slices, beam, microscope, shifts, False, dev)
    inv = 1.0 / (n * calib)
    return create_calibrated_array(patterns[0], inv, inv, False)


def stem_4d(pot_slice: PotentialSlices, beam: ProbeModes, microscope: MicroscopeConfig,
            detector: DetectorConfig) -> STEM4D:
    """4D-STEM over ``detector.scan_positions_px`` (``simulations.py:1330-1443``).

    As in the reference, the values of ``scan_positions_px`` are handed to
    the Fourier shift as Angstrom (``simulations.py:1390-1392``) and the
    stored ``scan_positions`` are those values times the calibration.
    """
    if detector.scan_positions_px is None:
        raise ValueError("detector.scan_positions_px is required")
    dev = _device_for(pot_slice.slices, beam.modes, detector.scan_positions_px)
    positions = detector.scan_positions_px.to(device=dev, dtype=torch.float64)
    if positions.shape[0] == 0:  # an empty scan maps to an empty stack, as vmap over zero positions does
        h, w = _grid_shape(beam.modes.shape[0], beam.modes.shape[1])
        calib = min(float(pot_slice.calib), float(beam.calib))
        return create_stem4d(torch.empty((0, h, w), dtype=torch.float32, device=dev), detector.real_space_calib_ang,
                             1.0 / (h * calib), positions * detector.real_space_calib_ang, microscope.voltage_kv)
    if not microscope.ensemble.is_empty():
        n = beam.modes.shape[0]
        calib = min(float(pot_slice.calib), float(beam.calib))
        needs_grad = torch.is_grad_enabled() and (pot_slice.slices.requires_grad or beam.modes.requires_grad
                                                  or positions.requires_grad
                                                  or _ensemble_requires_grad(microscope.ensemble))
        if needs_grad and _batched_gradient_ok(pot_slice, beam, microscope):
            # differentiable: every (position, sample) wave in batched launches with one batched adjoint
            patterns = _ensemble_patterns(pot_slice, beam, microscope, positions, detector.real_space_calib_ang,
                                          dev, grad=True)
        elif needs_grad:
            # differentiable: per position, the shifted modes through the ensemble cbed_image (its per-sample
            # binder), as the reference maps cbed_image over the positions (simulations.py:1395-1427)
            modes = beam.modes if beam.modes.ndim == 3 else beam.modes[..., None]
            patterns = torch.stack([
                cbed_image(pot_slice, ProbeModes(shift_beam_fourier(modes, positions[i:i + 1],
                                                                    detector.real_space_calib_ang)[0],
                                                 beam.weights, beam.calib), microscope).data_array
                for i in range(positions.shape[0])])
        else:
            patterns = _ensemble_patterns(pot_slice, beam, microscope, positions, detector.real_space_calib_ang,
                                          dev)
        return create_stem4d(patterns, detector.real_space_calib_ang, 1.0 / (n * calib),
                             positions * detector.real_space_calib_ang, microscope.voltage_kv)
    shifts = physics.shift_to_pixels(positions, detector.real_space_calib_ang)
    (patterns, _), calib, n = _run_positions(pot_slice, beam, microscope, shifts, False, dev)
    return create_stem4d(patterns, detector.real_space_calib_ang, 1.0 / (n * calib),
                         positions * detector.real_space_calib_ang, microscope.voltage_kv)


# ---------------------------------------------------------------------------
# Ensemble axes: probe modes, position jitter, partial coherence
# (simulations.py:682-945, reduce.py:92-201)
# ---------------------------------------------------------------------------

def probe_modes_to_distribution(probe: ProbeModes) -> Distribution:
    """Incoherent axis over the mode indices with the probe's weights (``simulations.py:682-711``)."""
    m = probe.modes.shape[2] if probe.modes.ndim == 3 else 1
    return create_distribution(torch.arange(m, dtype=torch.float64)[:, None], probe.weights.cpu(),
                               ReductionMode.INCOHERENT, AXIS_PROBE_MODES)


def _ensemble_requires_grad(ensemble) -> bool:
    axes = (ensemble.probe_modes, ensemble.position_jitter, ensemble.coherence)
    return any(a is not None and (a.samples.requires_grad or a.weights.requires_grad) for a in axes)


def _ensemble_axes_for_cbed(ensemble, beam: ProbeModes):
    """Axes in the reference's order: modes (explicit, or the probe's when M > 1), jitter, coherence; the probe
    modes alone when nothing else is set (``simulations.py:726-745``)."""
    axes = []
    m = beam.modes.shape[2] if beam.modes.ndim == 3 else 1
    if ensemble.probe_modes is not None:
        axes.append(ensemble.probe_modes)
    elif m > 1:
        axes.append(probe_modes_to_distribution(beam))
    if ensemble.position_jitter is not None:
        axes.append(ensemble.position_jitter)
    if ensemble.coherence is not None:
        axes.append(ensemble.coherence)
    if not axes:
        axes.append(probe_modes_to_distribution(beam))
    return tuple(axes)


_AXIS_DIMS = {AXIS_PROBE_MODES: PROBE_MODE_DIM, AXIS_POSITION_JITTER: POSITION_DIM, AXIS_COHERENCE: COHERENCE_DIM}


def _resolve_column_maps(axes, column_maps):
    if len(column_maps) == 0:
        out = []
        for axis in axes:
            if axis.axis_id is None:
                raise ValueError("distribution axis_id is required")
            out.append(axis.axis_id)
        return tuple(out)
    if len(column_maps) == len(axes):
        return tuple(column_maps)
    raise ValueError("column_maps must be empty or match axes length")


def _expected_axis_dim(axis_map: str) -> int:
    if axis_map not in _AXIS_DIMS:
        raise ValueError(f"unknown axis_id {axis_map!r}")
    return _AXIS_DIMS[axis_map]


def _validate_axis_maps(axes, axis_maps, probe_modes: ProbeModes) -> None:
    for i, (axis, axis_map) in enumerate(zip(axes, axis_maps)):
        expected = _expected_axis_dim(axis_map)
        if axis.samples.shape[1] != expected:
            raise ValueError(f"{axis_map!r} axis must have sample width {expected}")
        if axis_map == AXIS_PROBE_MODES and i != 0:
            raise ValueError("probe_modes axis must be first")
    m = probe_modes.modes.shape[2] if probe_modes.modes.ndim == 3 else 1
    if AXIS_PROBE_MODES not in axis_maps and m != 1:
        raise ValueError("multimode probes require a first probe_modes axis")


def _axis_update_from_sample(sample: torch.Tensor, axes, axis_maps):
    """Split one concatenated sample row into the combined update and the mode index
    (``simulations.py:948-990``)."""
    cursor = 0
    updates = []
    mode_idx = 0
    for axis, axis_map in zip(axes, axis_maps):
        dim = axis.samples.shape[1]
        cols = sample[cursor:cursor + dim]
        if axis_map == AXIS_PROBE_MODES:
            mode_idx = int(cols[0])
            updates.append(create_axis_update())
        elif axis_map == AXIS_POSITION_JITTER:
            updates.append(create_axis_update(position_delta_ang=cols))
        elif axis_map == AXIS_COHERENCE:
            updates.append(create_axis_update(energy_delta_ev=cols[0], tilt_delta_mrad=cols[1:3]))
        else:
            raise ValueError(f"unknown axis_id {axis_map!r}")
        cursor += dim
    return combine_axis_updates(tuple(updates)), mode_idx


def _tilt_ramp(n_y: int, n_x: int, tilt_delta_mrad, voltage_kv: float, calib_ang: float, device) -> torch.Tensor:
    """exp(i 2 pi / lambda (theta_x x + theta_y y)) on centred coordinates (i - (n-1)/2) calib, float64."""
    y = (torch.arange(n_y, dtype=torch.float64, device=device) - (n_y - 1) / 2.0) * calib_ang
    x = (torch.arange(n_x, dtype=torch.float64, device=device) - (n_x - 1) / 2.0) * calib_ang
    yy, xx = torch.meshgrid(y, x, indexing="ij")
    t = tilt_delta_mrad.to(device=device, dtype=torch.float64)
    phase = (2.0 * math.pi / physics.relativistic_wavelength_ang(voltage_kv)) * (t[0] / 1000.0 * xx
                                                                                 + t[1] / 1000.0 * yy)
    return torch.polar(torch.ones_like(phase), phase)


def _apply_tilt_phase_ramp(modes: torch.Tensor, tilt_delta_mrad, voltage_kv, calib_ang) -> torch.Tensor:
    """Beam tilt as a real-space phase ramp on every mode (``simulations.py:890-945``)."""
    ramp = _tilt_ramp(modes.shape[0], modes.shape[1], torch.as_tensor(tilt_delta_mrad, dtype=torch.float64),
                      float(voltage_kv), float(calib_ang), modes.device)
    return modes * ramp.to(modes.dtype)[..., None]


def bind_cbed_axes(pot_slices: PotentialSlices, probe_modes: ProbeModes, microscope: MicroscopeConfig,
                   detector: DetectorConfig, axes, column_maps=()):
    """Callable sample -> (H, W) complex amplitude of one ensemble member (``simulations.py:769-887``).

    Per sample: select the mode, shift it by the jitter (Angstrom, Fourier ramp at the detector
    calibration), tilt it (real-space ramp at the sample's energy), and run ``cbed_amplitude`` at the
    sample's voltage.  Differentiable w.r.t. the potential, the modes and the jitter/tilt samples; the
    energy enters as a number.  Each call is one HIP launch sequence; ``cbed_image`` and ``stem_4d``
    evaluate whole ensembles in one batched launch instead."""
    axis_maps = _resolve_column_maps(axes, column_maps)
    _validate_axis_maps(axes, axis_maps, probe_modes)
    expected_dim = sum(a.samples.shape[1] for a in axes)
    calib_ang = float(detector.real_space_calib_ang)

    def bound_amplitude_fn(sample):
        if sample.shape[0] != expected_dim:
            raise ValueError("sample width does not match bound axes")
        update, mode_idx = _axis_update_from_sample(sample, axes, axis_maps)
        kv = float(microscope.voltage_kv) + float(update.energy_delta_ev) / 1000.0
        modes = probe_modes.modes if probe_modes.modes.ndim == 3 else probe_modes.modes[..., None]
        selected = modes[..., mode_idx:mode_idx + 1]
        shifted = shift_beam_fourier(selected, update.position_delta_ang.to(selected.device)[None, :], calib_ang)[0]
        tilted = _apply_tilt_phase_ramp(shifted, update.tilt_delta_mrad, kv, calib_ang)
        probe = create_probe_modes(tilted, torch.ones(1, dtype=torch.float64), probe_modes.calib)
        return cbed_amplitude(pot_slices, probe, _replace(microscope, voltage_kv=kv))[..., 0]

    return bound_amplitude_fn


def decompose_beam_to_modes(beam: CalibratedArray, num_modes: int, key, first_mode_weight: float = 0.6) -> ProbeModes:
    """Split a beam into ``num_modes`` modes (``simulations.py:1446-1536``).

    Orthonormal random columns (reduced QR of a complex normal (H W, M) matrix drawn from ``key`` as
    ``jax.random`` would: two split keys for the real and imaginary parts, generated by the HIP Threefry
    kernel) scaled by the beam's amplitude |beam|; weights ``first_mode_weight`` then equal shares of the
    rest.  ``key`` is a raw ``uint32[2]`` key (``ptyrodactyl_amd.random.PRNGKey``) or an integer seed."""
    from . import random as _random
    hh, ww = beam.data_array.shape
    tp, m = hh * ww, int(num_modes)
    k1, k2 = _random.split(key, 2)
    dev = _device_for(beam.data_array)
    matrix = torch.complex(_random.normal(k1, (tp, m), dev), _random.normal(k2, (tp, m), dev))
    qq, _ = torch.linalg.qr(matrix, mode="reduced")
    amplitude = torch.abs(beam.data_array.reshape(-1).to(device=dev, dtype=torch.complex128))
    weights = torch.zeros(m, dtype=torch.float64)
    weights[0] = float(first_mode_weight)
    if m > 1:
        weights[1:] = (1.0 - float(first_mode_weight)) / (m - 1)
    modes = (qq * amplitude[:, None]).reshape(hh, ww, m)
    return create_probe_modes(modes, weights, beam.calib_y)


_ENSEMBLE_CHUNK_BYTES = 4 << 30
_DESC_DTYPE = np.dtype(_native.POSITION_DESC_FIELDS)


def _batched_gradient_ok(pot_slices: PotentialSlices, beam: ProbeModes, microscope: MicroscopeConfig) -> bool:
    """Whether the differentiable ensemble can run as one batched launch per chunk: a square power-of-two grid of
    the batched kernels (n <= 512) and incoherent axes only (the batched adjoint is seeded by dL/dI)."""
    h, w = _grid_shape(beam.modes.shape[0], beam.modes.shape[1])
    if not (engine.is_fast_grid(h, w) and h <= 512):
        return False
    axes = _ensemble_axes_for_cbed(microscope.ensemble, beam)
    return not any(a.reduction is ReductionMode.COHERENT for a in axes)


def _ensemble_patterns(pot_slices: PotentialSlices, beam: ProbeModes, microscope: MicroscopeConfig,
                       positions: Optional[torch.Tensor], det_calib: float, dev, grad: bool = False) -> torch.Tensor:
    """(P, h, w) float64 ensemble-reduced patterns; P = 1 without positions (``cbed_image``).

    Every (position, ensemble sample) pair becomes one wave of a batched HIP launch: its incident wave is
    built in complex128 in the reference's order (position shift at ``det_calib``, mode selection, jitter
    shift and tilt at the cbed calibration), its transmission and propagator are those of the sample's
    energy.  Incoherent axes are summed from the kernel's intensities, coherent axes from its amplitudes.
    Grids other than the square powers of two run each energy's waves as the modes of one runtime-size
    launch (amplitudes per wave).

    ``grad``: keep the potential, the modes, the positions and the axes' samples and weights in the autograd
    graph and run each chunk's waves through ``autograd.BatchedWavesFunction`` (forward with checkpoints, one
    batched adjoint); only for the grids and axes ``_batched_gradient_ok`` accepts."""
    from . import autograd
    h, w = _grid_shape(beam.modes.shape[0], beam.modes.shape[1])
    fast = engine.is_fast_grid(h, w)
    if tuple(pot_slices.slices.shape[:2]) != (h, w):
        raise ValueError("beam.modes spatial shape must match pot_slices.slices")
    calib = min(float(pot_slices.calib), float(beam.calib))
    axes = _ensemble_axes_for_cbed(microscope.ensemble, beam)
    maps = _resolve_column_maps(axes, ())
    _validate_axis_maps(axes, maps, beam)
    counts = [a.samples.shape[0] for a in axes]
    cart = cartesian_samples(axes)
    k_count = cart.shape[0]
    mode_idx, dpos, energy, tilt = [], [], [], []
    for row in cart:
        upd, m = _axis_update_from_sample(row, axes, maps)
        mode_idx.append(m)
        dpos.append(upd.position_delta_ang)
        energy.append(float(upd.energy_delta_ev))
        tilt.append(upd.tilt_delta_mrad)
    dpos = torch.stack(dpos).to(dev)
    tilt = torch.stack(tilt).to(dev)
    if not grad:
        dpos, tilt = dpos.detach(), tilt.detach()
    energies = sorted(set(energy))
    group = [energies.index(e) for e in energy]
    coherent = any(a.reduction is ReductionMode.COHERENT for a in axes)

    v = pot_slices.slices if pot_slices.slices.ndim == 3 else pot_slices.slices[..., None]
    v = v.to(device=dev, dtype=torch.float64)
    if not grad:
        v = v.detach()
    s_count = v.shape[2]
    props = torch.empty((len(energies), h + w), dtype=torch.complex64, device=dev)
    kvs = [float(microscope.voltage_kv) + e / 1000.0 for e in energies]
    slab_list = []
    for g, kv in enumerate(kvs):
        t = torch.polar(torch.ones_like(v), v * physics.interaction_sigma(kv))
        slab_list.append(t.permute(2, 0, 1).to(torch.complex64))
        fy, fx = physics.propagator_pair(h, w, pot_slices.slice_thickness, kv, calib, device=dev)
        props[g, :h] = fy
        props[g, h:] = fx
    slabs = torch.stack(slab_list).contiguous()
    qy = physics.fftfreq(h, calib, device=dev)
    qx = physics.fftfreq(w, calib, device=dev)
    jitter = None
    if bool((dpos != 0).any()):
        jitter = torch.polar(torch.ones((k_count, h, w), dtype=torch.float64, device=dev),
                             -2.0 * math.pi * (qy[None, :, None] * dpos[:, 0, None, None]
                                               + qx[None, None, :] * dpos[:, 1, None, None]))
    tilts = None
    if bool((tilt != 0).any()):
        tilts = torch.stack([_tilt_ramp(h, w, tilt[k], kvs[group[k]], calib, dev) for k in range(k_count)])
    weights = None
    if not coherent:
        grids = torch.meshgrid(*[torch.arange(c) for c in counts], indexing="ij")
        weights = torch.ones(k_count, dtype=torch.float64, device=dev)
        for a, gi in zip(axes, grids):
            aw = a.weights.to(device=dev, dtype=torch.float64)
            weights = weights * (aw if grad else aw.detach())[gi.reshape(-1).to(dev)]

    modes = beam.modes if beam.modes.ndim == 3 else beam.modes[..., None]
    modes = modes.to(device=dev, dtype=torch.complex128)
    if not grad:
        modes = modes.detach()
    p_count = 1 if positions is None else positions.shape[0]
    per_wave = h * w * (80 + (8 * ((s_count - 1) // autograd.DEFAULT_CKPT_EVERY) + 24 if grad else 0))
    chunk = max(1, min(p_count, _ENSEMBLE_CHUNK_BYTES // (k_count * per_wave)))
    out = torch.empty((p_count, h, w), dtype=torch.float64, device=dev)
    pieces = []
    base = engine.arena_base(dev)
    sel_idx = torch.as_tensor(mode_idx, device=dev)
    # The position shift (shift_beam_fourier at det_calib) and the jitter shift are both Fourier phase ramps, so
    # every incident wave is formed in Fourier space as fft2(mode) * ramp_p * jitter_k; only a tilt (a real-space
    # ramp) needs the inverse and forward transforms.
    modes_k = torch.fft.fft2(modes, dim=(0, 1))  # (h, w, M)
    qdy = physics.fftfreq(h, float(det_calib), device=dev)
    qdx = physics.fftfreq(w, float(det_calib), device=dev)
    for p0 in range(0, p_count, chunk):
        pc = min(chunk, p_count - p0)
        if positions is None:
            kp = modes_k[None]
        else:
            pp = positions[p0:p0 + pc].to(device=dev, dtype=torch.float64)
            if not grad:
                pp = pp.detach()
            phase = -2.0 * math.pi * (qdy[None, :, None] * pp[:, 0, None, None]
                                      + qdx[None, None, :] * pp[:, 1, None, None])
            kp = modes_k[None] * torch.polar(torch.ones_like(phase), phase)[..., None]
        waves_k = kp.index_select(3, sel_idx).permute(0, 3, 1, 2)  # (pc, K, h, w), Fourier space
        if jitter is not None:
            waves_k = waves_k * jitter[None]
        if tilts is not None:
            waves_k = torch.fft.fft2(torch.fft.ifft2(waves_k) * tilts[None])
        probe_k = waves_k.to(torch.complex64).reshape(pc * k_count, h, w).contiguous()
        del waves_k, kp
        wave_group = np.tile(np.asarray(group, dtype=np.int64), pc)
        if grad:
            patterns = autograd.batched_wave_patterns(slabs, probe_k, props, wave_group.tolist())
            pieces.append(torch.einsum("k,pkhw->phw", weights, patterns.to(torch.float64).reshape(pc, k_count, h, w)))
            continue
        if fast:
            n = h
            patterns = torch.empty((pc * k_count, n, n), dtype=torch.float32, device=dev)
            amps = torch.empty((pc * k_count, n, n), dtype=torch.complex64, device=dev) if coherent else None
            rec = np.zeros(pc * k_count, dtype=_DESC_DTYPE)
            rec["slab_offset"] = engine.element_offset(slabs, base) + wave_group * (s_count * n * n)
            rec["slab_h"] = n
            rec["slab_w"] = n
            rec["pattern_offset"] = engine.element_offset(patterns, base) + np.arange(pc * k_count) * n * n
            rec["probe_offset"] = engine.element_offset(probe_k, base) + np.arange(pc * k_count) * n * n
            rec["prop_offset"] = engine.element_offset(props, base) + wave_group * (2 * n)
            descs = torch.from_numpy(rec.view(np.uint8)).to(dev)
            engine.forward_batched(n, s_count, descs, base, amplitudes=amps)
        else:
            # each energy: its waves as the probe modes of one position at the origin, amplitudes per wave
            amps = torch.empty((pc * k_count, h, w), dtype=torch.complex64, device=dev)
            zero_o = torch.zeros((1, 2), dtype=torch.int32, device=dev)
            zero_s = torch.zeros((1, 2), dtype=torch.float64, device=dev)
            for g in range(len(energies)):
                sel = torch.as_tensor(np.nonzero(wave_group == g)[0], device=dev)
                pk_g = probe_k.index_select(0, sel).contiguous()
                ones = torch.ones(pk_g.shape[0], dtype=torch.float32, device=dev)
                _, amp_g = engine.multislice_forward(slabs[g], pk_g, ones, zero_o, zero_s,
                                                     (props[g, :h].contiguous(), props[g, h:].contiguous()),
                                                     want_patterns=False, want_amplitude=True)
                amps.index_copy_(0, sel, amp_g[0])
            patterns = amps.abs().square() if not coherent else None
        if coherent:
            a = amps.to(torch.complex128).reshape(pc, *counts, h, w)
            out[p0:p0 + pc] = torch.stack([_reduce_one(axes, a[i]) for i in range(pc)])
        else:
            out[p0:p0 + pc] = torch.einsum("k,pkhw->phw", weights,
                                           patterns.to(torch.float64).reshape(pc, k_count, h, w))
    return torch.cat(pieces) if grad else out


def _reduce_one(axes, amplitudes: torch.Tensor) -> torch.Tensor:
    from .reduce import reduce_amplitudes
    return reduce_amplitudes(axes, amplitudes)


def stem_4d_windows(slab_t: torch.Tensor, probe: torch.Tensor, origins: torch.Tensor, shifts_px: torch.Tensor,
                    slice_thickness: float, voltage_kv: float, calib_ang: float,
                    weights: Optional[torch.Tensor] = None) -> torch.Tensor:
    """4D-STEM with per-position transmission windows cropped from a larger slab.

    The pattern at position p equals ``cbed_image`` on the periodic window of
    ``slab_t`` (S, Hs, Ws complex64) at integer ``origins[p]`` with the probe
    (h, w) or (M, h, w) shifted by ``shifts_px[p]`` pixels.  Returns
    (P, h, w) float32 fftshifted patterns.
    """
    dev = slab_t.device
    pr = probe if probe.ndim == 3 else probe[None]
    pk = torch.fft.fft2(pr.to(device=dev, dtype=torch.complex128)).to(torch.complex64).contiguous()
    m_count = pk.shape[0]
    w = torch.full((m_count,), 1.0 / m_count, dtype=torch.float32, device=dev) if weights is None else \
        (weights / weights.sum()).to(device=dev, dtype=torch.float32).contiguous()
    prop = physics.propagator_pair(pk.shape[-2], pk.shape[-1], slice_thickness, voltage_kv, calib_ang, device=dev)
    patterns, _ = engine.multislice_forward(slab_t.contiguous(), pk, w, origins.to(dev, torch.int32).contiguous(),
                                            shifts_px.to(dev, torch.float64).contiguous(), prop)
    return patterns


def annular_mask(shape, fourier_space_calib: float, voltage_kv: float, inner_mrad: float, outer_mrad: float,
                 device=None) -> torch.Tensor:
    """0/1 (uint8) mask of the fftshifted pattern pixels inside the annulus, computed in float64 with the
    reference's operation order (``simulations.py:1596-1618``): radius = sqrt(qy^2 + qx^2) * calib with q
    measured from (H//2, W//2), inner_mrad / 1000 / lambda <= radius <= outer_mrad / 1000 / lambda."""
    hh, ww = int(shape[0]), int(shape[1])
    lam = physics.relativistic_wavelength_ang(voltage_kv)
    inner_k = (float(inner_mrad) / 1000.0) / lam
    outer_k = (float(outer_mrad) / 1000.0) / lam
    qy = torch.arange(hh, dtype=torch.float64, device=device) - hh // 2
    qx = torch.arange(ww, dtype=torch.float64, device=device) - ww // 2
    qya, qxa = torch.meshgrid(qy, qx, indexing="ij")
    radius = torch.sqrt(qya**2 + qxa**2) * float(fourier_space_calib)
    return ((radius >= inner_k) & (radius <= outer_k)).to(torch.uint8).contiguous()


def annular_detector(stem4d_data: STEM4D, detector: DetectorConfig) -> CalibratedArray:
    """Integrate 4D-STEM data with a virtual annular detector (``simulations.py:1540-1651``).

    Every pattern is summed over the pixels between ``detector.collection_inner_mrad`` and
    ``collection_outer_mrad`` (inclusive, converted with the wavelength at ``stem4d_data.voltage_kv``)
    by the HIP reduction ``pty_annular_integrate`` (FP64 accumulation, fixed order), and the sums are
    reshaped to ``detector.scan_shape``.  Returns a real-space ``CalibratedArray`` (float64) with the
    4D-STEM's real-space calibration on both axes.  Raises ``ValueError`` without a scan shape.
    """
    if detector.scan_shape is None:
        raise ValueError("detector.scan_shape is required")
    from . import _native
    data = stem4d_data.data
    if not (isinstance(data, torch.Tensor) and data.is_cuda):
        raise ValueError("stem4d_data.data must be a GPU tensor (the detector integration has no CPU fallback)")
    p_count, hh, ww = data.shape
    ny, nx = (int(v) for v in detector.scan_shape)
    if ny * nx != p_count:
        raise ValueError(f"scan_shape {detector.scan_shape} does not hold {p_count} patterns")
    dev = data.device
    mask = annular_mask((hh, ww), stem4d_data.fourier_space_calib, stem4d_data.voltage_kv,
                        detector.collection_inner_mrad, detector.collection_outer_mrad, device=dev)
    pats = data.to(torch.float32).contiguous()
    out = torch.empty(p_count, dtype=torch.float64, device=dev)
    lib = _native.load()
    with torch.cuda.device(dev):
        status = lib.pty_annular_integrate(_native.ptr(pats), p_count, hh * ww, _native.ptr(mask), _native.ptr(out),
                                           _native.stream_handle(dev))
    _native.check(status, "pty_annular_integrate")
    return create_calibrated_array(out.reshape(ny, nx), stem4d_data.real_space_calib, stem4d_data.real_space_calib,
                                   True)


def stem_4d_windows_detectors(slab_t: torch.Tensor, probe: torch.Tensor, origins: torch.Tensor,
                              shifts_px: torch.Tensor, slice_thickness: float, voltage_kv: float, calib_ang: float,
                              collection_mrad, weights: Optional[torch.Tensor] = None, keep_patterns: bool = False):
    """Windowed 4D-STEM (as :func:`stem_4d_windows`) with annular virtual detectors fused into the kernel
    epilogue.  collection_mrad: up to four (inner, outer) angle pairs; detector d integrates every
    pattern over the annulus as :func:`annular_detector` does (Fourier calibration 1 / (n calib_ang)).
    Returns (signals (P, D) float64, patterns (P, n, n) float32 or None); without ``keep_patterns``
    the pattern stack is never written."""
    dev = slab_t.device
    pr = probe if probe.ndim == 3 else probe[None]
    pk = torch.fft.fft2(pr.to(device=dev, dtype=torch.complex128)).to(torch.complex64).contiguous()
    m_count, n = pk.shape[0], pk.shape[-1]
    w = torch.full((m_count,), 1.0 / m_count, dtype=torch.float32, device=dev) if weights is None else \
        (weights / weights.sum()).to(device=dev, dtype=torch.float32).contiguous()
    prop = physics.propagator_factors(n, slice_thickness, voltage_kv, calib_ang, device=dev)
    masks = torch.stack([annular_mask((n, n), 1.0 / (n * calib_ang), voltage_kv, lo, hi, device=dev)
                         for lo, hi in collection_mrad]).contiguous()
    patterns, signals = engine.multislice_forward_detectors(
        slab_t.contiguous(), pk, w, origins.to(dev, torch.int32).contiguous(),
        shifts_px.to(dev, torch.float64).contiguous(), prop, masks, want_patterns=keep_patterns)
    return signals, patterns
