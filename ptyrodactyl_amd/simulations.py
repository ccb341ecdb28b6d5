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

from . import engine, physics
from .types import (STEM4D, CalibratedArray, DetectorConfig, MicroscopeConfig, PotentialSlices, ProbeModes,
                    create_calibrated_array, create_stem4d)


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
    chi = (float(defocus) * p**2) / 2 + (float(c3) * 1e7 * p**4) / 4 + (float(c5) * 1e7 * p**6) / 6
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


def _square_n(h: int, w: int) -> int:
    if h != w:
        raise ValueError(f"the multislice kernels need a square grid, got {h}x{w}")
    if h not in engine.SUPPORTED_N:
        raise ValueError(f"grid size {h} is not one of {engine.SUPPORTED_N}")
    return h


def _run_positions(pot_slices: PotentialSlices, beam: ProbeModes, microscope: MicroscopeConfig,
                   shifts_px: torch.Tensor, want_amplitude: bool, device):
    n = _square_n(beam.modes.shape[0], beam.modes.shape[1])
    if tuple(pot_slices.slices.shape[:2]) != (n, n):
        raise ValueError("beam.modes spatial shape must match pot_slices.slices")
    calib = min(float(pot_slices.calib), float(beam.calib))
    slab = _slab_from_potential(pot_slices, microscope.voltage_kv, device)
    pk = _probe_k(beam.modes, device)
    weights = beam.weights.to(device=device, dtype=torch.float32).contiguous()
    p_count = shifts_px.shape[0]
    origins = torch.zeros((p_count, 2), dtype=torch.int32, device=device)
    prop = physics.propagator_factors(n, pot_slices.slice_thickness, microscope.voltage_kv, calib, device=device)
    shifts = shifts_px.to(device).contiguous()
    if torch.is_grad_enabled() and (slab.requires_grad or pk.requires_grad or weights.requires_grad
                                    or shifts.requires_grad):
        # differentiable path: the potential, probe modes and positions reach the kernels through torch
        # (exp(i sigma V), fft2, positions / calibration) and the HIP adjoint supplies the gradient w.r.t.
        # the transmission, probe_k and the shifts in pixels
        from . import autograd
        if want_amplitude:
            return (None, autograd.multislice_amplitudes(slab, pk, origins, shifts, prop)), calib, n
        return (autograd.multislice_patterns(slab, pk, weights, origins, shifts, prop), None), calib, n
    return engine.multislice_forward(slab, pk, weights, origins, shifts, prop,
                                     want_patterns=True, want_amplitude=want_amplitude), calib, n


def cbed_amplitude(pot_slices: PotentialSlices, beam: ProbeModes, microscope: MicroscopeConfig) -> torch.Tensor:
    """Complex detector amplitude per mode, (H, W, M) complex64 (``simulations.py:606-679``)."""
    _check_no_ensemble(microscope)
    dev = _device_for(pot_slices.slices, beam.modes)
    shifts = torch.zeros((1, 2), dtype=torch.float64, device=dev)
    (_, amp), _, _ = _run_positions(pot_slices, beam, microscope, shifts, True, dev)
    return amp[0].permute(1, 2, 0)


def cbed_image(pot_slices: PotentialSlices, beam: ProbeModes, microscope: MicroscopeConfig) -> CalibratedArray:
    """CBED intensity sum_m w_m |A_m|^2 with Fourier calibration (``simulations.py:1130-1233``)."""
    _check_no_ensemble(microscope)
    dev = _device_for(pot_slices.slices, beam.modes)
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
    _check_no_ensemble(microscope)
    dev = _device_for(pot_slice.slices, beam.modes, detector.scan_positions_px)
    positions = detector.scan_positions_px.to(device=dev, dtype=torch.float64)
    shifts = physics.shift_to_pixels(positions, detector.real_space_calib_ang)
    (patterns, _), calib, n = _run_positions(pot_slice, beam, microscope, shifts, False, dev)
    return create_stem4d(patterns, detector.real_space_calib_ang, 1.0 / (n * calib),
                         positions * detector.real_space_calib_ang, microscope.voltage_kv)


def stem_4d_windows(slab_t: torch.Tensor, probe: torch.Tensor, origins: torch.Tensor, shifts_px: torch.Tensor,
                    slice_thickness: float, voltage_kv: float, calib_ang: float,
                    weights: Optional[torch.Tensor] = None) -> torch.Tensor:
    """4D-STEM with per-position transmission windows cropped from a larger slab.

    The pattern at position p equals ``cbed_image`` on the periodic window of
    ``slab_t`` (S, Hs, Ws complex64) at integer ``origins[p]`` with the probe
    (n, n) or (M, n, n) shifted by ``shifts_px[p]`` pixels.  Returns
    (P, n, n) float32 fftshifted patterns.
    """
    dev = slab_t.device
    pr = probe if probe.ndim == 3 else probe[None]
    pk = torch.fft.fft2(pr.to(device=dev, dtype=torch.complex128)).to(torch.complex64).contiguous()
    m_count, n = pk.shape[0], pk.shape[-1]
    w = torch.full((m_count,), 1.0 / m_count, dtype=torch.float32, device=dev) if weights is None else \
        (weights / weights.sum()).to(device=dev, dtype=torch.float32).contiguous()
    prop = physics.propagator_factors(n, slice_thickness, voltage_kv, calib_ang, device=dev)
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
