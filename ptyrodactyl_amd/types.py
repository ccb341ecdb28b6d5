"""Carriers and validated factories for the multislice path.

Mirrors the reference's Equinox carriers and ``create_*`` factories
(``src/ptyrodactyl/types/electron_types.py:103-373`` for the classes,
``:529-1504`` for the factories) with torch tensors.  Two validation tiers
as in the reference:

* structural problems (ranks, shapes, missing fields) raise ``ValueError``
  eagerly (e.g. ``electron_types.py:1283-1291``);
* value problems (non-finite data, non-positive scalars) raise
  :class:`RuntimeCheckError`, the counterpart of ``equinox.EquinoxRuntimeError``
  raised by ``eqx.error_if`` in the reference, with the same message text;
  setting ``PTYRODACTYL_DISABLE_RUNTIME_CHECKS=1`` turns this tier off
  (``src/ptyrodactyl/__init__.py:89-90``).
"""

from __future__ import annotations

import contextlib
import math
import os
import threading
from enum import Enum
from dataclasses import dataclass, field
from typing import Optional, Tuple

import torch


class RuntimeCheckError(RuntimeError):
    """A deferred value check failed (the reference raises EquinoxRuntimeError)."""


_CHECK_STATE = threading.local()


def runtime_checks_enabled() -> bool:
    if getattr(_CHECK_STATE, "suspended", 0):
        return False
    return os.environ.get("PTYRODACTYL_DISABLE_RUNTIME_CHECKS", "0").lower() not in ("1", "true", "yes")


@contextlib.contextmanager
def suspended_checks():
    """Tier-2 value checks off in this thread (for code recorded into a graph, which may not synchronise)."""
    _CHECK_STATE.suspended = getattr(_CHECK_STATE, "suspended", 0) + 1
    try:
        yield
    finally:
        _CHECK_STATE.suspended -= 1


def error_if(condition, message: str) -> None:
    """Raise RuntimeCheckError(message) when ``condition`` holds (tier 2)."""
    if not runtime_checks_enabled():
        return
    cond = bool(condition.item()) if isinstance(condition, torch.Tensor) else bool(condition)
    if cond:
        raise RuntimeCheckError(message)


class _ValueChecks:
    """Tier-2 checks of one factory, in order.  Conditions are queued as callables, so nothing is computed when the
    checks are disabled; when enabled, the device conditions are resolved with one synchronisation per device and
    the first failing check raises with its message."""

    def __init__(self):
        self.on = runtime_checks_enabled()
        self.items = []

    def add(self, condition, message: str) -> "_ValueChecks":
        if self.on:
            self.items.append((condition, message))
        return self

    def finite(self, value: float, name: str) -> "_ValueChecks":
        return self.add(lambda: not math.isfinite(float(value)), f"{name} must be finite")

    def positive(self, value: float, name: str) -> "_ValueChecks":
        self.finite(value, name)
        return self.add(lambda: value <= 0, f"{name} must be positive")

    def nonnegative(self, value: float, name: str) -> "_ValueChecks":
        self.finite(value, name)
        return self.add(lambda: value < 0, f"{name} must be non-negative")

    def run(self) -> None:
        if not self.on or not self.items:
            return
        values = [c() if callable(c) else c for c, _ in self.items]
        flags = [None] * len(values)
        by_device = {}
        for i, v in enumerate(values):
            if isinstance(v, torch.Tensor) and v.device.type != "cpu":
                by_device.setdefault(v.device, []).append(i)
        for idx in by_device.values():
            got = torch.stack([values[i].reshape(()).to(torch.bool) for i in idx]).cpu().tolist()
            for i, f in zip(idx, got):
                flags[i] = f
        for i, v in enumerate(values):
            if flags[i] is None:
                flags[i] = bool(v.item()) if isinstance(v, torch.Tensor) else bool(v)
        for f, (_, message) in zip(flags, self.items):
            if f:
                raise RuntimeCheckError(message)


def _as_tensor(x, dtype, device=None) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(dtype=dtype, device=device if device is not None else x.device)
    return torch.as_tensor(x, dtype=dtype, device=device)


def _scalar(value, name: str) -> float:
    t = torch.as_tensor(value.detach() if isinstance(value, torch.Tensor) else value, dtype=torch.float64)
    if t.ndim != 0:
        raise ValueError(f"{name} must be a scalar")
    return float(t)


def _keep_grad(original, checked: float):
    """The original 0-d tensor when it requires grad (validated through ``checked``), else the float."""
    if isinstance(original, torch.Tensor) and original.requires_grad:
        return original.to(torch.float64)
    return checked


def _checked_finite(value: float, name: str) -> float:
    error_if(not math.isfinite(float(value)), f"{name} must be finite")
    return value


def _checked_positive(value: float, name: str) -> float:
    _checked_finite(value, name)
    error_if(value <= 0, f"{name} must be positive")
    return value


def _checked_nonnegative(value: float, name: str) -> float:
    _checked_finite(value, name)
    error_if(value < 0, f"{name} must be non-negative")
    return value


def _optional_shape(shape, name: str) -> Optional[Tuple[int, int]]:
    if shape is None:
        return None
    vals = tuple(int(v) for v in (shape.tolist() if isinstance(shape, torch.Tensor) else shape))
    if len(vals) != 2:
        raise ValueError(f"{name} must have two entries")
    return vals


def _optional_positions(pos, name: str, checks: "_ValueChecks") -> Optional[torch.Tensor]:
    if pos is None:
        return None
    t = _as_tensor(pos, torch.float64)
    if t.ndim == 1 and t.shape[0] == 2:
        t = t[None, :]
    if t.ndim != 2:
        raise ValueError(f"{name} must be 2D")
    if t.shape[1] != 2:
        raise ValueError(f"{name} must have shape (P, 2)")
    checks.add(lambda: ~torch.isfinite(t).all(), f"{name} contain non-finite values")
    return t


class LossType(str, Enum):
    """Static loss selection of the reconstructors (``types/custom_types.py:50-70``): mean absolute,
    mean squared or root-mean-squared error.  ``LossType("bogus")`` raises ``ValueError``."""

    MAE = "mae"
    MSE = "mse"
    RMSE = "rmse"


@dataclass(frozen=True)
class EnsembleAxes:
    """Optional ensemble distributions (``electron_types.py:74-101``): probe modes (sample width 1, mode
    index), position jitter (width 2, Angstrom) and partial coherence (width 3: energy eV, tilt mrad x, y).
    Any axis set routes ``cbed_image`` / ``stem_4d`` through the ensemble reduction."""

    probe_modes: object = None
    position_jitter: object = None
    coherence: object = None

    def is_empty(self) -> bool:
        return self.probe_modes is None and self.position_jitter is None and self.coherence is None


@dataclass(frozen=True)
class MicroscopeConfig:
    """``electron_types.py:103-145``."""

    voltage_kv: float
    aperture_mrad: float
    defocus_ang: float
    c3_ang: float
    c5_ang: float
    ensemble: EnsembleAxes = field(default_factory=EnsembleAxes)
    probe_shape: Optional[Tuple[int, int]] = None


@dataclass(frozen=True)
class DetectorConfig:
    """``electron_types.py:147-190``."""

    real_space_calib_ang: float
    probe_calibration_pm: float
    collection_inner_mrad: float
    collection_outer_mrad: float
    scan_positions_px: Optional[torch.Tensor] = None
    scan_positions_ang: Optional[torch.Tensor] = None
    scan_shape: Optional[Tuple[int, int]] = None


@dataclass(frozen=True)
class CalibratedArray:
    """``electron_types.py:264-293``."""

    data_array: torch.Tensor
    calib_y: float
    calib_x: float
    real_space: bool


@dataclass(frozen=True)
class ProbeModes:
    """``electron_types.py:295-318``: modes (H, W, M) complex, weights (M,) normalised."""

    modes: torch.Tensor
    weights: torch.Tensor
    calib: float


@dataclass(frozen=True)
class PotentialSlices:
    """``electron_types.py:320-343``: slices (H, W, S) projected potential in V*Angstrom."""

    slices: torch.Tensor
    slice_thickness: float
    calib: float


@dataclass(frozen=True)
class STEM4D:
    """``electron_types.py:345-373``: data (P, H, W)."""

    data: torch.Tensor
    real_space_calib: float
    fourier_space_calib: float
    scan_positions: torch.Tensor
    voltage_kv: float


def _validate_distribution_axis(axis, expected_axis_id: str, expected_dim: int) -> None:
    if axis is None:
        return
    if axis.samples.ndim != 2:
        raise ValueError(f"{expected_axis_id}.samples must be 2D")
    if axis.samples.shape[1] != expected_dim:
        raise ValueError(f"{expected_axis_id}.samples must have width {expected_dim}")
    if axis.weights.ndim != 1 or tuple(axis.weights.shape) != (axis.samples.shape[0],):
        raise ValueError(f"{expected_axis_id}.weights must have shape (N,)")
    if axis.axis_id not in (None, expected_axis_id):
        raise ValueError(f"{expected_axis_id}.axis_id must be {expected_axis_id!r}")


def create_ensemble_axes(probe_modes=None, position_jitter=None, coherence=None) -> EnsembleAxes:
    """``electron_types.py:488-527``: each axis is a ``Distribution`` of the stated sample width or None."""
    _validate_distribution_axis(probe_modes, "probe_modes", 1)
    _validate_distribution_axis(position_jitter, "position_jitter", 2)
    _validate_distribution_axis(coherence, "coherence", 3)
    return EnsembleAxes(probe_modes, position_jitter, coherence)


def create_microscope_config(voltage_kv, aperture_mrad, defocus_ang=0.0, c3_ang=0.0, c5_ang=0.0,
                             ensemble: Optional[EnsembleAxes] = None, probe_shape=None) -> MicroscopeConfig:
    """``electron_types.py:529-614``."""
    voltage = _checked_positive(_scalar(voltage_kv, "voltage_kv"), "voltage_kv")
    aperture = _checked_positive(_scalar(aperture_mrad, "aperture_mrad"), "aperture_mrad")
    # the probe-forming aberrations stay tensors when they require grad: make_probe differentiates them
    defocus = _keep_grad(defocus_ang, _checked_finite(_scalar(defocus_ang, "defocus_ang"), "defocus_ang"))
    c3 = _keep_grad(c3_ang, _checked_finite(_scalar(c3_ang, "c3_ang"), "c3_ang"))
    c5 = _keep_grad(c5_ang, _checked_finite(_scalar(c5_ang, "c5_ang"), "c5_ang"))
    shape = _optional_shape(probe_shape, "probe_shape")
    return MicroscopeConfig(voltage, aperture, defocus, c3, c5,
                            create_ensemble_axes() if ensemble is None else ensemble, shape)


def create_detector_config(real_space_calib_ang, probe_calibration_pm=None, collection_inner_mrad=0.0,
                           collection_outer_mrad=0.0, scan_positions_px=None, scan_positions_ang=None,
                           scan_shape=None) -> DetectorConfig:
    """``electron_types.py:616-733``; probe_calibration_pm defaults to 100x the calibration."""
    real_calib = _scalar(real_space_calib_ang, "real_space_calib_ang")
    probe_calib = real_calib * 100.0 if probe_calibration_pm is None else _scalar(
        probe_calibration_pm, "probe_calibration_pm")
    inner = _scalar(collection_inner_mrad, "collection_inner_mrad")
    outer = _scalar(collection_outer_mrad, "collection_outer_mrad")
    checks = _ValueChecks()
    pos_px = _optional_positions(scan_positions_px, "scan_positions_px", checks)
    pos_ang = _optional_positions(scan_positions_ang, "scan_positions_ang", checks)
    shape = _optional_shape(scan_shape, "scan_shape")
    checks.positive(real_calib, "real_space_calib_ang").positive(probe_calib, "probe_calibration_pm")
    checks.nonnegative(inner, "collection_inner_mrad").nonnegative(outer, "collection_outer_mrad")
    checks.add(outer < inner, "collection_outer_mrad must be >= collection_inner_mrad").run()
    return DetectorConfig(real_calib, probe_calib, inner, outer, pos_px, pos_ang, shape)


def create_calibrated_array(data_array, calib_y, calib_x, real_space) -> CalibratedArray:
    """``electron_types.py:853-935``."""
    data = data_array if isinstance(data_array, torch.Tensor) else torch.as_tensor(data_array)
    if data.ndim != 2:
        raise ValueError("data_array must be 2D")
    cy, cx = _scalar(calib_y, "calib_y"), _scalar(calib_x, "calib_x")
    checks = _ValueChecks().positive(cy, "calib_y").positive(cx, "calib_x")
    checks.add(lambda: ~torch.isfinite(data).all(), "data_array contain non-finite values").run()
    return CalibratedArray(data, cy, cx, bool(real_space))


def create_probe_modes(modes, weights, calib) -> ProbeModes:
    """``electron_types.py:1238-1325``: weights are normalised to sum to one."""
    modes_t = modes if isinstance(modes, torch.Tensor) else torch.as_tensor(modes)
    if not modes_t.is_complex():
        modes_t = modes_t.to(torch.complex128)
    weights_t = _as_tensor(weights, torch.float64, device=modes_t.device)
    if modes_t.ndim != 3:
        raise ValueError("modes must be 3D")
    if weights_t.ndim != 1:
        raise ValueError("weights must be 1D")
    if weights_t.shape != (modes_t.shape[2],):
        raise ValueError("weights must have shape (M,)")
    calib_f = _scalar(calib, "calib")
    wsum = weights_t.sum()
    checks = _ValueChecks()
    checks.add(lambda: ~torch.isfinite(torch.view_as_real(modes_t)).all(), "modes contain non-finite values")
    checks.add(lambda: ~torch.isfinite(weights_t).all(), "weights contain non-finite values")
    checks.add(lambda: (weights_t < 0).any(), "weights must be non-negative")
    checks.add(lambda: wsum <= torch.finfo(torch.float64).eps, "weights must sum to a positive value")
    checks.positive(calib_f, "calib").run()
    return ProbeModes(modes_t, weights_t / wsum, calib_f)


def create_potential_slices(slices, slice_thickness, calib) -> PotentialSlices:
    """``electron_types.py:1328-1398``."""
    slices_t = _as_tensor(slices, torch.float64) if not (isinstance(slices, torch.Tensor)
                                                         and slices.dtype in (torch.float32, torch.float64)) else slices
    if slices_t.ndim != 3:
        raise ValueError("slices must be 3D")
    thick = _scalar(slice_thickness, "slice_thickness")
    calib_f = _scalar(calib, "calib")
    checks = _ValueChecks().add(lambda: ~torch.isfinite(slices_t).all(), "slices contain non-finite values")
    checks.positive(thick, "slice_thickness").positive(calib_f, "calib").run()
    return PotentialSlices(slices_t, thick, calib_f)


@dataclass(frozen=True)
class CrystalData:
    """``types/crystal_types.py:84-121``: positions (N, 3) A, atomic numbers (N,), lattice rows a, b, c (3, 3) A."""

    positions: torch.Tensor
    atomic_numbers: torch.Tensor
    lattice: Optional[torch.Tensor]
    stress: Optional[torch.Tensor] = None
    energy: Optional[float] = None
    properties: Optional[list] = None
    comment: Optional[str] = None


def create_crystal_data(positions, atomic_numbers, lattice=None, stress=None, energy=None, properties=None,
                        comment=None) -> CrystalData:
    """``types/crystal_types.py:244-400`` (same structural and value checks; lattice defaults to the identity)."""
    pos = _as_tensor(positions, torch.float64)
    nums = _as_tensor(atomic_numbers, torch.int32)
    lat = torch.eye(3, dtype=torch.float64) if lattice is None else _as_tensor(lattice, torch.float64)
    st = None if stress is None else _as_tensor(stress, torch.float64)
    en = None if energy is None else torch.as_tensor(energy, dtype=torch.float64)
    if pos.ndim != 2:
        raise ValueError("positions must be 2D")
    if pos.shape[1] != 3:
        raise ValueError("positions must have shape (N, 3)")
    if nums.ndim != 1:
        raise ValueError("atomic_numbers must be 1D")
    if nums.shape[0] != pos.shape[0]:
        raise ValueError("atomic_numbers must have shape (N,)")
    if tuple(lat.shape) != (3, 3):
        raise ValueError("lattice must have shape (3, 3)")
    if st is not None and tuple(st.shape) != (3, 3):
        raise ValueError("stress must have shape (3, 3)")
    if en is not None and en.ndim != 0:
        raise ValueError("energy must be a scalar")
    error_if(~torch.isfinite(pos).all(), "positions contain non-finite values")
    error_if((nums < 0).any(), "atomic_numbers must be non-negative")
    error_if(~torch.isfinite(lat).all(), "lattice contains non-finite values")
    if st is not None:
        error_if(~torch.isfinite(st).all(), "stress contains non-finite values")
    if en is not None:
        error_if(~torch.isfinite(en), "energy must be finite")
    return CrystalData(pos, nums, lat, st, None if en is None else float(en), properties, comment)


def create_stem4d(data, real_space_calib, fourier_space_calib, scan_positions, voltage_kv) -> STEM4D:
    """``electron_types.py:1401-1504``."""
    data_t = data if isinstance(data, torch.Tensor) else torch.as_tensor(data, dtype=torch.float64)
    pos_t = _as_tensor(scan_positions, torch.float64, device=data_t.device)
    if data_t.ndim != 3:
        raise ValueError("data must be 3D")
    if pos_t.ndim != 2:
        raise ValueError("scan_positions must be 2D")
    if pos_t.shape != (data_t.shape[0], 2):
        raise ValueError("scan_positions must have shape (P, 2)")
    rc = _scalar(real_space_calib, "real_space_calib")
    fc = _scalar(fourier_space_calib, "fourier_space_calib")
    kv = _scalar(voltage_kv, "voltage_kv")
    checks = _ValueChecks().add(lambda: ~torch.isfinite(data_t).all(), "data contain non-finite values")
    checks.add(lambda: ~torch.isfinite(pos_t).all(), "scan_positions contain non-finite values")
    checks.positive(rc, "real_space_calib").positive(fc, "fourier_space_calib").positive(kv, "voltage_kv").run()
    return STEM4D(data_t, rc, fc, pos_t, kv)
