"""Float64 physics set-up shared by the simulators (host side, not hot path).

Every function here builds a small table once per call (wavelength,
interaction constant, propagator factors, probe) in float64 with torch on
the caller's device; the per-position multislice work runs in the HIP
kernels.
"""

from __future__ import annotations

import math

import torch

# CODATA 2018, as in src/ptyrodactyl/types/constants.py:62-90
H_PLANCK = 6.62607015e-34
M_E = 9.1093837015e-31
E_CHARGE = 1.602176634e-19
C_LIGHT = 2.99792458e8


def relativistic_wavelength_ang(voltage_kv: float) -> float:
    """lambda = h c / sqrt(eV (2 m c^2 + eV)) in Angstrom (``types/constants.py:289-346``)."""
    ev = float(voltage_kv) * 1000.0 * E_CHARGE
    return 1e10 * math.sqrt((H_PLANCK**2 * C_LIGHT**2) / (ev * (2.0 * M_E * C_LIGHT**2 + ev)))


def interaction_sigma(voltage_kv: float) -> float:
    """sigma = (2 pi / (lambda V)) (mc^2 + eV)/(2 mc^2 + eV), rad/(V*Angstrom).

    The interaction constant of ``transmission_func``
    (``src/ptyrodactyl/multislice/simulations.py:139-157``); numerically equal
    to ``phase_interaction_parameter`` (``types/constants.py:179-244``).
    """
    voltage = float(voltage_kv) * 1000.0
    ev = E_CHARGE * voltage
    mc2 = M_E * C_LIGHT**2
    lam = relativistic_wavelength_ang(voltage_kv)
    return ((2.0 * math.pi / (lam * voltage)) * (mc2 + ev)) / (2.0 * mc2 + ev)


phase_interaction_parameter = interaction_sigma


def fftfreq(n: int, d: float, device=None) -> torch.Tensor:
    """numpy-compatible fftfreq in float64 (Nyquist bin negative for even n)."""
    return torch.fft.fftfreq(int(n), d=float(d), dtype=torch.float64, device=device)


def propagator_factors(n: int, thickness_ang: float, voltage_kv: float, calib_ang: float, device=None):
    """Separable Fresnel propagator exp(-i pi lambda dz q^2) / n on the fftfreq grid.

    ``propagation_func`` (``simulations.py:162-225``) is exp(-i pi lambda dz
    (qx^2 + qy^2)) = Py(qy) Px(qx); the 1/n per axis folds the inverse-FFT
    normalisation into the factor.  Returns a complex64 tensor of n values
    (used for both axes on a square grid).
    """
    q = fftfreq(n, calib_ang, device=device)
    lam = relativistic_wavelength_ang(voltage_kv)
    phase = -math.pi * lam * float(thickness_ang) * q * q
    f = torch.polar(torch.full_like(phase, 1.0 / n), phase)
    return f.to(torch.complex64).contiguous()


def propagator_pair(imsize_y: int, imsize_x: int, thickness_ang: float, voltage_kv: float, calib_ang: float,
                    device=None):
    """(Py / h, Px / w) of ``propagation_func`` on an h x w grid; one shared tensor when h == w."""
    fy = propagator_factors(imsize_y, thickness_ang, voltage_kv, calib_ang, device=device)
    if int(imsize_x) == int(imsize_y):
        return fy, fy
    return fy, propagator_factors(imsize_x, thickness_ang, voltage_kv, calib_ang, device=device)


def shift_to_pixels(positions_ang: torch.Tensor, calib_ang: float) -> torch.Tensor:
    """Probe shifts in pixels for ``shift_beam_fourier``'s Angstrom positions."""
    return (positions_ang.to(torch.float64) / float(calib_ang)).contiguous()
