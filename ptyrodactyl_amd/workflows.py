"""Crystal -> 4D-STEM workflows with the reference's signatures.

Mirrors ``src/ptyrodactyl/workflows/stem_4d.py``: ``crystal2stem4d``
(``:191-550``) and ``crystal2stem4d_tiled`` (``:553-1007``) with the same
arguments and defaults, and the memory estimate ``_estimate_memory_gb``
(``:57-138``).  The multislice runs in this package's HIP kernels, which take
grids of any shape with edges up to 4096 pixels (square powers of two on the
dedicated kernels, every other shape on the runtime-size ones); a workflow
whose grid exceeds that raises ``ValueError`` (choose the pixel size, padding
or ``grid_pixels`` / ``fourier_pixels`` accordingly).  ``force_parallel``
is accepted for signature compatibility: this build runs on one GPU, where the
reference's parallel and serial branches compute the same patterns.

Clip and resize: the pattern is cut to +-cbed_extent_mrad around its centre
and resampled to ``cbed_shape`` as ``jax.image.resize(..., "linear")`` does
(antialiased triangle kernel, per-output normalisation, equal-size axes
untouched), applied on the GPU as two FP64 matrix products per pattern batch.
"""

from __future__ import annotations

import math
from typing import Optional, Tuple

import torch

from . import engine
from .atom_potentials import single_atom_potential
from .parallelized import cbed_image_from_atoms, create_atomic_slice_data, stem4d_sharded
from .physics import relativistic_wavelength_ang
from .simulations import _device_for, make_probe, probe_modes_to_distribution
from .types import (STEM4D, CrystalData, _as_tensor, create_detector_config, create_microscope_config,
                    create_probe_modes, create_stem4d)

# square grids that run on the dedicated power-of-two kernels; any other h x w up to GRID_MAX_EDGE also runs
SUPPORTED_GRIDS = engine.SUPPORTED_N
GRID_MAX_EDGE = engine.GRID_MAX_EDGE
LARGE_POSITION_THRESHOLD = 100
_POTENTIAL_EXTENT_ANG = 10.0


def estimate_memory_gb(num_positions: int, height: int, width: int, num_modes: int, num_slices: int) -> float:
    """The reference's footprint estimate: (complex128 shifted beams + float64 patterns + float64 slices) x 2.5."""
    total = (num_positions * height * width * num_modes * 16 + num_positions * height * width * 8
             + height * width * num_slices * 8) * 2.5
    return total / (1024 ** 3)


def device_memory_gb(device=None) -> float:
    """Total memory of the GPU in GiB (16 when none is visible, as the reference's fallback)."""
    if not torch.cuda.is_available():
        return 16.0
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    return torch.cuda.get_device_properties(dev).total_memory / 1024 ** 3


def resize_weights(in_size: int, out_size: int, device=None) -> torch.Tensor:
    """(in, out) float64 weights of the antialiased linear resize of one axis."""
    scale = out_size / in_size
    inv = 1.0 / scale
    kscale = max(inv, 1.0)
    sample = (torch.arange(out_size, dtype=torch.float64, device=device) + 0.5) * inv - 0.5
    x = (sample[None, :] - torch.arange(in_size, dtype=torch.float64, device=device)[:, None]).abs() / kscale
    w = torch.clamp(1.0 - x.abs(), min=0.0)
    tot = w.sum(dim=0, keepdim=True)
    w = torch.where(tot.abs() > 1000.0 * torch.finfo(torch.float32).eps, w / torch.where(tot != 0, tot, 1.0),
                    torch.zeros_like(w))
    inside = (sample >= -0.5) & (sample <= in_size - 0.5)
    return torch.where(inside[None, :], w, torch.zeros_like(w))


def resize_linear(images: torch.Tensor, shape: Tuple[int, int]) -> torch.Tensor:
    """``jax.image.resize(images, (..., *shape), "linear")`` over the last two axes, float64."""
    out = images.to(torch.float64)
    h, w = out.shape[-2:]
    if int(shape[0]) != h:
        out = torch.matmul(resize_weights(h, int(shape[0]), out.device).T, out)
    if int(shape[1]) != w:
        out = torch.matmul(out, resize_weights(w, int(shape[1]), out.device))
    return out


def _clip_window(size: int, fourier_calib: float, extent_mrad: float, voltage_kv: float) -> Tuple[int, int]:
    extent_px = int(math.ceil(float(extent_mrad) / (relativistic_wavelength_ang(voltage_kv) * 1000.0)
                              / fourier_calib))
    c = size // 2
    return max(0, c - extent_px), min(size, c + extent_px)


def _wrapped_kernels(elements, pixel: float, grid_hw: Tuple[int, int], supersampling: int,
                     device) -> torch.Tensor:
    k = int(math.ceil(_POTENTIAL_EXTENT_ANG / pixel))
    k += 1 - k % 2
    hk = k // 2
    if k > min(grid_hw):
        raise ValueError(f"the {k}-pixel atom kernel ({_POTENTIAL_EXTENT_ANG} A at {pixel} A) exceeds the "
                         f"{grid_hw} grid")
    out = torch.zeros((len(elements), *grid_hw), dtype=torch.float64, device=device)
    for i, z in enumerate(elements):
        small = single_atom_potential(int(z), pixel, (k, k), torch.zeros(2, dtype=torch.float64, device=device),
                                      supersampling)
        out[i, :hk + 1, :hk + 1] = small[hk:, hk:]
        out[i, :hk + 1, -hk:] = small[hk:, :hk]
        out[i, -hk:, :hk + 1] = small[:hk, hk:]
        out[i, -hk:, -hk:] = small[:hk, :hk]
    return out


def _slice_bounds(z: torch.Tensor, dz: float) -> torch.Tensor:
    z_min = float(z.min())
    count = max(1, int(math.ceil(float(z.max() - z.min()) / dz)))
    return torch.tensor([[z_min + i * dz, z_min + i * dz + dz] for i in range(count)], dtype=torch.float64)


def _element_types(numbers: torch.Tensor):
    elements = sorted({int(v) for v in numbers.tolist()})
    index = {z: i for i, z in enumerate(elements)}
    return elements, torch.tensor([index[int(v)] for v in numbers.tolist()], dtype=torch.int64)


def _check_grid(name: str, h: int, w: int) -> None:
    if not (1 <= h <= GRID_MAX_EDGE and 1 <= w <= GRID_MAX_EDGE):
        raise ValueError(f"{name}: the multislice grid is {h} x {w} pixels; the GPU kernels take edges of 1 to "
                         f"{GRID_MAX_EDGE} pixels")


def _probe_modes(microscope, detector, num_modes: int, pixel: float, device):
    probe = make_probe(microscope, detector).to(device)
    modes = torch.stack([probe] * int(num_modes), dim=-1)
    return create_probe_modes(modes, torch.ones(modes.shape[-1], dtype=torch.float64), pixel)


def crystal2stem4d(crystal_data: CrystalData, scan_positions, voltage_kv, cbed_aperture_mrad,
                   cbed_extent_mrad=50.0, cbed_shape: Tuple[int, int] = (256, 256),
                   real_space_pixel_size_ang=0.02, slice_thickness=1.0, num_modes: int = 1, probe_defocus=0.0,
                   probe_c3=0.0, probe_c5=0.0, padding: float = 4.0, supersampling: int = 4,
                   force_parallel: Optional[bool] = None) -> STEM4D:
    """4D-STEM of a crystal on one grid spanning it plus ``padding``: per-element potentials from the Lobato
    table, atoms binned into slices of ``slice_thickness``, ``stem4d_sharded`` over ``scan_positions`` (A), each
    pattern clipped to +-``cbed_extent_mrad`` and resized to ``cbed_shape``."""
    pixel, dz = float(real_space_pixel_size_ang), float(slice_thickness)
    pos = crystal_data.positions.to(torch.float64)
    width = int(math.ceil((float(pos[:, 0].max() - pos[:, 0].min()) + 2 * padding) / pixel))
    height = int(math.ceil((float(pos[:, 1].max() - pos[:, 1].min()) + 2 * padding) / pixel))
    _check_grid("crystal2stem4d", height, width)
    dev = _device_for(crystal_data.positions)
    scan = _as_tensor(scan_positions, torch.float64)
    microscope = create_microscope_config(voltage_kv=voltage_kv, aperture_mrad=cbed_aperture_mrad,
                                          defocus_ang=probe_defocus, c3_ang=probe_c3, c5_ang=probe_c5,
                                          probe_shape=(height, width))
    detector = create_detector_config(real_space_calib_ang=pixel, probe_calibration_pm=pixel * 100.0,
                                      scan_positions_ang=scan)
    beam = _probe_modes(microscope, detector, num_modes, pixel, dev)
    elements, types = _element_types(crystal_data.atomic_numbers)
    kern = _wrapped_kernels(elements, pixel, (height, width), supersampling, dev)
    sample = create_atomic_slice_data(pos.to(dev), types.to(dev), _slice_bounds(pos[:, 2], dz).to(dev), kern)
    raw = stem4d_sharded(beam, sample, microscope, detector)
    y0, y1 = _clip_window(height, float(raw.fourier_space_calib), cbed_extent_mrad, voltage_kv)
    x0, x1 = _clip_window(width, float(raw.fourier_space_calib), cbed_extent_mrad, voltage_kv)
    data = resize_linear(raw.data[:, y0:y1, x0:x1], cbed_shape)
    out_cal = 2.0 * float(cbed_extent_mrad) / cbed_shape[0] / (relativistic_wavelength_ang(voltage_kv) * 1000.0)
    return create_stem4d(data, pixel, out_cal, raw.scan_positions, voltage_kv)


def fourier_pixels_for(voltage_kv, cbed_extent_mrad, cbed_shape, pixel_size_ang, grid_pixels: int) -> int:
    """Smallest power of two whose FFT grid samples the pattern at 2 cbed_extent_mrad / cbed_shape[0], at least
    ``grid_pixels`` (``stem_4d.py:735-760``)."""
    target = 2.0 * float(cbed_extent_mrad) / float(cbed_shape[0]) / (relativistic_wavelength_ang(voltage_kv) * 1000.0)
    need = int(math.ceil((1.0 / target) / float(pixel_size_ang)))
    n = 1
    while n < need:
        n *= 2
    return max(n, int(grid_pixels))


def crystal2stem4d_tiled(crystal_data: CrystalData, scan_positions, voltage_kv, cbed_aperture_mrad,
                         cbed_extent_mrad=50.0, cbed_shape: Tuple[int, int] = (256, 256), tile_size_ang: float = 40.0,
                         grid_pixels: int = 4096, pixel_size_ang: float = 0.02, fourier_pixels: Optional[int] = None,
                         slice_thickness=1.0, num_modes: int = 1, probe_defocus=0.0, probe_c3=0.0, probe_c5=0.0,
                         supersampling: int = 4) -> STEM4D:
    """4D-STEM of a large crystal with one tile per scan position: the atoms within the grid_pixels x
    pixel_size_ang square centred on the position (local coordinates), the probe at the tile centre, an FFT grid
    of ``fourier_pixels`` (default: the power of two that samples the output pattern), then the same clip and
    resize as :func:`crystal2stem4d`."""
    pixel, dz = float(pixel_size_ang), float(slice_thickness)
    grid_size = int(grid_pixels) * pixel
    if (grid_size - float(tile_size_ang)) / 2.0 < 0:
        raise ValueError(f"Tile size ({tile_size_ang} Å) exceeds grid size ({grid_size} Å). Increase grid_pixels "
                         f"or decrease tile_size_ang.")
    if fourier_pixels is None:
        n = fourier_pixels_for(voltage_kv, cbed_extent_mrad, cbed_shape, pixel, grid_pixels)
    else:
        if int(fourier_pixels) < int(grid_pixels):
            raise ValueError(f"fourier_pixels ({fourier_pixels}) must be >= grid_pixels ({grid_pixels}).")
        n = int(fourier_pixels)
    _check_grid("crystal2stem4d_tiled", n, n)
    dev = _device_for(crystal_data.positions)
    pos = crystal_data.positions.to(device=dev, dtype=torch.float64)
    bounds = _slice_bounds(pos[:, 2], dz).to(dev)
    elements, types = _element_types(crystal_data.atomic_numbers)
    types = types.to(dev)
    kern = _wrapped_kernels(elements, pixel, (n, n), supersampling, dev)
    microscope = create_microscope_config(voltage_kv=voltage_kv, aperture_mrad=cbed_aperture_mrad,
                                          defocus_ang=probe_defocus, c3_ang=probe_c3, c5_ang=probe_c5,
                                          probe_shape=(n, n))
    detector = create_detector_config(real_space_calib_ang=pixel, probe_calibration_pm=pixel * 100.0)
    beam = _probe_modes(microscope, detector, num_modes, pixel, dev)
    distribution = probe_modes_to_distribution(beam)
    half = grid_size / 2.0
    q = torch.fft.fftfreq(n, d=pixel, dtype=torch.float64, device=dev)
    ramp = torch.exp(-2j * math.pi * (q[:, None] * half + q[None, :] * half))
    modes = beam.modes.to(device=dev, dtype=torch.complex128)
    shifted = torch.fft.ifft2(torch.fft.fft2(modes, dim=(0, 1)) * ramp[..., None], dim=(0, 1))
    scan = _as_tensor(scan_positions, torch.float64)
    raw = []
    for py, px in scan.tolist():
        ty, tx = py - half, px - half
        mask = ((pos[:, 0] >= tx) & (pos[:, 0] < tx + grid_size) & (pos[:, 1] >= ty)
                & (pos[:, 1] < ty + grid_size)).to(torch.float64)
        local = torch.stack([pos[:, 0] - tx, pos[:, 1] - ty, pos[:, 2]], dim=1)
        raw.append(cbed_image_from_atoms(shifted, distribution, local, types, bounds, kern, voltage_kv, pixel,
                                         atom_mask=mask))
    raw = torch.stack(raw)
    y0, y1 = _clip_window(n, 1.0 / (n * pixel), cbed_extent_mrad, voltage_kv)
    data = resize_linear(raw[:, y0:y1, y0:y1], cbed_shape)
    out_cal = 2.0 * float(cbed_extent_mrad) / cbed_shape[0] / (relativistic_wavelength_ang(voltage_kv) * 1000.0)
    return create_stem4d(data, pixel, out_cal, scan, voltage_kv)


__all__ = ["crystal2stem4d", "crystal2stem4d_tiled", "device_memory_gb", "estimate_memory_gb", "fourier_pixels_for",
           "resize_linear", "resize_weights"]
