"""CPU oracle for the crystal -> 4D-STEM workflows.

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` may import this module, and only
as the checker; the product path never calls into it.

A float64 numpy restatement (paths relative to the reference root) of
``workflows/stem_4d.py``:

* ``_estimate_memory_gb`` (``:57-138``);
* ``crystal2stem4d`` (``:191-550``): grid from the crystal extent plus
  padding, per-element potentials of a 10 A kernel wrapped into the grid's
  corners, ``stem4d_sharded`` over the scan, clip of +-cbed_extent_mrad around
  the pattern centre, linear resize to ``cbed_shape``;
* ``crystal2stem4d_tiled`` (``:553-1007``): a power-of-two FFT grid, per
  position only the atoms inside the tile around it (local coordinates), the
  probe at the tile centre, then the same clip and resize.

``jax.image.resize(..., method="linear")`` (jax 0.9.2, antialias on, the
dependency the reference calls at ``:505-520, 968-975``) is restated from its
published algorithm: per axis an (in, out) weight matrix of the triangle kernel
at |(i + 0.5) / scale - 0.5 - j| / max(1 / scale, 1), normalised per output
sample, with axes of equal size left untouched.

Pins (``tests/test_workflows.py``): the resize against torch's independent
antialiased bilinear interpolation and against its closed forms (identity,
the [1, 3, 3, 1] / 8 kernel of a 2x reduction, exact linear interpolation on
a ramp); the pipelines against compositions of ``ms_oracle`` and
``potential_oracle`` functions that are pinned on their own.
"""

from __future__ import annotations

import math

import numpy as np

from . import ms_oracle as o
from . import potential_oracle as po


def estimate_memory_gb(num_positions, height, width, num_modes, num_slices) -> float:
    total = (num_positions * height * width * num_modes * 16 + num_positions * height * width * 8
             + height * width * num_slices * 8) * 2.5
    return total / (1024 ** 3)


def resize_weights(in_size: int, out_size: int) -> np.ndarray:
    """(in, out) weights of jax.image.resize's linear kernel with antialiasing."""
    scale = out_size / in_size
    inv = 1.0 / scale
    kscale = max(inv, 1.0)
    sample = (np.arange(out_size) + 0.5) * inv - 0.5
    x = np.abs(sample[None, :] - np.arange(in_size)[:, None]) / kscale
    w = np.maximum(0.0, 1.0 - np.abs(x))
    tot = w.sum(axis=0, keepdims=True)
    w = np.where(np.abs(tot) > 1000.0 * float(np.finfo(np.float32).eps), w / np.where(tot != 0, tot, 1.0), 0.0)
    inside = (sample >= -0.5) & (sample <= in_size - 0.5)
    return np.where(inside[None, :], w, 0.0)


def resize_linear(images, shape):
    """Resize the last two axes to ``shape``."""
    out = np.asarray(images, np.float64)
    h, w = out.shape[-2:]
    if shape[0] != h:
        out = np.einsum("...hw,ho->...ow", out, resize_weights(h, shape[0]))
    if shape[1] != w:
        out = np.einsum("...hw,wo->...ho", out, resize_weights(w, shape[1]))
    return out


def clip_window(size: int, fourier_calib: float, extent_mrad: float, voltage_kv: float):
    extent_px = int(math.ceil(extent_mrad / (o.wavelength_ang(voltage_kv) * 1000.0) / fourier_calib))
    c = size // 2
    return max(0, c - extent_px), min(size, c + extent_px)


def wrapped_kernels(rows, kind, elements, pixel, grid_hw, supersampling):
    """Each element's potential on a (K, K) grid, K = ceil(10 A / pixel) rounded up to odd, moved so its centre
    sample sits at pixel (0, 0) of the (H, W) grid (``stem_4d.py:407-452``)."""
    k = int(math.ceil(10.0 / pixel))
    k += 1 - k % 2
    hk = k // 2
    out = []
    for z in elements:
        small = po.single_atom_potential(kind, rows[int(z)], pixel, (k, k), (0.0, 0.0), supersampling)
        pad = np.zeros(grid_hw)
        pad[:hk + 1, :hk + 1] = small[hk:, hk:]
        pad[:hk + 1, -hk:] = small[hk:, :hk]
        pad[-hk:, :hk + 1] = small[:hk, hk:]
        pad[-hk:, -hk:] = small[:hk, :hk]
        out.append(pad)
    return np.stack(out)


def slice_bounds(z, dz):
    z_min, z_range = float(np.min(z)), float(np.max(z) - np.min(z))
    s = max(1, int(math.ceil(z_range / dz)))
    return np.asarray([[z_min + i * dz, z_min + i * dz + dz] for i in range(s)])


def crystal2stem4d(rows, positions, numbers, scan_positions, voltage_kv, aperture_mrad, extent_mrad, cbed_shape,
                   pixel, dz, num_modes, defocus, c3, c5, padding, supersampling, kind="lobato"):
    """(data (P, Ho, Wo), fourier calib of the output, raw patterns)."""
    pos = np.asarray(positions, np.float64)
    width = int(np.ceil((float(pos[:, 0].max() - pos[:, 0].min()) + 2 * padding) / pixel))
    height = int(np.ceil((float(pos[:, 1].max() - pos[:, 1].min()) + 2 * padding) / pixel))
    probe = o.make_probe(voltage_kv, aperture_mrad, defocus, c3, c5, (height, width), pixel * 100.0)
    modes = np.stack([probe] * num_modes, axis=-1)
    elements = sorted({int(v) for v in numbers})
    types = np.asarray([elements.index(int(v)) for v in numbers])
    kern = wrapped_kernels(rows, kind, elements, pixel, (height, width), supersampling)
    raw = o.stem4d_from_atoms(modes, np.full(num_modes, 1.0 / num_modes), pos, types, slice_bounds(pos[:, 2], dz),
                              kern, voltage_kv, pixel, np.asarray(scan_positions, np.float64))
    fcal = 1.0 / (height * pixel)
    y0, y1 = clip_window(height, fcal, extent_mrad, voltage_kv)
    x0, x1 = clip_window(width, fcal, extent_mrad, voltage_kv)
    data = resize_linear(raw[:, y0:y1, x0:x1], cbed_shape)
    out_cal = 2.0 * extent_mrad / cbed_shape[0] / (o.wavelength_ang(voltage_kv) * 1000.0)
    return data, out_cal, raw


def fft_pixels_for(voltage_kv, extent_mrad, cbed_shape, pixel, grid_pixels):
    target = 2.0 * extent_mrad / cbed_shape[0] / (o.wavelength_ang(voltage_kv) * 1000.0)
    need = int(np.ceil((1.0 / target) / pixel))
    n = 1
    while n < need:
        n *= 2
    return max(n, grid_pixels)


def crystal2stem4d_tiled(rows, positions, numbers, scan_positions, voltage_kv, aperture_mrad, extent_mrad,
                         cbed_shape, tile_size, grid_pixels, pixel, fft_pixels, dz, num_modes, defocus, supersampling,
                         kind="lobato"):
    pos = np.asarray(positions, np.float64)
    g = grid_pixels * pixel
    n = fft_pixels
    probe = o.make_probe(voltage_kv, aperture_mrad, defocus, 0.0, 0.0, (n, n), pixel * 100.0)
    modes = np.stack([probe] * num_modes, axis=-1)
    elements = sorted({int(v) for v in numbers})
    types = np.asarray([elements.index(int(v)) for v in numbers])
    kern = wrapped_kernels(rows, kind, elements, pixel, (n, n), supersampling)
    bounds = slice_bounds(pos[:, 2], dz)
    raw = []
    for py, px in np.asarray(scan_positions, np.float64):
        ty, tx = py - g / 2.0, px - g / 2.0
        mask = ((pos[:, 0] >= tx) & (pos[:, 0] < tx + g) & (pos[:, 1] >= ty) & (pos[:, 1] < ty + g)).astype(float)
        local = np.column_stack([pos[:, 0] - tx, pos[:, 1] - ty, pos[:, 2]])
        raw.append(o.stem4d_from_atoms(modes, np.full(num_modes, 1.0 / num_modes), local, types, bounds, kern,
                                       voltage_kv, pixel, np.asarray([[g / 2.0, g / 2.0]]), atom_mask=mask)[0])
    raw = np.stack(raw)
    y0, y1 = clip_window(n, 1.0 / (n * pixel), extent_mrad, voltage_kv)
    data = resize_linear(raw[:, y0:y1, y0:y1], cbed_shape)
    return data, raw
