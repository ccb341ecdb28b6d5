"""Seeded synthetic specimens and scan geometries for the benchmark workloads.

A cubic perovskite-like crystal (a = 3.9 Angstrom; Z = 38 at (0,0,0),
Z = 22 at (1/2,1/2,1/2), Z = 8 at the three face centres (1/2,1/2,0),
(1/2,0,1/2), (0,1/2,1/2)) with seeded Gaussian thermal jitter; each atom
contributes a Z-scaled Gaussian blob of width 0.3 Angstrom to the projected
potential of the slice its z coordinate falls in.  The potential is periodic
over the slab (atoms wrap modulo the field of view).  All arithmetic is
float64 numpy; the transmission exp(i sigma V) is returned as complex64.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import physics

LATTICE_A = 3.9
SITES = ((38, (0.0, 0.0, 0.0)), (22, (0.5, 0.5, 0.5)), (8, (0.5, 0.5, 0.0)), (8, (0.5, 0.0, 0.5)),
         (8, (0.0, 0.5, 0.5)))
BLOB_WIDTH = 0.3          # Angstrom (Gaussian sigma)
PEAK_PER_Z = 10.0         # V*Angstrom peak of the projected blob per unit Z
THERMAL_SIGMA = 0.05      # Angstrom


def crystal_atoms(cells_y: int, cells_x: int, cells_z: int, seed: int = 0, jitter: float = THERMAL_SIGMA):
    """Atom list (N, 4) = (y, x, z, Z) in Angstrom for a block of unit cells."""
    rng = np.random.default_rng(seed)
    iy, ix, iz = np.meshgrid(np.arange(cells_y), np.arange(cells_x), np.arange(cells_z), indexing="ij")
    cells = np.stack([iy.ravel(), ix.ravel(), iz.ravel()], axis=1).astype(np.float64)
    atoms = []
    for z_num, frac in SITES:
        pos = (cells + np.asarray(frac)[[1, 0, 2]]) * LATTICE_A  # frac given as (x, y, z)
        atoms.append(np.concatenate([pos, np.full((len(pos), 1), float(z_num))], axis=1))
    atoms = np.concatenate(atoms, axis=0)
    atoms[:, :3] += rng.normal(0.0, jitter, size=(len(atoms), 3))
    return atoms


def projected_potential(atoms, grid: int, calib: float, num_slices: int, slice_thickness: float,
                        width: float = BLOB_WIDTH, peak_per_z: float = PEAK_PER_Z):
    """Periodic (S, grid, grid) float64 projected potential in V*Angstrom."""
    fov = grid * calib
    pot = np.zeros((num_slices, grid * grid), np.float64)
    sl = np.floor(atoms[:, 2] / slice_thickness).astype(np.int64)
    keep = (sl >= 0) & (sl < num_slices)
    radius = int(math.ceil(5 * width / calib))
    offs = np.arange(-radius, radius + 1)
    for s in range(num_slices):
        sel = atoms[keep & (sl == s)]
        if len(sel) == 0:
            continue
        y = np.mod(sel[:, 0], fov) / calib
        x = np.mod(sel[:, 1], fov) / calib
        cy = np.floor(y).astype(np.int64)
        cx = np.floor(x).astype(np.int64)
        py = cy[:, None] + offs[None, :]
        px = cx[:, None] + offs[None, :]
        gy = np.exp(-((py - y[:, None]) * calib) ** 2 / (2 * width**2))
        gx = np.exp(-((px - x[:, None]) * calib) ** 2 / (2 * width**2))
        amp = peak_per_z * sel[:, 3]
        stamps = amp[:, None, None] * gy[:, :, None] * gx[:, None, :]
        flat = (np.mod(py, grid)[:, :, None] * grid + np.mod(px, grid)[:, None, :])
        pot[s] += np.bincount(flat.ravel(), weights=stamps.ravel(), minlength=grid * grid)
    return pot.reshape(num_slices, grid, grid)


def transmission(pot: np.ndarray, voltage_kv: float) -> np.ndarray:
    """exp(i sigma V) as complex64 (S, H, W)."""
    sigma = physics.interaction_sigma(voltage_kv)
    return np.exp(1j * sigma * pot).astype(np.complex64)


@dataclass
class Workload:
    name: str
    n: int
    num_slices: int
    slice_thickness: float
    calib: float
    voltage_kv: float
    aperture_mrad: float
    defocus_ang: float
    c3_ang: float           # reference units (x1e7 -> Angstrom)
    slab_t: np.ndarray      # (S, Hs, Ws) complex64
    origins: np.ndarray     # (P, 2) int32
    shifts_px: np.ndarray   # (P, 2) float64
    probe: np.ndarray       # (n, n) complex128 real-space probe (make_probe convention)


def probe_for(n: int, calib: float, voltage_kv: float, aperture_mrad: float, defocus_ang: float,
              c3_ang: float) -> np.ndarray:
    """make_probe (``simulations.py:334-431``) evaluated in numpy float64."""
    lam = physics.relativistic_wavelength_ang(voltage_kv)
    fov = n * calib
    q1 = np.roll(np.arange(-n / 2, n / 2, 1.0) / fov, n // 2)
    qy, qx = np.meshgrid(q1, q1, indexing="ij")
    q = np.sqrt(qx**2 + qy**2)
    p = lam * q
    chi = (2 * np.pi / lam) * (defocus_ang * p**2 / 2 + c3_ang * 1e7 * p**4 / 4)
    a = (q <= (aperture_mrad / 1000.0) / lam) * np.exp(-1j * chi)
    return np.fft.ifftshift(np.fft.ifft2(a))


def config1() -> Workload:
    """128^2 grid at 0.15 A, 20 slices of 1.95 A, 200 kV, 20 mrad, 16x16 scan (slab = window)."""
    n, calib, s_count, dz, kv = 128, 0.15, 20, 1.95, 200.0
    atoms = crystal_atoms(5, 5, 10, seed=0)
    pot = projected_potential(atoms, n, calib, s_count, dz)
    step = n // 16
    iy, ix = np.meshgrid(np.arange(16), np.arange(16), indexing="ij")
    shifts = np.stack([iy.ravel() * step, ix.ravel() * step], axis=1).astype(np.float64)
    return Workload("stem4d_128x20_p256", n, s_count, dz, calib, kv, 20.0, 0.0, 0.0, transmission(pot, kv),
                    np.zeros((len(shifts), 2), np.int32), shifts, probe_for(n, calib, kv, 20.0, 0.0, 0.0))


def config2(scan: int = 64, step: int = 12) -> Workload:
    """1024^2 slab at 0.1 A, 64 slices of 1 A, 256^2 windows, 300 kV / 25 mrad / C1 50 A / Cs 1 um,
    scan x scan positions on a `step`-pixel raster with seeded sub-pixel remainders."""
    n, grid, calib, s_count, dz, kv = 256, 1024, 0.1, 64, 1.0, 300.0
    cells = int(math.ceil(grid * calib / LATTICE_A))
    atoms = crystal_atoms(cells, cells, int(math.ceil(s_count * dz / LATTICE_A)) + 1, seed=0)
    pot = projected_potential(atoms, grid, calib, s_count, dz)
    rng = np.random.default_rng(1)
    iy, ix = np.meshgrid(np.arange(scan), np.arange(scan), indexing="ij")
    origins = np.stack([iy.ravel() * step, ix.ravel() * step], axis=1).astype(np.int32)
    shifts = rng.uniform(-0.5, 0.5, size=(scan * scan, 2))
    return Workload("stem4d_256x64_p4096_slab1024", n, s_count, dz, calib, kv, 25.0, 50.0, 1e-3,
                    transmission(pot, kv), origins, shifts, probe_for(n, calib, kv, 25.0, 50.0, 1e-3))


ABERRATIONS_TRUE = (50.0, 12.0, -8.0, 150.0, -90.0, 1.0e4, 400.0, 250.0, -1200.0, 800.0, 3000.0, -2000.0)


@dataclass
class AberrationWorkload:
    n: int
    num_slices: int
    slice_thickness: float
    calib: float
    voltage_kv: float
    aperture_mrad: float
    slab_t: np.ndarray       # (S, H, W) complex64: the one object slice repeated S times
    origins: np.ndarray      # (P, 2) int32
    shifts_px: np.ndarray    # (P, 2) float64
    coeffs_true: np.ndarray  # (12,) Angstrom
    coeffs_init: np.ndarray  # (12,) seeded +-20 % offsets
    object_true: np.ndarray = None  # (H, W) complex64 object slice (each of the S slices repeats it)
    object_init: np.ndarray = None  # (H, W) complex64 start of the joint fit: seeded +-5 % amplitude noise


def config4(scan: int = 24, n: int = 128, object_px: int = 256, num_slices: int = 8) -> AberrationWorkload:
    """Aberration fit: one 256^2 object slice (config-1 generator, 0.15 A) applied at each of 8 slices,
    128^2 windows on a scan x scan raster, 300 kV / 20 mrad; initial coefficients offset by seeded +-20 %,
    and for the joint fit an initial object with seeded +-5 % amplitude noise."""
    calib, kv, dz = 0.15, 300.0, 1.95
    cells = int(math.ceil(object_px * calib / LATTICE_A))
    atoms = crystal_atoms(cells, cells, 1, seed=0)
    pot = projected_potential(atoms, object_px, calib, 1, LATTICE_A)
    t1 = transmission(pot, kv)[0]
    slab = np.repeat(t1[None], num_slices, axis=0)
    span = object_px - n
    lin = np.round(np.linspace(0, span, scan)).astype(np.int32)
    iy, ix = np.meshgrid(lin, lin, indexing="ij")
    origins = np.stack([iy.ravel(), ix.ravel()], axis=1).astype(np.int32)
    shifts = np.zeros((len(origins), 2))
    rng = np.random.default_rng(3)
    true = np.asarray(ABERRATIONS_TRUE, np.float64)
    init = true * (1.0 + rng.uniform(-0.2, 0.2, size=12))
    obj0 = (t1 * (1.0 + np.random.default_rng(4).uniform(-0.05, 0.05, size=t1.shape))).astype(np.complex64)
    return AberrationWorkload(n, num_slices, dz, calib, kv, 20.0, slab, origins, shifts, true, init, t1, obj0)


def projected_potential_torch(atoms: np.ndarray, grid: int, calib: float, num_slices: int, slice_thickness: float,
                              device, width: float = BLOB_WIDTH, peak_per_z: float = PEAK_PER_Z):
    """GPU version of :func:`projected_potential` (same blobs, periodic, float64)."""
    import torch
    a = torch.as_tensor(atoms, dtype=torch.float64, device=device)
    sl = torch.floor(a[:, 2] / slice_thickness).to(torch.int64)
    keep = (sl >= 0) & (sl < num_slices)
    a, sl = a[keep], sl[keep]
    fov = grid * calib
    y = torch.remainder(a[:, 0], fov) / calib
    x = torch.remainder(a[:, 1], fov) / calib
    radius = int(math.ceil(5 * width / calib))
    offs = torch.arange(-radius, radius + 1, device=device)
    py = torch.floor(y).to(torch.int64)[:, None] + offs[None, :]
    px = torch.floor(x).to(torch.int64)[:, None] + offs[None, :]
    gy = torch.exp(-((py - y[:, None]) * calib) ** 2 / (2 * width**2))
    gx = torch.exp(-((px - x[:, None]) * calib) ** 2 / (2 * width**2))
    stamps = (peak_per_z * a[:, 3])[:, None, None] * gy[:, :, None] * gx[:, None, :]
    flat = (sl[:, None, None] * grid * grid + torch.remainder(py, grid)[:, :, None] * grid
            + torch.remainder(px, grid)[:, None, :])
    pot = torch.zeros(num_slices * grid * grid, dtype=torch.float64, device=device)
    pot.index_add_(0, flat.reshape(-1), stamps.reshape(-1))
    return pot.reshape(num_slices, grid, grid)


@dataclass
class GradientWorkload:
    n: int
    num_slices: int
    slice_thickness: float
    calib: float
    voltage_kv: float
    object_t: np.ndarray     # (S, H, W) complex64 object to differentiate at (+-5 % amplitude noise)
    object_true: np.ndarray  # (S, H, W) complex64 unperturbed object
    probe: np.ndarray        # (n, n) complex128
    origins: np.ndarray
    shifts_px: np.ndarray
    measured: np.ndarray     # (P, n, n) float64 fftshifted Poisson intensities (units of the simulation)
    ckpt_every: int = 4


def config3(scan: int = 32, n: int = 128, object_px: int = 512, num_slices: int = 16) -> GradientWorkload:
    """Gradient workload: 16 object slices of 512^2 (config-1 generator at 0.15 A) with seeded +-5 %
    amplitude noise, 128^2 probe at 300 kV / 20 mrad / C1 = 30 A, 32 x 32 scan with seeded +-0.2 px
    jitter; measured intensities from the unperturbed object with Poisson noise at 1e5 e- per pattern
    (computed by the caller-supplied forward; filled in by :func:`config3_measure`)."""
    calib, kv, dz = 0.15, 300.0, 1.95
    cells = int(math.ceil(object_px * calib / LATTICE_A))
    atoms = crystal_atoms(cells, cells, int(math.ceil(num_slices * dz / LATTICE_A)) + 1, seed=0)
    pot = projected_potential(atoms, object_px, calib, num_slices, dz)
    t_true = transmission(pot, kv)
    rng = np.random.default_rng(2)
    amp = 1.0 + rng.uniform(-0.05, 0.05, size=t_true.shape)
    t_obj = (t_true * amp).astype(np.complex64)
    span = object_px - n
    lin = np.linspace(0, span, scan)
    iy, ix = np.meshgrid(lin, lin, indexing="ij")
    pos = np.stack([iy.ravel(), ix.ravel()], axis=1) + rng.uniform(-0.2, 0.2, size=(scan * scan, 2))
    origins = np.floor(pos).astype(np.int32)
    shifts = pos - origins
    probe = probe_for(n, calib, kv, 20.0, 30.0, 0.0)
    return GradientWorkload(n, num_slices, dz, calib, kv, t_obj, t_true, probe, origins, shifts,
                            np.zeros((0, n, n)))


def poisson_measure(intensity: np.ndarray, dose: float = 1e5, seed: int = 2) -> np.ndarray:
    """Poisson counts at `dose` electrons per pattern, rescaled back to simulation units."""
    rng = np.random.default_rng(seed)
    total = intensity.sum(axis=(-2, -1), keepdims=True)
    counts = rng.poisson(dose * intensity / total)
    return counts * total / dose
