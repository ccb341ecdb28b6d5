"""Projected-potential slices of a crystal, built on the GPU.

Mirrors ``src/ptyrodactyl/multislice/atom_potentials.py``: ``bessel_kv``
(``:451-536``), ``single_atom_potential`` (``:588-681``) and
``crystal_potential_slices`` (``:1364-1633``) with the same arguments,
defaults (Lobato parameterisation, supersampling 4, padding 4 A) and grid
conventions.

The reference adds each atom to its slice by a Fourier shift of the
element's centred potential (one FFT pair per atom per slice).  Here the
element potentials are evaluated once by ``pty_atom_potential_grid``, their
spectra Khat_t taken once, and every slice is formed in Fourier space by
``pty_potential_slices_fourier`` as sum_t Khat_t sum_a exp(2 pi i k.s_a),
followed by one inverse FFT per slice; the two are equal by linearity.  The
slices are accumulated in float64 (the reference accumulates in float32).

Kept from the reference: atoms land at the point reflection 2c - p of their
pixel position p about the grid centre c = (W/2, H/2), because the shift
phase is exp(+2 pi i k.(p - c)) (``atom_potentials.py:950-955``); the
supersampled grid's origin is sample (H*ss/2, W*ss/2), so the block average
puts an element's centre (ss-1)/(2 ss) pixel before the pixel centre.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np
import torch

from . import _native
from .form_factors import _CODES, _check_parameterization, _gpu, coefficient_row
from .types import CrystalData, PotentialSlices, _as_tensor, create_potential_slices

default_repeats = (1, 1, 1)
_FOURIER_BUDGET_BYTES = 2 << 30


# ---------------------------------------------------------------------------
# bessel_kv
# ---------------------------------------------------------------------------

def _series_i(x: torch.Tensor, order: int) -> torch.Tensor:
    q = x * x / 4.0
    term = torch.ones_like(x) if order == 0 else x / 2.0
    total = term.clone()
    for k in range(1, 40):
        term = term * q / (k * (k + order))
        total = total + term
    return total


def _poly(y: torch.Tensor, coeffs) -> torch.Tensor:
    acc = torch.full_like(y, coeffs[-1])
    for c in coeffs[-2::-1]:
        acc = c + y * acc
    return acc


_K0_SMALL = (-0.57721566, 0.42278420, 0.23069756, 0.03488590, 0.00262698, 0.00010750, 0.00000740)
_K1_SMALL = (1.0, 0.15443144, -0.67278579, -0.18156897, -0.01919402, -0.00110404, -0.00004686)
_K0_LARGE = (1.25331414, -0.07832358, 0.02189568, -0.01062446, 0.00587872, -0.00251540, 0.00053208)
_K1_LARGE = (1.25331414, 0.23498619, -0.03655620, 0.01504268, -0.00780353, 0.00325614, -0.00068245)


def bessel_kv(v, x) -> torch.Tensor:
    """Modified Bessel function of the second kind K_v(x), x > 0, evaluated as the reference does: orders 0 and 1
    by Abramowitz & Stegun 9.8.5-9.8.8, higher integer orders by forward recurrence (x <= 2) or the 5-term
    asymptotic series (x > 2), non-integer orders by the reflection formula with a 20-term I_v series (x <= 2) or
    the asymptotic series, order 1/2 in closed form.  Runs on the GPU."""
    v = float(v)
    x = _as_tensor(x, torch.float64)
    x = x.to(_gpu(x.device))
    if abs(v - 0.5) < 1e-10:
        return torch.sqrt(math.pi / (2.0 * x)) * torch.exp(-x)
    small = x <= 2.0
    xs = torch.where(small, x, torch.ones_like(x))
    xl = torch.where(small, torch.full_like(x, 3.0), x)
    vi = round(v)
    if abs(v - vi) < 1e-10:
        k0 = -torch.log(xs / 2.0) * _series_i(xs, 0) + _poly(xs * xs / 4.0, _K0_SMALL)
        k1 = torch.log(xs / 2.0) * _series_i(xs, 1) + _poly(xs * xs / 4.0, _K1_SMALL) / xs
        n = abs(int(vi))
        if n == 0:
            s_val = k0
        elif n == 1:
            s_val = k1
        else:
            km, kc = k0, k1
            for i in range(1, n):
                km, kc = kc, 2.0 * i / xs * kc + km
            s_val = kc
    else:
        s = math.sin(math.pi * v)
        if abs(s) > 1e-10:
            q = xs * xs / 4.0

            def iv(order):
                g1 = math.gamma(order + 1.0)
                tot = torch.zeros_like(xs)
                for k in range(20):
                    tot = tot + q ** k / (math.factorial(k) * math.gamma(k + order + 1.0) / g1)
                return (xs / 2.0) ** order / g1 * tot

            s_val = (math.pi / (2.0 * s)) * (iv(-v) - iv(v))
        else:
            s_val = torch.zeros_like(xs)
    if abs(v) < 1e-10:
        l_val = torch.exp(-xl) / torch.sqrt(xl) * _poly(2.0 / xl, _K0_LARGE)
    elif abs(abs(v) - 1.0) < 1e-10:
        l_val = torch.exp(-xl) / torch.sqrt(xl) * _poly(2.0 / xl, _K1_LARGE)
    else:
        m = 4.0 * v * v
        a = (1.0, (m - 1.0) / 8.0, (m - 1.0) * (m - 9.0) / 128.0, (m - 1.0) * (m - 9.0) * (m - 25.0) / 3072.0,
             (m - 1.0) * (m - 9.0) * (m - 25.0) * (m - 49.0) / 98304.0)
        l_val = torch.sqrt(math.pi / (2.0 * xl)) * torch.exp(-xl) * _poly(1.0 / xl, a)
    return torch.where(small, s_val, l_val)


# ---------------------------------------------------------------------------
# Element potentials on a grid (HIP)
# ---------------------------------------------------------------------------

def _atom_grids(rows: np.ndarray, parameterization: str, pixel_size: float, height: int, width: int,
                supersampling: int, centers: Optional[torch.Tensor], device) -> torch.Tensor:
    rows_t = torch.as_tensor(rows, dtype=torch.float64).reshape(-1, 12).to(device).contiguous()
    out = torch.empty((rows_t.shape[0], height, width), dtype=torch.float64, device=device)
    ctr = None if centers is None else centers.to(device=device, dtype=torch.float64).contiguous()
    lib = _native.load()
    with torch.cuda.device(device):
        status = lib.pty_atom_potential_grid(_CODES[parameterization], rows_t.shape[0], _native.ptr(rows_t),
                                             float(pixel_size), int(height), int(width), int(supersampling),
                                             _native.ptr(ctr), _native.ptr(out), _native.stream_handle(device))
    _native.check(status, "pty_atom_potential_grid")
    return out


def _atom_center_grad(rows_t: torch.Tensor, parameterization: str, pixel_size: float, height: int, width: int,
                      supersampling: int, centers: torch.Tensor, grad_out: torch.Tensor) -> torch.Tensor:
    """(T, 2) float64 d(sum grad_out * grids)/d(centers) on the GPU (``pty_atom_potential_grid_center_grad``)."""
    dev = grad_out.device
    lib = _native.load()
    t = rows_t.shape[0]
    ws = torch.empty(max(1, int(lib.pty_atom_center_grad_workspace_bytes(t))), dtype=torch.uint8, device=dev)
    out = torch.empty((t, 2), dtype=torch.float64, device=dev)
    g = grad_out.to(torch.float64).contiguous()
    with torch.cuda.device(dev):
        status = lib.pty_atom_potential_grid_center_grad(_CODES[parameterization], t, _native.ptr(rows_t),
                                                         float(pixel_size), int(height), int(width),
                                                         int(supersampling), _native.ptr(centers), _native.ptr(g),
                                                         _native.ptr(out), _native.ptr(ws), ws.numel(),
                                                         _native.stream_handle(dev))
    _native.check(status, "pty_atom_potential_grid_center_grad")
    return out


class _AtomGridFunction(torch.autograd.Function):
    """Element potentials on the grid as a function of their centres: forward ``pty_atom_potential_grid``,
    backward ``pty_atom_potential_grid_center_grad`` (the reference differentiates ``single_atom_potential`` in
    ``center_coords`` with jax.grad, ``atom_potentials.py:588-681``)."""

    @staticmethod
    def forward(centers, rows, parameterization, pixel_size, height, width, supersampling):
        return _atom_grids(rows, parameterization, pixel_size, height, width, supersampling, centers, centers.device)

    @staticmethod
    def setup_context(ctx, inputs, output):
        centers, rows, parameterization, pixel_size, height, width, supersampling = inputs
        ctx.save_for_backward(centers.detach())
        ctx.args = (rows, parameterization, pixel_size, height, width, supersampling)

    @staticmethod
    def backward(ctx, grad_out):
        (centers,) = ctx.saved_tensors
        rows, parameterization, pixel_size, height, width, supersampling = ctx.args
        rows_t = torch.as_tensor(rows, dtype=torch.float64).reshape(-1, 12).to(grad_out.device).contiguous()
        g = _atom_center_grad(rows_t, parameterization, pixel_size, height, width, supersampling,
                              centers.to(grad_out.device, torch.float64).contiguous(), grad_out)
        return g.to(centers.dtype), None, None, None, None, None, None


def single_atom_potential(atom_no, pixel_size, grid_shape: Tuple[int, int], center_coords=None,
                          supersampling: int = 4, *, parameterization: str = "lobato") -> torch.Tensor:
    """(H, W) float64 projected potential (V A) of one atom: the supersampled grid (step pixel_size /
    supersampling, origin at sample (H ss/2, W ss/2)) averaged over supersampling^2 blocks; centred at
    ``center_coords`` = (x, y) A, default the origin.  Differentiable in ``center_coords``."""
    _check_parameterization(parameterization)
    h, w = (int(v) for v in grid_shape)
    dev = _gpu(center_coords.device if isinstance(center_coords, torch.Tensor) else None)
    row = coefficient_row(atom_no, parameterization)
    if isinstance(center_coords, torch.Tensor) and center_coords.requires_grad:
        ctr = center_coords.to(device=dev, dtype=torch.float64).reshape(1, 2)
        return _AtomGridFunction.apply(ctr, row, parameterization, float(pixel_size), h, w, int(supersampling))[0]
    ctr = None if center_coords is None else _as_tensor(center_coords, torch.float64).reshape(1, 2)
    return _atom_grids(row, parameterization, float(pixel_size), h, w, int(supersampling), ctr, dev)[0]


# ---------------------------------------------------------------------------
# crystal_potential_slices
# ---------------------------------------------------------------------------

def _apply_repeats(positions: torch.Tensor, numbers: torch.Tensor, lattice: torch.Tensor, repeats):
    """All lattice shifts (ix, iy, iz) in C order, each copying every atom (``atom_potentials.py:1126-1241``)."""
    reps = [int(r) for r in (repeats.tolist() if hasattr(repeats, "tolist") else repeats)]
    if len(reps) != 3:
        raise ValueError("repeats must have three entries")
    if not any(r > 1 for r in reps):
        return positions, numbers
    dev = positions.device
    grids = torch.meshgrid(*[torch.arange(r, device=dev) for r in reps], indexing="ij")
    idx = torch.stack([g.reshape(-1) for g in grids], dim=-1).to(torch.float64)
    shifts = idx @ lattice.to(dev)
    return (positions[None] + shifts[:, None]).reshape(-1, 3), numbers.repeat(shifts.shape[0])


def _slice_atoms(coords: torch.Tensor, atom_numbers: torch.Tensor, slice_thickness) -> torch.Tensor:
    """(N, 4) [x, y, slice index, atomic number] sorted stably by slice index floor((z - z_min) / dz)
    (``atom_potentials.py:684-745``)."""
    z = coords[:, 2]
    idx = torch.floor((z - z.min()) / float(slice_thickness))
    order = torch.argsort(idx, stable=True)
    out = torch.stack([coords[:, 0], coords[:, 1], idx.to(torch.float32).to(torch.float64),
                       atom_numbers.to(torch.float64)], dim=1)
    return out[order]


def _grid_dimensions(x: torch.Tensor, y: torch.Tensor, padding: float, pixel_size: float,
                     grid_hw: Optional[Tuple[int, int]]):
    """(x_min, y_min, width, height) (``atom_potentials.py:748-823``)."""
    x, y = x.detach(), y.detach()
    x_min = float(x.min()) - padding
    y_min = float(y.min()) - padding
    if grid_hw is not None:
        return x_min, y_min, int(grid_hw[1]), int(grid_hw[0])
    width = int(math.ceil(((float(x.max()) + padding) - x_min) / pixel_size))
    height = int(math.ceil(((float(y.max()) + padding) - y_min) / pixel_size))
    crop = int(round(padding / pixel_size))
    return x_min, y_min, max(width, 2 * crop + 1), max(height, 2 * crop + 1)


@dataclass(frozen=True)
class SlicePlan:
    """Host-side layout of a crystal's slices: grid, elements, and atoms sorted by (slice, element) with their
    pixel shifts from the grid centre and the segment offsets of every (slice, element) pair."""

    x_min: float
    y_min: float
    width: int
    height: int
    num_slices: int
    crop: int
    elements: torch.Tensor   # (T,) sorted atomic numbers
    shifts: torch.Tensor     # (N, 2) float64 (sx, sy) pixels
    segments: torch.Tensor   # (S*T + 1,) int64


def plan_slices(crystal_data: CrystalData, pixel_size, slice_thickness=1.0, repeats=default_repeats, padding=4.0,
                grid_shape: Optional[Tuple[int, int, int]] = None, device=None) -> SlicePlan:
    """Repeats, z slicing, grid dimensions and the (slice, element) atom order of ``crystal_potential_slices``
    (``atom_potentials.py:1364-1633``), on ``device``."""
    if crystal_data.lattice is None:
        raise ValueError("CrystalData.lattice must not be None")
    pixel, dz, pad = float(pixel_size), float(slice_thickness), float(padding)
    dev = crystal_data.positions.device if device is None else torch.device(device)
    pos = crystal_data.positions.to(dev, torch.float64)
    nums = crystal_data.atomic_numbers.to(dev, torch.int64)
    pos, nums = _apply_repeats(pos, nums, crystal_data.lattice.to(dev, torch.float64), repeats)
    sliced = _slice_atoms(pos, nums, dz)
    x, y = sliced[:, 0], sliced[:, 1]
    sidx = sliced[:, 2].to(torch.int64)
    anum = sliced[:, 3].to(torch.int64)
    x_min, y_min, width, height = _grid_dimensions(x, y, pad, pixel, None if grid_shape is None
                                                   else (grid_shape[0], grid_shape[1]))
    if pos.requires_grad:
        # the grid origin follows the lowest atom, as the reference's x_min = min(x) - padding does under
        # jax.grad (amin splits the gradient evenly between tied minima, as jax's min does)
        x_min_t, y_min_t = torch.amin(x) - pad, torch.amin(y) - pad
    else:
        x_min_t, y_min_t = x_min, y_min
    n_slices = int(grid_shape[2]) if grid_shape is not None else int(sidx.max()) + 1
    elements = torch.unique(anum)
    n_types = elements.numel()
    tidx = torch.searchsorted(elements, anum)
    keep = sidx < n_slices
    key = (sidx * n_types + tidx)[keep]
    order = torch.argsort(key, stable=True)
    sx = ((x - x_min_t) / pixel - width / 2.0)[keep][order]
    sy = ((y - y_min_t) / pixel - height / 2.0)[keep][order]
    counts = torch.bincount(key, minlength=n_slices * n_types)
    segments = torch.zeros(n_slices * n_types + 1, dtype=torch.int64, device=dev)
    segments[1:] = torch.cumsum(counts, 0)
    crop = int(round(pad / pixel))
    if height - 2 * crop < 1 or width - 2 * crop < 1:
        raise ValueError("padding removes the whole grid")
    return SlicePlan(x_min, y_min, width, height, n_slices, crop, elements,
                     torch.stack([sx, sy], dim=1).contiguous(), segments)


def _slices_from_plan(plan: SlicePlan, shifts: torch.Tensor, kern_k: torch.Tensor, n_types: int) -> torch.Tensor:
    """(H', W', S) float64 slices of ``plan`` with the given atom shifts (pixels): ``pty_potential_slices_fourier``
    per chunk of slices, then the inverse FFT and the crop."""
    h, w, n_slices, crop = plan.height, plan.width, plan.num_slices, plan.crop
    dev = kern_k.device
    out_h, out_w = h - 2 * crop, w - 2 * crop
    result = torch.empty((out_h, out_w, n_slices), dtype=torch.float64, device=dev)
    chunk = max(1, min(n_slices, _FOURIER_BUDGET_BYTES // (16 * h * w)))
    lib = _native.load()
    for s0 in range(0, n_slices, chunk):
        ns = min(chunk, n_slices - s0)
        seg = plan.segments[s0 * n_types:(s0 + ns) * n_types + 1].contiguous()
        fk = torch.empty((ns, h, w), dtype=torch.complex128, device=dev)
        with torch.cuda.device(dev):
            status = lib.pty_potential_slices_fourier(ns, n_types, h, w, shifts.shape[0], _native.ptr(shifts),
                                                      _native.ptr(seg), _native.ptr(kern_k), _native.ptr(fk),
                                                      _native.stream_handle(dev))
        _native.check(status, "pty_potential_slices_fourier")
        real = torch.fft.ifft2(fk).real
        result[:, :, s0:s0 + ns] = real[:, crop:crop + out_h, crop:crop + out_w].permute(1, 2, 0)
    return result


class _SlicesFunction(torch.autograd.Function):
    """Crystal slices as a function of the atoms' pixel shifts.  Backward: A_s = ifft2 of the padded slice
    gradient, then ``pty_potential_slices_shift_grad`` (dL/dsx, dL/dsy per atom) — the gradient the reference
    gets from jax.grad through its per-atom Fourier shifts (``atom_potentials.py:826-993``)."""

    @staticmethod
    def forward(shifts, plan, kern_k, n_types):
        return _slices_from_plan(plan, shifts.detach().contiguous(), kern_k, n_types)

    @staticmethod
    def setup_context(ctx, inputs, output):
        shifts, plan, kern_k, n_types = inputs
        ctx.save_for_backward(shifts.detach())
        ctx.plan, ctx.kern_k, ctx.n_types = plan, kern_k, n_types

    @staticmethod
    def backward(ctx, grad_out):
        (shifts,) = ctx.saved_tensors
        plan, kern_k, n_types = ctx.plan, ctx.kern_k, ctx.n_types
        h, w, n_slices, crop = plan.height, plan.width, plan.num_slices, plan.crop
        dev = kern_k.device
        shifts = shifts.contiguous()
        n_atoms = shifts.shape[0]
        grad = torch.zeros((n_atoms, 2), dtype=torch.float64, device=dev)
        if n_atoms == 0:
            return grad, None, None, None
        lib = _native.load()
        ws = torch.empty(max(1, int(lib.pty_potential_shift_grad_workspace_bytes(h, w, n_atoms))), dtype=torch.uint8,
                         device=dev)
        g = grad_out.to(device=dev, dtype=torch.float64)
        out_h, out_w = h - 2 * crop, w - 2 * crop
        chunk = max(1, min(n_slices, _FOURIER_BUDGET_BYTES // (16 * h * w)))
        for s0 in range(0, n_slices, chunk):
            ns = min(chunk, n_slices - s0)
            pad = torch.zeros((ns, h, w), dtype=torch.float64, device=dev)
            pad[:, crop:crop + out_h, crop:crop + out_w] = g[:, :, s0:s0 + ns].permute(2, 0, 1)
            adj = torch.fft.ifft2(pad).contiguous()
            seg = plan.segments[s0 * n_types:(s0 + ns) * n_types + 1].contiguous()
            with torch.cuda.device(dev):
                status = lib.pty_potential_slices_shift_grad(ns, n_types, h, w, n_atoms, _native.ptr(shifts),
                                                             _native.ptr(seg), _native.ptr(kern_k), _native.ptr(adj),
                                                             _native.ptr(grad), _native.ptr(ws), ws.numel(),
                                                             _native.stream_handle(dev))
            _native.check(status, "pty_potential_slices_shift_grad")
        return grad.to(shifts.dtype), None, None, None


def crystal_potential_slices(crystal_data: CrystalData, pixel_size, slice_thickness=1.0, repeats=default_repeats,
                             padding=4.0, supersampling: int = 4, grid_shape: Optional[Tuple[int, int, int]] = None,
                             *, parameterization: str = "lobato") -> PotentialSlices:
    """Projected-potential slices (H, W, S) float64 of a crystal: lattice repeats, atoms binned into slices of
    ``slice_thickness`` from the lowest z, a grid spanning the atoms plus ``padding`` on each side (or
    ``grid_shape`` = (H, W, S) before cropping), every atom's element potential placed by a Fourier shift, and
    round(padding / pixel_size) pixels cropped from each edge."""
    _check_parameterization(parameterization)
    dev = _gpu(crystal_data.positions.device)
    plan = plan_slices(crystal_data, pixel_size, slice_thickness, repeats, padding, grid_shape, dev)
    pixel, h, w, n_slices = float(pixel_size), plan.height, plan.width, plan.num_slices
    rows = np.stack([coefficient_row(int(z), parameterization) for z in plan.elements.tolist()])
    n_types = rows.shape[0]
    kern_k = torch.fft.fft2(_atom_grids(rows, parameterization, pixel, h, w, int(supersampling), None, dev))
    kern_k = kern_k.contiguous()
    if plan.shifts.requires_grad:  # differentiable in the atom positions
        result = _SlicesFunction.apply(plan.shifts, plan, kern_k, n_types)
    else:
        result = _slices_from_plan(plan, plan.shifts, kern_k, n_types)
    return create_potential_slices(result, float(slice_thickness), pixel)


__all__ = ["SlicePlan", "bessel_kv", "crystal_potential_slices", "default_repeats", "plan_slices",
           "single_atom_potential"]
