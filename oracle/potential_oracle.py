"""CPU oracle for the independent-atom potential builders.

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` may import this module, and only
as the checker; the product path never calls into it.

A float64 numpy restatement (paths relative to the reference root) of:

* ``multislice/atom_potentials.py:44-536``: ``bessel_kv``; the small-argument
  branch (x <= 2) of orders 0 and 1 and the large-argument branch (x > 2) are
  the polynomial approximations of Abramowitz & Stegun 9.8.5-9.8.8, the
  reference's evaluation; integer orders above 1 by forward recurrence from
  K0, K1 (small x); non-integer orders by the reflection formula with a
  20-term I_v series (small x) or the 5-term asymptotic series (large x);
  order 1/2 in closed form;
* ``multislice/form_factors.py:204-584``: Lobato and Kirkland form factors,
  projected potentials and the band-limited Lobato peak;
* ``multislice/atom_potentials.py:537-683``: ``single_atom_potential``
  (supersampled grid, average down-sampling) and its gradient in the centre
  (``projected_dr``, ``single_atom_center_grad``), which the reference takes
  with jax.grad;
* ``multislice/atom_potentials.py:684-1640``: ``crystal_potential_slices``
  (lattice repeats, z slicing, grid from the coordinate range plus padding,
  per-atom Fourier shift of the element's centred potential accumulated in
  float32, padding cropped).

Pins (``tests/test_potential_oracle.py``): the reference's captured Kirkland
C/Si/Au single-atom sums and maxima (``test_atom_potentials.py:32-64``) and
its published band-limited Lobato peaks (``test_form_factors.py:53-81``);
f_H(0) = a_0 and the Lobato Bethe normalisation; the A&S approximations
against scipy's K0/K1 within the published error bounds; both projected
potentials against the Hankel transform of their form factors; Lobato and
Kirkland agreement; the slice builder against integer-pixel rolls of the
centred potential and against the structure-factor sum; ``projected_dr``
and ``single_atom_center_grad`` against central differences of ``projected``
and ``single_atom_potential`` (the reference's own check of this gradient,
``test_atom_potentials.py:91-151``, on its Si case).
"""

from __future__ import annotations

import math

import numpy as np

A_BOHR = 0.529177210903
MOTT_BETHE = 47.87801
KIRKLAND_A0 = 0.5292
KIRKLAND_EK = 14.4

_K0_SMALL = (-0.57721566, 0.42278420, 0.23069756, 0.03488590, 0.00262698, 0.00010750, 0.00000740)
_K1_SMALL = (1.0, 0.15443144, -0.67278579, -0.18156897, -0.01919402, -0.00110404, -0.00004686)
_K0_LARGE = (1.25331414, -0.07832358, 0.02189568, -0.01062446, 0.00587872, -0.00251540, 0.00053208)
_K1_LARGE = (1.25331414, 0.23498619, -0.03655620, 0.01504268, -0.00780353, 0.00325614, -0.00068245)


def _series(x2, coeffs):
    return sum(c * x2 ** k for k, c in enumerate(coeffs))


def _horner(y, coeffs):
    acc = np.zeros_like(y) + coeffs[-1]
    for c in coeffs[-2::-1]:
        acc = c + y * acc
    return acc


def i0(x):
    x = np.asarray(x, np.float64)
    q = x * x / 4.0
    term, total = np.ones_like(x), np.ones_like(x)
    for k in range(1, 40):
        term = term * q / (k * k)
        total = total + term
    return total


def i1(x):
    x = np.asarray(x, np.float64)
    q = x * x / 4.0
    term = x / 2.0
    total = term.copy()
    for k in range(1, 40):
        term = term * q / (k * (k + 1))
        total = total + term
    return total


def k0_small(x):
    return -np.log(x / 2.0) * i0(x) + _series(x * x / 4.0, _K0_SMALL)


def k1_small(x):
    return np.log(x / 2.0) * i1(x) + _series(x * x / 4.0, _K1_SMALL) / x


def k0_large(x):
    return np.exp(-x) / np.sqrt(x) * _horner(2.0 / x, _K0_LARGE)


def k1_large(x):
    return np.exp(-x) / np.sqrt(x) * _horner(2.0 / x, _K1_LARGE)


def _iv_series(v, x):
    q = x * x / 4.0
    g1 = math.gamma(v + 1.0)
    total = np.zeros_like(x)
    for k in range(20):
        total = total + q ** k / (math.factorial(k) * math.gamma(k + v + 1.0) / g1)
    return (x / 2.0) ** v / g1 * total


def bessel_kv(v: float, x):
    """``atom_potentials.py:451-536`` for a scalar order ``v``."""
    x = np.asarray(x, np.float64)
    v = float(v)
    if abs(v - 0.5) < 1e-10:
        return np.sqrt(np.pi / (2.0 * x)) * np.exp(-x)
    vi = round(v)
    small = x <= 2.0
    xs = np.where(small, x, 1.0)
    xl = np.where(small, 3.0, x)
    if abs(v - vi) < 1e-10:
        n = abs(int(vi))
        k0, k1 = k0_small(xs), k1_small(xs)
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
        s_val = (math.pi / (2.0 * s)) * (_iv_series(-v, xs) - _iv_series(v, xs)) if abs(s) > 1e-10 \
            else np.zeros_like(xs)
    if abs(v) < 1e-10:
        l_val = k0_large(xl)
    elif abs(abs(v) - 1.0) < 1e-10:
        l_val = k1_large(xl)
    else:
        m = 4.0 * v * v
        a = (1.0, (m - 1.0) / 8.0, (m - 1.0) * (m - 9.0) / 128.0, (m - 1.0) * (m - 9.0) * (m - 25.0) / 3072.0,
             (m - 1.0) * (m - 9.0) * (m - 25.0) * (m - 49.0) / 98304.0)
        z = 1.0 / xl
        poly = a[0] + z * (a[1] + z * (a[2] + z * (a[3] + z * a[4])))
        l_val = np.sqrt(np.pi / (2.0 * xl)) * np.exp(-xl) * poly
    return np.where(small, s_val, l_val)


def lobato_split(row):
    row = np.asarray(row, np.float64)
    return row[0::2], row[1::2]


def kirkland_split(row):
    row = np.asarray(row, np.float64)
    amp, sc = row[0::2], row[1::2]
    return amp[:3], sc[:3], amp[3:], sc[3:]


def lobato_form_factor(row, q):
    a, b = lobato_split(row)
    g2 = (np.asarray(q, np.float64) / (2.0 * np.pi)) ** 2
    s = g2[..., None] * b
    return np.sum(a * (2.0 + s) / (1.0 + s) ** 2, axis=-1)


def kirkland_form_factor(row, q):
    la, lb, ga, gb = kirkland_split(row)
    g2 = ((np.asarray(q, np.float64) / (2.0 * np.pi)) ** 2)[..., None]
    return np.sum(la / (g2 + lb), axis=-1) + np.sum(ga * np.exp(-gb * g2), axis=-1)


def lobato_projected_potential(row, r, k0=None, k1=None):
    """Eq. 16 of Lobato & Van Dyck (2014) as ``form_factors.py:323-372``; ``k0``/``k1`` default to the A&S
    evaluation the reference uses."""
    k0 = k0 or (lambda t: bessel_kv(0.0, t))
    k1 = k1 or (lambda t: bessel_kv(1.0, t))
    a, b = lobato_split(row)
    rr = np.maximum(np.asarray(r, np.float64), 1e-10)[..., None]
    arg = 2.0 * np.pi * rr / np.sqrt(b)
    terms = a / b * k0(arg) + a * np.pi * rr / b ** 1.5 * k1(arg)
    return 2.0 * np.pi * MOTT_BETHE * np.sum(terms, axis=-1)


def kirkland_projected_potential(row, r, k0=None):
    """``form_factors.py:375-439`` (legacy 0.5292 * 14.4 normalisation, three-term sums in order)."""
    k0 = k0 or (lambda t: bessel_kv(0.0, t))
    la, lb, ga, gb = kirkland_split(row)
    rr = np.maximum(np.asarray(r, np.float64), 1e-10)[..., None]
    lt = la * k0(2.0 * np.pi * rr * np.sqrt(lb))
    gt = ga / gb * np.exp(-(np.pi ** 2) * rr * rr / gb)
    lp = 4.0 * np.pi ** 2 * KIRKLAND_A0 * KIRKLAND_EK
    gp = 2.0 * np.pi ** 2 * KIRKLAND_A0 * KIRKLAND_EK
    return lp * (lt[..., 0] + lt[..., 1] + lt[..., 2]) + gp * (gt[..., 0] + gt[..., 1] + gt[..., 2])


def projected(kind, row, r):
    return lobato_projected_potential(row, r) if kind == "lobato" else kirkland_projected_potential(row, r)


def form_factor(kind, row, q):
    return lobato_form_factor(row, q) if kind == "lobato" else kirkland_form_factor(row, q)


def lobato_bandlimited_peak(row, g_max):
    """``form_factors.py:526-584`` (Eqs. E.51-E.52)."""
    a, b = lobato_split(row)
    x = g_max * np.sqrt(b)
    anti = x - x / (2.0 * (1.0 + x * x)) - 0.5 * np.arctan(x)
    return 4.0 * np.pi * MOTT_BETHE * np.sum(a * b ** -1.5 * anti)


def single_atom_potential(kind, row, pixel_size, grid_shape, center=(0.0, 0.0), supersampling=4):
    """``atom_potentials.py:588-681``: evaluate on the supersampled grid, average ``supersampling``^2 blocks."""
    ss = int(supersampling)
    h, w = grid_shape
    step = pixel_size / ss
    gh, gw = h * ss, w * ss
    y = (np.arange(gh) - gh // 2) * step
    x = (np.arange(gw) - gw // 2) * step
    ya, xa = np.meshgrid(y, x, indexing="ij")
    r = np.sqrt((xa - center[0]) ** 2 + (ya - center[1]) ** 2 + 1e-10)
    v = projected(kind, row, r)
    return v.reshape(h, ss, w, ss).mean(axis=(1, 3))


def projected_dr(kind, row, r):
    """d/dr of ``projected`` by the Bessel identities K0'(x) = -K1(x) and (x K1(x))' = -x K0(x), applied to the A&S
    evaluations (the reference differentiates those evaluations with jax.grad; the two derivatives differ by the
    approximations' own error).  Zero below the 1e-10 radius floor."""
    r = np.asarray(r, np.float64)
    rr = np.maximum(r, 1e-10)[..., None]
    k0 = lambda t: bessel_kv(0.0, t)  # noqa: E731
    k1 = lambda t: bessel_kv(1.0, t)  # noqa: E731
    if kind == "lobato":
        a, b = lobato_split(row)
        sb = np.sqrt(b)
        arg = 2.0 * np.pi * rr / sb
        terms = -(a / b) * (2.0 * np.pi / sb) * k1(arg) - a * np.pi / b ** 1.5 * arg * k0(arg)
        d = 2.0 * np.pi * MOTT_BETHE * np.sum(terms, axis=-1)
    else:
        la, lb, ga, gb = kirkland_split(row)
        sb = np.sqrt(lb)
        lt = -la * 2.0 * np.pi * sb * k1(2.0 * np.pi * rr * sb)
        gt = -ga / gb * np.exp(-(np.pi ** 2) * rr * rr / gb) * (2.0 * np.pi ** 2 * rr / gb)
        d = (4.0 * np.pi ** 2 * KIRKLAND_A0 * KIRKLAND_EK * np.sum(lt, axis=-1)
             + 2.0 * np.pi ** 2 * KIRKLAND_A0 * KIRKLAND_EK * np.sum(gt, axis=-1))
    return np.where(r < 1e-10, 0.0, d)


def single_atom_center_grad(kind, row, pixel_size, grid_shape, center, supersampling, grad_out):
    """(dL/dx, dL/dy) of L = sum grad_out * single_atom_potential(..., center): the chain rule through
    r = sqrt((x - cx)^2 + (y - cy)^2 + 1e-10) and the block average of ``single_atom_potential``."""
    ss = int(supersampling)
    h, w = grid_shape
    step = pixel_size / ss
    gh, gw = h * ss, w * ss
    y = (np.arange(gh) - gh // 2) * step
    x = (np.arange(gw) - gw // 2) * step
    ya, xa = np.meshgrid(y, x, indexing="ij")
    dx, dy = xa - center[0], ya - center[1]
    r = np.sqrt(dx * dx + dy * dy + 1e-10)
    dv = projected_dr(kind, row, r) / r
    g = np.repeat(np.repeat(np.asarray(grad_out, np.float64), ss, axis=0), ss, axis=1) / (ss * ss)
    return np.array([np.sum(g * dv * -dx), np.sum(g * dv * -dy)])


def tile_repeats(positions, numbers, lattice, repeats):
    """``atom_potentials.py:1126-1241``: every lattice shift (ix, iy, iz) in C order, each copying all atoms."""
    positions = np.asarray(positions, np.float64)
    numbers = np.asarray(numbers, np.int64)
    nx, ny, nz = (int(v) for v in repeats)
    if max(nx, ny, nz) <= 1:
        return positions, numbers
    ii = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    shifts = ii.astype(np.float64) @ np.asarray(lattice, np.float64)
    return (positions[None] + shifts[:, None]).reshape(-1, 3), np.tile(numbers, shifts.shape[0])


def slice_atoms(positions, numbers, slice_thickness):
    """``atom_potentials.py:684-745``: slice index floor((z - z_min) / dz), atoms stably sorted by it."""
    z = positions[:, 2]
    idx = np.floor((z - z.min()) / slice_thickness)
    order = np.argsort(idx, kind="stable")
    return positions[order, 0], positions[order, 1], idx[order].astype(np.float32).astype(np.int64), numbers[order]


def grid_dimensions(x, y, padding, pixel_size, grid_hw=None):
    """``atom_potentials.py:748-823``: (x_min, y_min, width, height)."""
    x_min, y_min = x.min() - padding, y.min() - padding
    if grid_hw is not None:
        return x_min, y_min, grid_hw[1], grid_hw[0]
    width = int(np.ceil((x.max() + padding - x_min) / pixel_size))
    height = int(np.ceil((y.max() + padding - y_min) / pixel_size))
    crop = int(np.round(padding / pixel_size))
    return x_min, y_min, max(width, 2 * crop + 1), max(height, 2 * crop + 1)


def crystal_potential_slices(kind, rows, positions, numbers, lattice, pixel_size, slice_thickness=1.0,
                             repeats=(1, 1, 1), padding=4.0, supersampling=4, grid_shape=None,
                             accumulate=np.float32):
    """``atom_potentials.py:1364-1633``: (H, W, S) float64 slices (float32-accumulated values).

    ``rows`` maps atomic number -> coefficient row.  Each atom adds
    Re ifft2(fft2(V_Z) exp(2 pi i (kx sx + ky sy))) with V_Z the element's
    potential centred on the grid and (sx, sy) = pixel position - (W/2, H/2)."""
    pos, num = tile_repeats(positions, numbers, lattice, repeats)
    x, y, sidx, num = slice_atoms(pos, num, slice_thickness)
    gh = None if grid_shape is None else grid_shape[:2]
    x_min, y_min, w, h = grid_dimensions(x, y, padding, pixel_size, gh)
    n_slices = int(grid_shape[2]) if grid_shape is not None else int(sidx.max()) + 1
    kernels = {int(z): single_atom_potential(kind, rows[int(z)], pixel_size, (h, w), (0.0, 0.0), supersampling)
               for z in np.unique(num)}
    kern_k = {z: np.fft.fft2(v) for z, v in kernels.items()}
    ky = np.fft.fftfreq(h, d=1.0).reshape(-1, 1)
    kx = np.fft.fftfreq(w, d=1.0).reshape(1, -1)
    out = np.zeros((h, w, n_slices), accumulate)
    for s in range(n_slices):
        acc = np.zeros((h, w), accumulate)
        for a in np.nonzero(sidx == s)[0]:
            sx = (x[a] - x_min) / pixel_size - w / 2.0
            sy = (y[a] - y_min) / pixel_size - h / 2.0
            phase = np.exp(2j * np.pi * (kx * sx + ky * sy))
            acc = (acc + np.real(np.fft.ifft2(kern_k[int(num[a])] * phase)).astype(accumulate)).astype(accumulate)
        out[:, :, s] = acc
    crop = round(float(padding) / float(pixel_size))
    return out[crop:h - crop, crop:w - crop, :].astype(np.float64), (x_min, y_min, w, h)


def crystal_potential_slices_position_grad(kind, rows, positions, numbers, lattice, pixel_size, grad_slices,
                                           slice_thickness=1.0, repeats=(1, 1, 1), padding=4.0, supersampling=4,
                                           grid_shape=None):
    """(N, 3) dL/dpositions of L = sum grad_slices * crystal_potential_slices(...), the gradient jax.grad takes
    through ``atom_potentials.py:1364-1633``.  Each atom enters through its Fourier shift
    (sx, sy) = ((x - x_min) / pixel, (y - y_min) / pixel) - (W/2, H/2):
        dL/dsx_a = Re sum_k 2 pi i kx Khat(k) A_s(k) exp(2 pi i (kx sx_a + ky sy_a)),  A_s = ifft2(padded dL/dV_s),
    and x_min = min(x) - padding passes -sum_a dL/dsx_a / pixel to the atoms at the minimum, split evenly between
    ties (jax's min rule); lattice copies add into their source atom; z enters only through the slice index
    (piecewise constant), so its column is zero."""
    positions = np.asarray(positions, np.float64)
    n_in = positions.shape[0]
    pos, num = tile_repeats(positions, numbers, lattice, repeats)
    z = pos[:, 2]
    idx = np.floor((z - z.min()) / float(slice_thickness)).astype(np.int64)
    order = np.argsort(idx, kind="stable")
    x, y, sidx, num = pos[order, 0], pos[order, 1], idx[order], np.asarray(num)[order]
    gh = None if grid_shape is None else grid_shape[:2]
    x_min, y_min, w, h = grid_dimensions(x, y, padding, pixel_size, gh)
    n_slices = int(grid_shape[2]) if grid_shape is not None else int(sidx.max()) + 1
    kern_k = {int(zz): np.fft.fft2(single_atom_potential(kind, rows[int(zz)], pixel_size, (h, w), (0.0, 0.0),
                                                         supersampling)) for zz in np.unique(num)}
    ky = np.fft.fftfreq(h, d=1.0).reshape(-1, 1)
    kx = np.fft.fftfreq(w, d=1.0).reshape(1, -1)
    crop = round(float(padding) / float(pixel_size))
    gsx = np.zeros(x.shape[0])
    gsy = np.zeros(x.shape[0])
    for s in range(n_slices):
        pad = np.zeros((h, w))
        pad[crop:h - crop, crop:w - crop] = np.asarray(grad_slices, np.float64)[:, :, s]
        adj = np.fft.ifft2(pad)
        for a in np.nonzero(sidx == s)[0]:
            sx = (x[a] - x_min) / pixel_size - w / 2.0
            sy = (y[a] - y_min) / pixel_size - h / 2.0
            term = 2j * np.pi * kern_k[int(num[a])] * adj * np.exp(2j * np.pi * (kx * sx + ky * sy))
            gsx[a] = np.real(np.sum(term * kx))
            gsy[a] = np.real(np.sum(term * ky))
    gx, gy = gsx / pixel_size, gsy / pixel_size
    at_min_x = x == x.min()
    at_min_y = y == y.min()
    gx = gx - at_min_x * gx.sum() / at_min_x.sum()
    gy = gy - at_min_y * gy.sum() / at_min_y.sum()
    tiled = np.zeros((x.shape[0], 3))
    tiled[order, 0] = gx
    tiled[order, 1] = gy
    return tiled.reshape(-1, n_in, 3).sum(axis=0)
