"""CPU oracle for the multislice 4D-STEM forward model and its gradient.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import this module, and only as
the checker (or the timed CPU baseline).  The product path in
``ptyrodactyl_amd`` never calls into it.

Everything here is a complex128 / float64 numpy restatement of the
reference's algorithm, written from its documented behaviour; each function
cites the reference file:line it follows (paths relative to the reference
repository root).

Parity pins (see ``tests/test_oracle.py``):

* ``wavelength_ang`` / ``interaction_sigma``: published values from
  ``tests/test_ptyrodactyl/test_types/test_constants.py:173-239``.
* ``make_probe`` / ``cbed_image`` / ``stem_4d``: the reference's golden
  capture ``tests/test_data/plan01_ug_reference.npz`` (arrays ``cbed`` and
  ``stem4d``), regenerated from the capture recipe
  ``tests/test_data/plan01_ug_capture.py:41-70`` with the Threefry-2x32
  restatement below; plus closed forms (V=0 gives the aperture indicator,
  Parseval unitarity, S=1 reduces to |FFT(psi T)|^2, brute-force O(N^4) DFT
  on tiny grids, translation equivariance).
* ``multislice_vjp`` (the gradient): central finite differences in both
  real and imaginary directions, an independent torch autograd twin, and
  the adjoint dot-product identity; its shift gradient against central
  differences of ``window_stem`` in the shifts and against the closed form
  for a vacuum slab (intensity is shift-invariant, so the gradient is 0
  there while the probe gradient is not).
* ``window_stem``: equals ``stem_4d`` when the slab is one window tiled
  periodically; equals ``stem_4d`` on an explicitly cropped window.
* ``aberration12`` / ``probe_from_aberrations``: reduces to the reference
  ``aberration`` (C1/C3 terms) when only C1 and C3 are set; every other
  coefficient by the properties its name fixes (radial order 2/3/4, m-fold
  rotational symmetry with a sign flip under a rotation by pi/m, the
  Krivanek-form prefactor of its maximum, and the orientation of its real and
  imaginary parts).
* ``lm_step_dense`` / ``lm_gain_ratio`` / ``lm_update_damping`` and the loop of
  ``lm_aberration_fit``: the closed forms of
  ``tests/test_ptyrodactyl/test_jacobian/test_solvers.py:53-226`` (GN step 1,
  LM step 4/(4+lambda), damping/3 on an accepted quadratic step, frozen state
  when already converged), and the fit recovers known coefficients.
"""

from __future__ import annotations

import numpy as np

try:  # pocketfft with a worker pool when scipy is present
    import scipy.fft as _fft

    def _fft2(a, axes=(-2, -1), workers=1):
        return _fft.fft2(a, axes=axes, workers=workers)

    def _ifft2(a, axes=(-2, -1), workers=1):
        return _fft.ifft2(a, axes=axes, workers=workers)

except ImportError:  # pragma: no cover - numpy fallback for the checker only

    def _fft2(a, axes=(-2, -1), workers=1):
        return np.fft.fft2(a, axes=axes)

    def _ifft2(a, axes=(-2, -1), workers=1):
        return np.fft.ifft2(a, axes=axes)


# CODATA 2018 constants, as in types/constants.py:62-90.
H_PLANCK = 6.62607015e-34
M_E = 9.1093837015e-31
E_CHARGE = 1.602176634e-19
C_LIGHT = 2.99792458e8


def wavelength_ang(voltage_kv: float) -> float:
    """Relativistic electron wavelength in Angstrom.

    Follows ``types/constants.py:289-346``:
    lambda = h c / sqrt(eV (2 m c^2 + eV)) * 1e10.
    """
    ev = float(voltage_kv) * 1000.0 * E_CHARGE
    num = H_PLANCK**2 * C_LIGHT**2
    den = ev * (2.0 * M_E * C_LIGHT**2 + ev)
    return 1e10 * np.sqrt(num / den)


def interaction_sigma(voltage_kv: float) -> float:
    """Interaction constant sigma in rad/(V*Angstrom).

    Follows ``multislice/simulations.py:139-157`` (transmission_func):
    sigma = (2 pi / (lambda V)) (m c^2 + eV) / (2 m c^2 + eV).
    """
    voltage = float(voltage_kv) * 1000.0
    ev = E_CHARGE * voltage
    lam = wavelength_ang(voltage_kv)
    mc2 = M_E * C_LIGHT**2
    return ((2.0 * np.pi / (lam * voltage)) * (mc2 + ev)) / (2.0 * mc2 + ev)


def transmission_func(pot_slice, voltage_kv: float):
    """T = exp(i sigma V); ``multislice/simulations.py:105-159``."""
    return np.exp(1j * interaction_sigma(voltage_kv) * np.asarray(pot_slice, np.float64))


def fftfreq(n: int, d: float):
    """numpy/jax ``fftfreq``: Nyquist bin is negative for even n."""
    return np.fft.fftfreq(int(n), d=float(d))


def propagation_func(imsize_y: int, imsize_x: int, thickness_ang: float,
                     voltage_kv: float, calib_ang: float):
    """Fresnel propagator exp(-i pi lambda dz |q|^2) on the unshifted grid.

    Follows ``multislice/simulations.py:162-225`` (meshgrid ij-indexing).
    """
    qy = fftfreq(imsize_y, calib_ang)
    qx = fftfreq(imsize_x, calib_ang)
    lya, lxa = np.meshgrid(qy, qx, indexing="ij")
    l_sq = lxa**2 + lya**2
    lam = wavelength_ang(voltage_kv)
    return np.exp(-1j * np.pi * lam * float(thickness_ang) * l_sq)


def aberration(fourier_coord, lambda_angstrom: float, defocus: float = 0.0,
               c3: float = 0.0, c5: float = 0.0):
    """chi(q) = 2pi/lambda (C1 p^2/2 + C3 1e7 p^4/4 + C5 1e7 p^6/6), p = lambda q.

    Follows ``multislice/simulations.py:434-498`` including the 1e7 factor
    on C3/C5 (their unit is effectively mm).
    """
    p = lambda_angstrom * np.asarray(fourier_coord, np.float64)
    chi = (defocus * p**2) / 2 + (c3 * 1e7 * p**4) / 4 + (c5 * 1e7 * p**6) / 6
    return (2 * np.pi * chi) / lambda_angstrom


def probe_fourier_grid(image_y: int, image_x: int, probe_calibration_pm: float):
    """|q| grid used by make_probe, ``simulations.py:400-413``.

    q = roll(arange(-n/2, n/2)/fov, n//2) with fov = n * 0.01 * calib_pm;
    equal to fftfreq for even n.
    """
    x_fov = image_x * 0.01 * probe_calibration_pm
    y_fov = image_y * 0.01 * probe_calibration_pm
    qx = np.arange(-image_x / 2, image_x / 2, 1.0) / x_fov
    qy = np.arange(-image_y / 2, image_y / 2, 1.0) / y_fov
    lx = np.roll(qx, image_x // 2)
    ly = np.roll(qy, image_y // 2)
    lya, lxa = np.meshgrid(lx, ly)  # xy-indexing, as in the reference
    return np.sqrt(lxa * lxa + lya * lya)


def make_probe(voltage_kv: float, aperture_mrad: float, defocus_ang: float,
               c3_ang: float, c5_ang: float, probe_shape, probe_calibration_pm: float):
    """Real-space probe ifftshift(ifft2(A exp(-i chi))).

    Follows ``multislice/simulations.py:334-431``: hard aperture ``|q| <= alpha/lambda``.
    """
    image_y, image_x = int(probe_shape[0]), int(probe_shape[1])
    lam = wavelength_ang(voltage_kv)
    l_max = (aperture_mrad / 1000.0) / lam
    q = probe_fourier_grid(image_y, image_x, probe_calibration_pm)
    a_dist = (q <= l_max).astype(np.complex128)
    chi = aberration(q, lam, defocus_ang, c3_ang, c5_ang)
    a_dist = a_dist * np.exp(-1j * chi)
    return np.fft.ifftshift(np.fft.ifft2(a_dist))


def shift_beam_fourier(beam, pos, calib_ang: float):
    """Shift beam modes by ``pos`` (Angstrom, (y, x)) with a Fourier phase ramp.

    Follows ``multislice/simulations.py:1236-1327``.  ``beam`` is (H, W[, M]);
    returns (P, H, W, M) complex128.
    """
    b = np.asarray(beam, np.complex128)
    if b.ndim == 2:
        b = b[..., None]
    pos = np.atleast_2d(np.asarray(pos, np.float64))
    hh, ww = b.shape[0], b.shape[1]
    qy = fftfreq(hh, calib_ang)
    qx = fftfreq(ww, calib_ang)
    qya, qxa = np.meshgrid(qy, qx, indexing="ij")
    beam_k = np.fft.fft2(b, axes=(0, 1))
    out = np.empty((pos.shape[0], hh, ww, b.shape[2]), np.complex128)
    for i, (ys, xs) in enumerate(pos):
        phase = -2.0 * np.pi * (qya * ys + qxa * xs)
        out[i] = np.fft.ifft2(beam_k * np.exp(1j * phase)[..., None], axes=(0, 1))
    return out


def multislice_amplitude(trans, wave0, prop, workers: int = 1):
    """Exit-plane detector amplitude for precomputed transmissions.

    The scan body of ``_cbed_amplitude_from_slice_provider``
    (``multislice/simulations.py:501-603``): for each slice multiply by T;
    for every slice except the last apply ifft2(fft2(.) * P).  Result is
    fftshift(fft2(wave)).

    trans: (S, H, W) complex; wave0: (..., H, W) complex (leading batch axes
    allowed); prop: (H, W) complex.
    """
    wave = np.array(wave0, dtype=np.result_type(wave0, np.complex64), copy=True)
    s_count = trans.shape[0]
    for s in range(s_count):
        wave = wave * trans[s]
        if s != s_count - 1:
            wave = _ifft2(_fft2(wave, workers=workers) * prop, workers=workers)
    return np.fft.fftshift(_fft2(wave, workers=workers), axes=(-2, -1))


def cbed_amplitude(pot_slices, slice_thickness: float, pot_calib: float,
                   beam_modes, beam_calib: float, voltage_kv: float):
    """``multislice/simulations.py:606-679``: reference layout (H, W, S) / (H, W, M).

    calib = min(pot.calib, beam.calib).  Returns (H, W, M) complex128.
    """
    pot = np.asarray(pot_slices, np.float64)
    if pot.ndim == 2:
        pot = pot[..., None]
    modes = np.asarray(beam_modes, np.complex128)
    if modes.ndim == 2:
        modes = modes[..., None]
    calib = min(float(pot_calib), float(beam_calib))
    h, w = modes.shape[0], modes.shape[1]
    prop = propagation_func(h, w, slice_thickness, voltage_kv, calib)
    trans = np.stack([transmission_func(pot[..., s], voltage_kv) for s in range(pot.shape[2])])
    amp = multislice_amplitude(trans, np.moveaxis(modes, 2, 0), prop)
    return np.moveaxis(amp, 0, 2)


def incoherent_sum(amplitudes, weights):
    """I = sum_m w_m |A_m|^2 over the last axis; ``multislice/reduce.py:81-86``."""
    return np.einsum("m,...m->...", np.asarray(weights, np.float64),
                     np.abs(amplitudes) ** 2)


def cbed_image(pot_slices, slice_thickness, pot_calib, beam_modes, beam_weights,
               beam_calib, voltage_kv):
    """``multislice/simulations.py:1130-1233`` without ensemble axes.

    Returns (pattern (H, W) float64, fourier calibration 1/(H calib)).
    """
    amp = cbed_amplitude(pot_slices, slice_thickness, pot_calib, beam_modes,
                         beam_calib, voltage_kv)
    calib = min(float(pot_calib), float(beam_calib))
    return incoherent_sum(amp, beam_weights), 1.0 / (amp.shape[0] * calib)


def stem_4d(pot_slices, slice_thickness, pot_calib, beam_modes, beam_weights,
            beam_calib, voltage_kv, scan_positions_px, detector_calib_ang):
    """``multislice/simulations.py:1330-1443``.

    The positions in ``scan_positions_px`` are passed unchanged to
    ``shift_beam_fourier`` which treats them as Angstrom (``:1390-1392``).
    Returns dict with ``data`` (P, H, W) float64 and the carrier fields.
    """
    pos = np.atleast_2d(np.asarray(scan_positions_px, np.float64))
    shifted = shift_beam_fourier(beam_modes, pos, detector_calib_ang)
    pot = np.asarray(pot_slices, np.float64)
    if pot.ndim == 2:
        pot = pot[..., None]
    calib = min(float(pot_calib), float(beam_calib))
    h, w = shifted.shape[1], shifted.shape[2]
    prop = propagation_func(h, w, slice_thickness, voltage_kv, calib)
    trans = np.stack([transmission_func(pot[..., s], voltage_kv) for s in range(pot.shape[2])])
    amp = multislice_amplitude(trans, np.moveaxis(shifted, 3, 1), prop)  # (P, M, H, W)
    data = np.einsum("m,pmhw->phw", np.asarray(beam_weights, np.float64), np.abs(amp) ** 2)
    return {
        "data": data,
        "real_space_calib": float(detector_calib_ang),
        "fourier_space_calib": 1.0 / (h * calib),
        "scan_positions": pos * float(detector_calib_ang),
    }


# ---------------------------------------------------------------------------
# Window-crop semantics (a slab larger than the probe window).
# ---------------------------------------------------------------------------

def annular_detector(data, fourier_space_calib: float, voltage_kv: float, inner_mrad: float, outer_mrad: float,
                     scan_shape):
    """Virtual annular detector (``multislice/simulations.py:1540-1651``): per pattern, the sum of
    the pixels whose radius sqrt(qy^2 + qx^2) * fourier_space_calib, with q measured from
    (H//2, W//2), lies in [inner, outer] (inclusive) in inverse Angstrom (mrad / 1000 / lambda);
    the sums are reshaped to the scan raster."""
    data = np.asarray(data, np.float64)
    lam = wavelength_ang(voltage_kv)
    inner_k = (inner_mrad / 1000.0) / lam
    outer_k = (outer_mrad / 1000.0) / lam
    _, hh, ww = data.shape
    qy = np.arange(hh) - hh // 2
    qx = np.arange(ww) - ww // 2
    qya, qxa = np.meshgrid(qy, qx, indexing="ij")
    radius = np.sqrt(qya**2 + qxa**2) * fourier_space_calib
    mask = (radius >= inner_k) & (radius <= outer_k)
    return (data * mask).sum(axis=(1, 2)).reshape(scan_shape)


def ramp_1d(n: int, shift_px: float):
    """exp(-2 pi i k shift/n) with k = fftfreq(n)*n (Nyquist negative)."""
    k = np.fft.fftfreq(n) * n
    return np.exp(-2j * np.pi * k * float(shift_px) / n)


def shifted_probe(probe, shift_px):
    """Probe shifted by (sy, sx) pixels with the Fourier phase ramp of
    ``shift_beam_fourier`` (``simulations.py:1287-1321``); shift in pixels."""
    n_y, n_x = probe.shape[-2], probe.shape[-1]
    ramp = ramp_1d(n_y, shift_px[0])[:, None] * ramp_1d(n_x, shift_px[1])[None, :]
    return np.fft.ifft2(np.fft.fft2(probe) * ramp)


def window_of(slab, origin, n):
    """(S, h, w) periodic window of ``slab`` (S, Hs, Ws) at (oy, ox); ``n`` is h = w or (h, w)."""
    h, w = (n, n) if np.ndim(n) == 0 else n
    hs, ws = slab.shape[-2], slab.shape[-1]
    rows = (int(origin[0]) + np.arange(h)) % hs
    cols = (int(origin[1]) + np.arange(w)) % ws
    return slab[..., rows[:, None], cols[None, :]]


def window_stem(slab_t, probe, origins, shifts_px, prop, workers: int = 1,
                return_amplitude: bool = False):
    """4D-STEM patterns with per-position transmission windows.

    The window semantics used by the production configs: the pattern at
    scan position p equals the reference ``cbed_image``
    (``simulations.py:1130-1233``) evaluated on the periodic T window of
    ``slab_t`` at integer ``origins[p]`` with the probe shifted by
    ``shifts_px[p]`` pixels through ``shift_beam_fourier``.

    slab_t: (S, Hs, Ws) complex transmission; probe: (h, w) complex;
    origins: (P, 2) int; shifts_px: (P, 2) float; prop: (h, w) complex.
    Returns (P, h, w) float64 patterns (and amplitudes if requested).
    """
    h, w = probe.shape[-2], probe.shape[-1]
    origins = np.asarray(origins)
    shifts_px = np.asarray(shifts_px, np.float64)
    out = np.empty((origins.shape[0], h, w), np.float64)
    amps = [] if return_amplitude else None
    probe_k = _fft2(np.asarray(probe, np.complex128))
    for p in range(origins.shape[0]):
        ramp = ramp_1d(h, shifts_px[p, 0])[:, None] * ramp_1d(w, shifts_px[p, 1])[None, :]
        wave0 = _ifft2(probe_k * ramp, workers=workers)
        t_win = window_of(slab_t, origins[p], (h, w))
        amp = multislice_amplitude(t_win, wave0, prop, workers=workers)
        out[p] = np.abs(amp) ** 2
        if return_amplitude:
            amps.append(amp)
    if return_amplitude:
        return out, np.stack(amps)
    return out


def window_stem_batched(slab_t, probe, origins, shifts_px, prop, workers: int = 1,
                        dtype=np.complex128):
    """Same as ``window_stem`` but processes all positions as one batch.

    Used as the timed CPU baseline (pocketfft over ``workers`` threads);
    ``dtype`` selects complex128 (default) or complex64 arithmetic.
    """
    n = probe.shape[-1]
    origins = np.asarray(origins)
    shifts_px = np.asarray(shifts_px, np.float64)
    probe_k = _fft2(np.asarray(probe, np.complex128))
    ramps = np.stack([ramp_1d(n, sy)[:, None] * ramp_1d(n, sx)[None, :] for sy, sx in shifts_px])
    wave = _ifft2(probe_k[None] * ramps, workers=workers).astype(dtype)
    prop_c = prop.astype(dtype)
    s_count = slab_t.shape[0]
    for s in range(s_count):
        t_s = np.stack([window_of(slab_t[s], o, n) for o in origins]).astype(dtype)
        wave *= t_s
        if s != s_count - 1:
            wave = _ifft2(_fft2(wave, workers=workers) * prop_c, workers=workers)
    amp = np.fft.fftshift(_fft2(wave, workers=workers), axes=(-2, -1))
    return (np.abs(amp) ** 2).astype(np.float64)


def window_stem_perturbed(slab_t, probe, origins, shifts_px, prop, index, deltas, workers: int = 1):
    """``window_stem_batched`` of ``slab_t`` with ``deltas[j]`` added to the one element ``index`` = (s, y, x),
    for every j: (len(deltas), P, h, w) patterns.  The waves entering slice s do not depend on the
    perturbation, so they are computed once and shared by all variants (central differences in the
    real and imaginary directions need four variants of one pixel)."""
    n = probe.shape[-1]
    s_pix, y_pix, x_pix = (int(v) for v in index)
    origins = np.asarray(origins)
    shifts_px = np.asarray(shifts_px, np.float64)
    probe_k = _fft2(np.asarray(probe, np.complex128))
    ramps = np.stack([ramp_1d(n, sy)[:, None] * ramp_1d(n, sx)[None, :] for sy, sx in shifts_px])
    wave = _ifft2(probe_k[None] * ramps, workers=workers)
    s_count = slab_t.shape[0]
    for s in range(s_pix):
        wave = wave * np.stack([window_of(slab_t[s], o, n) for o in origins])
        wave = _ifft2(_fft2(wave, workers=workers) * prop, workers=workers)
    out = []
    for d in deltas:
        sl = np.array(slab_t[s_pix:], np.complex128)
        sl[0, y_pix, x_pix] += d
        w = wave.copy()
        for s in range(sl.shape[0]):
            w = w * np.stack([window_of(sl[s], o, n) for o in origins])
            if s_pix + s != s_count - 1:
                w = _ifft2(_fft2(w, workers=workers) * prop, workers=workers)
        amp = np.fft.fftshift(_fft2(w, workers=workers), axes=(-2, -1))
        out.append(np.abs(amp) ** 2)
    return np.stack(out)


# ---------------------------------------------------------------------------
# Gradient (reverse mode) of the windowed forward, complex128.
# ---------------------------------------------------------------------------

def multislice_vjp(slab_t, probe, origins, shifts_px, prop, grad_intensity, want_shifts: bool = False):
    """Vector-Jacobian product of ``window_stem`` w.r.t. the slab, probe and shifts.

    Convention: for a real loss L, returns dL/dRe + i dL/dIm for every complex
    input (the torch convention; ``jax.grad`` returns the conjugate).  The
    reference differentiates the same scan with ``jax.value_and_grad``
    (``multislice/multislice_recon.py:1915-1951``), including the scan
    positions in its position-correcting reconstructors
    (``multislice_recon.py:849-885``).

    grad_intensity: (P, h, w) real dL/dI.  Returns (grad_slab, grad_probe),
    plus grad_shifts (P, 2) = dL/d(shift_y, shift_x) in pixels when
    ``want_shifts``.
    """
    h, w = probe.shape[-2], probe.shape[-1]
    s_count = slab_t.shape[0]
    hs, ws = slab_t.shape[-2], slab_t.shape[-1]
    grad_slab = np.zeros(slab_t.shape, np.complex128)
    probe_k = np.fft.fft2(np.asarray(probe, np.complex128))
    grad_probe_k = np.zeros((h, w), np.complex128)
    grad_shifts = np.zeros((len(origins), 2))
    ky = np.fft.fftfreq(h) * h  # Nyquist negative, as in ramp_1d
    kx = np.fft.fftfreq(w) * w
    for p in range(len(origins)):
        ramp = ramp_1d(h, shifts_px[p][0])[:, None] * ramp_1d(w, shifts_px[p][1])[None, :]
        psi = [np.fft.ifft2(probe_k * ramp)]
        t_win = window_of(slab_t, origins[p], (h, w))
        for s in range(s_count):
            phi = psi[s] * t_win[s]
            if s != s_count - 1:
                psi.append(np.fft.ifft2(np.fft.fft2(phi) * prop))
        amp = np.fft.fftshift(np.fft.fft2(phi))
        g_amp = 2.0 * grad_intensity[p] * amp
        # adjoint of fftshift(fft2): F^H = n^2 ifft2, applied after ifftshift
        g_phi = np.fft.ifft2(np.fft.ifftshift(g_amp)) * (h * w)
        rows = (int(origins[p][0]) + np.arange(h)) % hs
        cols = (int(origins[p][1]) + np.arange(w)) % ws
        for s in range(s_count - 1, -1, -1):
            if s != s_count - 1:
                g_phi = np.fft.ifft2(np.fft.fft2(g_psi) * np.conj(prop))
            np.add.at(grad_slab[s], (rows[:, None], cols[None, :]), np.conj(psi[s]) * g_phi)
            g_psi = np.conj(t_win[s]) * g_phi
        # psi0 = ifft2(u), u = probe_k * ramp, probe_k = fft2(probe)
        g_u = np.fft.fft2(g_psi) / (h * w)
        grad_probe_k += np.conj(ramp) * g_u
        # d ramp / d shift_y = ramp * (-2 pi i k_y / h); dL/ds = Re sum conj(g_u) du/ds
        du = probe_k * ramp * (-2j * np.pi)
        grad_shifts[p, 0] = np.real(np.sum(np.conj(g_u) * du * (ky / h)[:, None]))
        grad_shifts[p, 1] = np.real(np.sum(np.conj(g_u) * du * (kx / w)[None, :]))
    grad_probe = np.fft.ifft2(grad_probe_k) * (h * w)
    if want_shifts:
        return grad_slab, grad_probe, grad_shifts
    return grad_slab, grad_probe


def multislice_jvp(slab_t, probe, origins, shifts_px, prop, d_slab, d_probe):
    """Jacobian-vector product of ``window_stem`` (intensity) in complex128.

    Tangents d_slab (S, Hs, Ws), d_probe (N, N) complex.  Returns dI (P, N, N).
    """
    n = probe.shape[-1]
    s_count = slab_t.shape[0]
    probe_k = np.fft.fft2(np.asarray(probe, np.complex128))
    d_probe_k = np.fft.fft2(np.asarray(d_probe, np.complex128))
    out = np.empty((len(origins), n, n))
    for p in range(len(origins)):
        ramp = ramp_1d(n, shifts_px[p][0])[:, None] * ramp_1d(n, shifts_px[p][1])[None, :]
        psi = np.fft.ifft2(probe_k * ramp)
        dpsi = np.fft.ifft2(d_probe_k * ramp)
        t_win = window_of(slab_t, origins[p], n)
        dt_win = window_of(d_slab, origins[p], n)
        for s in range(s_count):
            phi = psi * t_win[s]
            dphi = dpsi * t_win[s] + psi * dt_win[s]
            if s != s_count - 1:
                psi = np.fft.ifft2(np.fft.fft2(phi) * prop)
                dpsi = np.fft.ifft2(np.fft.fft2(dphi) * prop)
        amp = np.fft.fft2(phi)
        damp = np.fft.fft2(dphi)
        out[p] = np.fft.fftshift(2.0 * np.real(np.conj(amp) * damp))
    return out


# ---------------------------------------------------------------------------
# Twelve-coefficient aberration function (build extension for the GN/LM
# config); reduces to ``aberration`` for C1/C3 only.
# ---------------------------------------------------------------------------

ABERRATION_NAMES = ("C1", "A1a", "A1b", "B2a", "B2b", "C3", "A2a", "A2b",
                    "S3a", "S3b", "A3a", "A3b")


def aberration12(qy, qx, lambda_angstrom: float, coeffs):
    """chi(q) for the axial aberrations up to third order, Angstrom units.

    omega = lambda (qx + i qy) is the complex scattering angle and
    chi = (2 pi / lambda) Re[ C1 |w|^2/2 + A1 conj(w)^2/2 + B2 w conj(w)^2
          + A2 conj(w)^3/3 + C3 |w|^4/4 + S3 w conj(w)^3 + A3 conj(w)^4/4 ]
    with complex A1 = A1a + i A1b (likewise B2, A2, S3, A3).  With only C1
    and C3 = 1e7 c3_ang set it equals the reference ``aberration``
    (``simulations.py:434-498``).
    """
    c = [float(v) for v in coeffs]
    w = lambda_angstrom * (np.asarray(qx) + 1j * np.asarray(qy))
    wc = np.conj(w)
    ww = (w * wc).real
    a1 = c[1] + 1j * c[2]
    b2 = c[3] + 1j * c[4]
    a2 = c[6] + 1j * c[7]
    s3 = c[8] + 1j * c[9]
    a3 = c[10] + 1j * c[11]
    chi = (c[0] * ww / 2 + a1 * wc**2 / 2 + b2 * w * wc**2 + a2 * wc**3 / 3
           + c[5] * ww**2 / 4 + s3 * w * wc**3 + a3 * wc**4 / 4)
    return (2 * np.pi / lambda_angstrom) * np.real(chi)


def probe_from_aberrations(n: int, calib_ang: float, voltage_kv: float,
                           aperture_mrad: float, coeffs):
    """ifftshift(ifft2(A exp(-i chi12))) on the fftfreq grid (``make_probe`` form)."""
    lam = wavelength_ang(voltage_kv)
    q = fftfreq(n, calib_ang)
    qy, qx = np.meshgrid(q, q, indexing="ij")
    aperture = (np.sqrt(qx**2 + qy**2) <= (aperture_mrad / 1000.0) / lam).astype(np.float64)
    chi = aberration12(qy, qx, lam, coeffs)
    return np.fft.ifftshift(np.fft.ifft2(aperture * np.exp(-1j * chi)))


# ---------------------------------------------------------------------------
# Dense Levenberg-Marquardt step rules (``jacobian/solvers.py:387-521``).
# ---------------------------------------------------------------------------

def lm_step_dense(jac, residual, damping):
    """One LM step on a dense real Jacobian (``solvers.py:450-474``): g = J^T r with r = f(theta) - data,
    s solves (J^T J + lambda I) s = g, and the candidate is theta - s.  Returns (s, J^T J, g)."""
    jtj = jac.T @ jac
    g = jac.T @ residual
    step = np.linalg.solve(jtj + damping * np.eye(jtj.shape[0]), g)
    return step, jtj, g


def lm_gain_ratio(rss: float, rss_new: float, step, g, jtj) -> float:
    """(|r|^2 - |r_new|^2) / (<s, g> - 1/2 <s, J^T J s> + 1e-12) (``solvers.py:480-489``).

    The numerator is a reduction of |r|^2 and the denominator that of |r|^2 / 2 under the quadratic
    model, so an exact quadratic step gives 2 (the reference's factor, kept)."""
    predicted = float(step @ g - 0.5 * step @ (jtj @ step))
    return (float(rss) - float(rss_new)) / (predicted + 1e-12)


def lm_update_damping(rho: float, damping: float) -> float:
    """lambda/3 if rho > 0.75, 2 lambda if rho < 0.25 (``solvers.py:491-503``)."""
    if rho > 0.75:
        return damping / 3.0
    if rho < 0.25:
        return damping * 2.0
    return damping


# ---------------------------------------------------------------------------
# JAX Threefry-2x32 PRNG restatement (needed only to regenerate the input of
# the reference's golden capture; Salmon et al., SC'11, Random123).
# ---------------------------------------------------------------------------

_ROT = ((13, 15, 26, 6), (17, 29, 16, 24))


def _rotl32(x, r):
    return ((x << np.uint32(r)) | (x >> np.uint32(32 - r))).astype(np.uint32)


def threefry2x32(key, x0, x1):
    """Threefry-2x32 with 20 rounds on uint32 arrays (x0, x1)."""
    k0, k1 = np.uint32(key[0]), np.uint32(key[1])
    ks = (k0, k1, np.uint32(k0 ^ k1 ^ np.uint32(0x1BD11BDA)))
    x0 = (np.asarray(x0, np.uint32) + ks[0]).astype(np.uint32)
    x1 = (np.asarray(x1, np.uint32) + ks[1]).astype(np.uint32)
    with np.errstate(over="ignore"):
        for block in range(5):
            for r in _ROT[block % 2]:
                x0 = (x0 + x1).astype(np.uint32)
                x1 = _rotl32(x1, r)
                x1 = (x1 ^ x0).astype(np.uint32)
            x0 = (x0 + ks[(block + 1) % 3]).astype(np.uint32)
            x1 = (x1 + ks[(block + 2) % 3] + np.uint32(block + 1)).astype(np.uint32)
    return x0, x1


def jax_uniform_f64(seed: int, shape):
    """``jax.random.uniform(PRNGKey(seed), shape, float64)`` restated.

    PRNGKey(seed) = (seed >> 32, seed & 0xffffffff); partitionable Threefry
    bits: counter = (hi, lo) words of the flat C-order index; 64-bit word =
    out0 << 32 | out1; uniform = bitcast((bits >> 12) | bits(1.0)) - 1.
    """
    key = (np.uint32((seed >> 32) & 0xFFFFFFFF), np.uint32(seed & 0xFFFFFFFF))
    count = int(np.prod(shape))
    idx = np.arange(count, dtype=np.uint64)
    hi = (idx >> np.uint64(32)).astype(np.uint32)
    lo = (idx & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    b0, b1 = threefry2x32(key, hi, lo)
    bits = (b0.astype(np.uint64) << np.uint64(32)) | b1.astype(np.uint64)
    one = np.array(1.0, np.float64).view(np.uint64)
    fbits = (bits >> np.uint64(12)) | one
    return (fbits.view(np.float64) - 1.0).reshape(shape)


# ---------------------------------------------------------------------------
# FP64 Levenberg-Marquardt over the twelve probe aberrations (the oracle run
# the GPU mixed-precision fit is compared with).  Same update rules as
# ``levenberg_marquardt_step`` (``jacobian/solvers.py:387-521``), through
# ``lm_step_dense`` (step from (J^T J + lambda I) s = J^T r, candidate
# theta - s), ``lm_gain_ratio`` and ``lm_update_damping`` (lambda/3 above
# 0.75, 2 lambda below 0.25); accept when the ratio is positive.
# ---------------------------------------------------------------------------

def aberration_basis12(n: int, calib_ang: float, voltage_kv: float):
    """(12, n, n) d chi / d c_j on the fftfreq grid (chi is linear in the coefficients)."""
    lam = wavelength_ang(voltage_kv)
    q = fftfreq(n, calib_ang)
    qy, qx = np.meshgrid(q, q, indexing="ij")
    eye = np.eye(12)
    return np.stack([aberration12(qy, qx, lam, eye[j]) for j in range(12)])


def probe_k_from_coeffs(n, calib_ang, voltage_kv, aperture_mrad, coeffs):
    """fft2 of probe_from_aberrations (computed through the FFTs, not the closed form)."""
    return np.fft.fft2(probe_from_aberrations(n, calib_ang, voltage_kv, aperture_mrad, coeffs))


def window_amplitudes(slab_t, probe_k_modes, origins, shifts_px, prop):
    """(P, M, n, n) unshifted exit amplitudes F(phi_{S-1}) for every probe mode (given as fft2)."""
    n = probe_k_modes.shape[-1]
    out = np.empty((len(origins), probe_k_modes.shape[0], n, n), np.complex128)
    for p in range(len(origins)):
        ramp = ramp_1d(n, shifts_px[p][0])[:, None] * ramp_1d(n, shifts_px[p][1])[None, :]
        wave = np.fft.ifft2(probe_k_modes * ramp[None])
        t_win = window_of(slab_t, origins[p], n)
        for s in range(slab_t.shape[0]):
            wave = wave * t_win[s][None]
            if s != slab_t.shape[0] - 1:
                wave = np.fft.ifft2(np.fft.fft2(wave) * prop[None])
        out[p] = np.fft.fft2(wave)
    return out


def lm_aberration_fit(slab_t, origins, shifts_px, data, coeffs_init, calib_ang, voltage_kv, aperture_mrad,
                      slice_thickness, iterations=30, damping_init=1e-3, tolerance=1e-8):
    """FP64 LM on the 12 aberration coefficients, parameters scaled to max |dchi| = 1 rad in the aperture.

    data: (P, n, n) fftshifted intensities.  Returns (coefficients, residual norm history).
    """
    n = data.shape[-1]
    lam = wavelength_ang(voltage_kv)
    basis = aberration_basis12(n, calib_ang, voltage_kv)
    q = fftfreq(n, calib_ang)
    qy, qx = np.meshgrid(q, q, indexing="ij")
    aperture = (np.sqrt(qx**2 + qy**2) <= (aperture_mrad / 1000.0) / lam).astype(float)
    unit = 1.0 / np.max(np.abs(basis) * aperture[None], axis=(1, 2))
    sbasis = basis * unit[:, None, None]
    prop = propagation_func(n, n, slice_thickness, voltage_kv, calib_ang)
    data_u = np.fft.ifftshift(np.asarray(data, np.float64), axes=(-2, -1))

    def pk_of(theta):
        return probe_k_from_coeffs(n, calib_ang, voltage_kv, aperture_mrad, theta * unit)

    def rss_of(theta):
        a = window_amplitudes(slab_t, pk_of(theta)[None], origins, shifts_px, prop)[:, 0]
        return float(np.sum((np.abs(a) ** 2 - data_u) ** 2))

    theta = np.asarray(coeffs_init, np.float64) / unit
    damping = float(damping_init)
    norm = np.sqrt(rss_of(theta))
    hist = [norm]
    for _ in range(iterations):
        if norm < tolerance:
            break
        pk0 = pk_of(theta)
        modes = np.concatenate([pk0[None], (-1j) * sbasis * pk0[None]], axis=0)
        amp = window_amplitudes(slab_t, modes, origins, shifts_px, prop)
        a0 = amp[:, 0]
        jac = 2.0 * np.real(np.conj(a0)[:, None] * amp[:, 1:])          # (P, 12, n, n)
        r = np.abs(a0) ** 2 - data_u
        jmat = jac.transpose(1, 0, 2, 3).reshape(12, -1)
        rss = float(np.sum(r**2))
        step, jtj, g = lm_step_dense(jmat.T, r.ravel(), damping)
        cand = theta - step
        rss_new = rss_of(cand)
        gain = lm_gain_ratio(rss, rss_new, step, g, jtj)
        damping = lm_update_damping(gain, damping)
        if gain > 0:
            theta, norm = cand, np.sqrt(rss_new)
        else:
            norm = np.sqrt(rss)
        hist.append(norm)
    return theta * unit, hist


# ---------------------------------------------------------------------------
# Atom-based slice provider and stem4d_sharded (``multislice/parallelized.py``)
# ---------------------------------------------------------------------------

def slice_potential_from_atoms(atom_coords, atom_types, z_min, z_max, atom_potentials, grid_shape, calib_ang,
                               atom_mask=None):
    """``_compute_slice_potential`` (``parallelized.py:54-169``).

    Atoms with z_min <= z < z_max are scattered to a delta grid per type at
    (floor(y/calib) % H, floor(x/calib) % W) (coordinates are (x, y, z)) and
    FFT-convolved with that type's kernel; the real parts are summed.
    """
    h, w = grid_shape
    coords = np.asarray(atom_coords, np.float64)
    types = np.asarray(atom_types)
    kern = np.asarray(atom_potentials, np.float64)
    in_slice = ((coords[:, 2] >= z_min) & (coords[:, 2] < z_max)).astype(np.float64)
    if atom_mask is not None:
        in_slice = in_slice * np.asarray(atom_mask, np.float64)
    x_idx = np.floor(coords[:, 0] / calib_ang).astype(np.int64) % w
    y_idx = np.floor(coords[:, 1] / calib_ang).astype(np.int64) % h
    total = np.zeros((h, w))
    for t in range(kern.shape[0]):
        grid = np.zeros((h, w))
        np.add.at(grid, (y_idx, x_idx), (types == t) * in_slice)
        total += np.real(np.fft.ifft2(np.fft.fft2(grid) * np.fft.fft2(kern[t])))
    return total


def stem4d_from_atoms(probe_modes, weights, atom_coords, atom_types, slice_z_bounds, atom_potentials, voltage_kv,
                      calib_ang, scan_positions_ang, atom_mask=None):
    """``stem4d_sharded`` (``parallelized.py:348-536``) without a mesh: (P, H, W) float64 patterns."""
    modes = np.asarray(probe_modes, np.complex128)
    if modes.ndim == 2:
        modes = modes[..., None]
    h, w = modes.shape[0], modes.shape[1]
    bounds = np.asarray(slice_z_bounds, np.float64)
    dz = bounds[0, 1] - bounds[0, 0]
    trans = np.stack([transmission_func(slice_potential_from_atoms(atom_coords, atom_types, b[0], b[1],
                                                                   atom_potentials, (h, w), calib_ang, atom_mask),
                                        voltage_kv) for b in bounds])
    prop = propagation_func(h, w, dz, voltage_kv, calib_ang)
    shifted = shift_beam_fourier(modes, scan_positions_ang, calib_ang)       # (P, H, W, M)
    amp = multislice_amplitude(trans, np.moveaxis(shifted, 3, 1), prop)       # (P, M, H, W)
    return np.einsum("m,pmhw->phw", np.asarray(weights, np.float64), np.abs(amp) ** 2)
