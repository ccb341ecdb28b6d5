"""FP64 oracle of the joint aberration + object Levenberg-Marquardt reconstruction.

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` and the fixture script under
``tests/golden/`` import this module, as the checker.  The product path in
``ptyrodactyl_amd`` never calls into it.

Unknowns: the twelve probe aberration coefficients (scaled so that a unit step
moves chi by at most 1 rad inside the aperture) and one complex object slice
that every slice of the specimen repeats.  Every iteration is one LM step on
the coefficients with the object fixed (dense normal equations, as
``ms_oracle.lm_aberration_fit``), then one LM step on the object with the
coefficients fixed, matrix-free as the reference's
``levenberg_marquardt_step`` (``src/ptyrodactyl/jacobian/solvers.py:387-521``):
the gradient g = J^T r, a conjugate-gradient solve of (J^T J + lambda I) s = g
for a fixed number of iterations with the reference's recurrence
(``solvers.py:62-180``, ``alpha = rr/(pAp + 1e-12)``), candidate O - s, the
gain ratio (|r|^2 - |r_new|^2) / (<s, g> - 1/2 <s, J^T J s> + 1e-12), lambda/3
above 0.75 and 2 lambda below 0.25, acceptance when the ratio is positive.
J^T J v is the VJP of the JVP, as ``jtj_operator`` (``jacobian/operators.py:132-185``)
builds it.  All arithmetic is complex128 / float64.

Pins (``tests/test_joint_oracle.py``): the batched forward, JVP and VJP equal
``ms_oracle``'s per-position ``window_stem`` / ``multislice_jvp`` /
``multislice_vjp`` (themselves pinned by the reference's golden capture, central
differences and the dot-product identity); the object JVP / VJP pass the
dot-product identity and central differences; the CG solve matches a dense solve
of the same damped system; the fit recovers the coefficients and the object from
noise-free data.
"""

from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import ms_oracle as o

try:
    import scipy.fft as _sfft

    def _fft2(a, workers):
        return _sfft.fft2(a, axes=(-2, -1), workers=workers)

    def _ifft2(a, workers):
        return _sfft.ifft2(a, axes=(-2, -1), workers=workers)
except ImportError:  # pragma: no cover
    def _fft2(a, workers):
        return np.fft.fft2(a, axes=(-2, -1))

    def _ifft2(a, workers):
        return np.fft.ifft2(a, axes=(-2, -1))


def _windows(obj, origins, n):
    """(P, n, n) periodic windows of a 2-D array at the integer origins."""
    h, w = obj.shape
    rows = (origins[:, 0][:, None] + np.arange(n)[None, :]) % h
    cols = (origins[:, 1][:, None] + np.arange(n)[None, :]) % w
    return obj[rows[:, :, None], cols[:, None, :]]


def _scatter_windows(contrib, origins, shape):
    """Sum of (P, n, n) window contributions into a 2-D array (adjoint of ``_windows``)."""
    out = np.zeros(shape, np.complex128)
    n = contrib.shape[-1]
    h, w = shape
    for p in range(contrib.shape[0]):
        oy, ox = int(origins[p, 0]), int(origins[p, 1])
        if oy + n <= h and ox + n <= w:
            out[oy:oy + n, ox:ox + n] += contrib[p]
        else:
            rows = (oy + np.arange(n)) % h
            cols = (ox + np.arange(n)) % w
            np.add.at(out, (rows[:, None], cols[None, :]), contrib[p])
    return out


def _incident(probe_k, shifts_px, n, workers):
    ramps = np.stack([o.ramp_1d(n, sy)[:, None] * o.ramp_1d(n, sx)[None, :] for sy, sx in shifts_px])
    return _ifft2(np.asarray(probe_k, np.complex128)[None] * ramps, workers)


def _position_chunks(count, workers):
    return [c for c in np.array_split(np.arange(count), max(1, min(int(workers), count))) if c.size]


def _by_chunks(fn, count, workers, combine):
    """fn(index array) over position chunks on ``workers`` threads (numpy and pocketfft release the GIL);
    ``combine`` joins the chunk results in chunk order."""
    chunks = _position_chunks(count, workers)
    if len(chunks) == 1:
        return fn(chunks[0])
    with ThreadPoolExecutor(len(chunks)) as ex:
        return combine(list(ex.map(fn, chunks)))


def forward_amplitudes(obj, num_slices, probe_k, origins, shifts_px, prop, workers=8, keep_waves=False):
    """Exit amplitudes F(phi_{S-1}) (P, n, n), unshifted, of a specimen repeating ``obj`` ``num_slices``
    times; with ``keep_waves`` also the waves psi_s entering each slice (S, P, n, n)."""
    if not keep_waves and workers > 1:
        return _by_chunks(lambda i: forward_amplitudes(obj, num_slices, probe_k, origins[i], shifts_px[i], prop, 1),
                          len(origins), workers, np.concatenate)
    workers = 1
    n = probe_k.shape[-1]
    t_win = _windows(obj, origins, n)
    wave = _incident(probe_k, shifts_px, n, workers)
    waves = []
    for s in range(num_slices):
        if keep_waves:
            waves.append(wave)
        phi = wave * t_win
        if s != num_slices - 1:
            wave = _ifft2(_fft2(phi, workers) * prop, workers)
    amp = _fft2(phi, workers)
    return (amp, np.stack(waves)) if keep_waves else amp


def patterns(obj, num_slices, probe_k, origins, shifts_px, prop, workers=8):
    """fftshifted intensities (P, n, n)."""
    a = forward_amplitudes(obj, num_slices, probe_k, origins, shifts_px, prop, workers)
    return np.fft.fftshift(np.abs(a) ** 2, axes=(-2, -1))


def object_jvp(obj, num_slices, probe_k, origins, shifts_px, prop, d_obj, workers=8):
    """Tangent of the fftshifted intensities for the object tangent ``d_obj`` (every slice moves with it)."""
    if workers > 1:
        return _by_chunks(lambda i: object_jvp(obj, num_slices, probe_k, origins[i], shifts_px[i], prop, d_obj, 1),
                          len(origins), workers, np.concatenate)
    n = probe_k.shape[-1]
    t_win = _windows(obj, origins, n)
    dt_win = _windows(d_obj, origins, n)
    wave = _incident(probe_k, shifts_px, n, workers)
    dwave = np.zeros_like(wave)
    for s in range(num_slices):
        phi = wave * t_win
        dphi = dwave * t_win + wave * dt_win
        if s != num_slices - 1:
            wave = _ifft2(_fft2(phi, workers) * prop, workers)
            dwave = _ifft2(_fft2(dphi, workers) * prop, workers)
    amp = _fft2(phi, workers)
    damp = _fft2(dphi, workers)
    return np.fft.fftshift(2.0 * np.real(np.conj(amp) * damp), axes=(-2, -1))


def object_vjp(obj, num_slices, probe_k, origins, shifts_px, prop, grad_i, workers=8):
    """dL/dRe(O) + i dL/dIm(O) for dL/dI = ``grad_i`` (fftshifted, (P, n, n)); the slices' gradients add
    (position chunks are summed in chunk order)."""
    if workers > 1:
        return _by_chunks(lambda i: object_vjp(obj, num_slices, probe_k, origins[i], shifts_px[i], prop, grad_i[i], 1),
                          len(origins), workers, lambda parts: np.sum(np.stack(parts), axis=0))
    n = probe_k.shape[-1]
    t_win = _windows(obj, origins, n)
    amp, waves = forward_amplitudes(obj, num_slices, probe_k, origins, shifts_px, prop, workers, keep_waves=True)
    g_amp = 2.0 * np.fft.ifftshift(np.asarray(grad_i, np.float64), axes=(-2, -1)) * amp
    g_phi = _ifft2(g_amp, workers) * (n * n)  # adjoint of fft2 (unnormalised): n^2 ifft2
    contrib = np.zeros_like(g_phi)
    for s in range(num_slices - 1, -1, -1):
        if s != num_slices - 1:
            g_phi = _ifft2(_fft2(g_psi, workers) * np.conj(prop), workers)
        contrib += np.conj(waves[s]) * g_phi
        g_psi = np.conj(t_win) * g_phi
    return _scatter_windows(contrib, origins, obj.shape)


def cg_fixed(apply_a, rhs, iterations):
    """The reference's conjugate gradient from x0 = 0 for a fixed number of iterations (no early stop)."""
    x = np.zeros_like(rhs)
    r = rhs.copy()
    p = r.copy()
    rr = float(np.real(np.vdot(r, r)))
    for _ in range(iterations):
        ap = apply_a(p)
        alpha = rr / (float(np.real(np.vdot(p, ap))) + 1e-12)
        x = x + alpha * p
        r = r - alpha * ap
        rr_new = float(np.real(np.vdot(r, r)))
        beta = rr_new / (rr + 1e-12)
        p = r + beta * p
        rr = rr_new
    return x


def joint_fit(obj_init, num_slices, origins, shifts_px, data, coeffs_init, calib_ang, voltage_kv, aperture_mrad,
              slice_thickness, iterations=30, damping_coeffs=1e-3, damping_object=1e-3, cg_iterations=3,
              workers=8, log=None):
    """Block-coordinate LM over (12 aberrations, object); returns (coefficients, object, history).

    history[t] = (|r| after the coefficient step, |r| after the object step) of iteration t, with the
    residual norm of the start first."""
    n = data.shape[-1]
    lam = o.wavelength_ang(voltage_kv)
    basis = o.aberration_basis12(n, calib_ang, voltage_kv)
    q = o.fftfreq(n, calib_ang)
    qy, qx = np.meshgrid(q, q, indexing="ij")
    aperture = (np.sqrt(qx**2 + qy**2) <= (aperture_mrad / 1000.0) / lam).astype(float)
    unit = 1.0 / np.max(np.abs(basis) * aperture[None], axis=(1, 2))
    sbasis = basis * unit[:, None, None]
    prop = o.propagation_func(n, n, slice_thickness, voltage_kv, calib_ang)
    data = np.asarray(data, np.float64)
    theta = np.asarray(coeffs_init, np.float64) / unit
    obj = np.asarray(obj_init, np.complex128).copy()
    lam_c, lam_o = float(damping_coeffs), float(damping_object)

    def pk_of(th):
        return o.probe_k_from_coeffs(n, calib_ang, voltage_kv, aperture_mrad, th * unit)

    def residual(th, ob):
        return patterns(ob, num_slices, pk_of(th), origins, shifts_px, prop, workers) - data

    r = residual(theta, obj)
    rss = float(np.sum(r**2))
    history = [(np.sqrt(rss), np.sqrt(rss))]
    for it in range(iterations):
        # coefficients, object fixed: dense FP64 normal equations
        pk0 = pk_of(theta)
        amp0 = forward_amplitudes(obj, num_slices, pk0, origins, shifts_px, prop, workers)
        cols = []
        for j in range(12):
            amp_j = forward_amplitudes(obj, num_slices, (-1j) * sbasis[j] * pk0, origins, shifts_px, prop, workers)
            cols.append(np.fft.fftshift(2.0 * np.real(np.conj(amp0) * amp_j), axes=(-2, -1)).ravel())
        jac = np.stack(cols, axis=1)
        r = np.fft.fftshift(np.abs(amp0) ** 2, axes=(-2, -1)) - data
        rss = float(np.sum(r**2))
        step, jtj, g = o.lm_step_dense(jac, r.ravel(), lam_c)
        cand = theta - step
        r_new = residual(cand, obj)
        rss_new = float(np.sum(r_new**2))
        rho = o.lm_gain_ratio(rss, rss_new, step, g, jtj)
        lam_c = o.lm_update_damping(rho, lam_c)
        if rho > 0:
            theta, r, rss = cand, r_new, rss_new
        after_c = np.sqrt(rss)
        # object, coefficients fixed: matrix-free LM step with a fixed-length CG
        pk = pk_of(theta)
        g_obj = object_vjp(obj, num_slices, pk, origins, shifts_px, prop, r, workers)

        def jtj_obj(v):
            return object_vjp(obj, num_slices, pk, origins, shifts_px, prop,
                              object_jvp(obj, num_slices, pk, origins, shifts_px, prop, v, workers), workers)

        s_obj = cg_fixed(lambda v: jtj_obj(v) + lam_o * v, g_obj, cg_iterations)
        cand_obj = obj - s_obj
        r_new = residual(theta, cand_obj)
        rss_new = float(np.sum(r_new**2))
        predicted = float(np.real(np.vdot(s_obj, g_obj))) - 0.5 * float(np.real(np.vdot(s_obj, jtj_obj(s_obj))))
        rho = (rss - rss_new) / (predicted + 1e-12)
        lam_o = o.lm_update_damping(rho, lam_o)
        if rho > 0:
            obj, r, rss = cand_obj, r_new, rss_new
        history.append((after_c, np.sqrt(rss)))
        if log is not None:
            log(f"iteration {it + 1}: |r| {after_c:.6e} (coefficients)  {np.sqrt(rss):.6e} (object)  "
                f"lambda {lam_c:.3e} / {lam_o:.3e}")
    return theta * unit, obj, history
