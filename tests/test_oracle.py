"""Pins of the CPU oracle (``oracle/ms_oracle.py``) — no GPU needed.

Each check is chosen so that a plausible mistake in the oracle (dropped
term, wrong sign or index, transposed operand) fails at least one of them:
golden arrays captured from the reference, published constants, closed
forms, brute-force DFT evaluation and independent autodiff.
"""

from pathlib import Path

import numpy as np
import pytest

from oracle import ms_oracle as o

GOLDEN = Path(__file__).resolve().parent / "golden" / "plan01_ug_reference.npz"


def _plan01_inputs():
    """Inputs of tests/test_data/plan01_ug_capture.py:41-70 in the reference."""
    pot = o.jax_uniform_f64(42, (32, 32, 2)) * 10.0
    probe = o.make_probe(100.0, 20.0, 0.0, 0.0, 0.0, (32, 32), 0.2 * 100.0)
    return pot, probe


def test_golden_cbed_matches_reference_capture():
    ref = np.load(GOLDEN)
    pot, probe = _plan01_inputs()
    cbed, fcal = o.cbed_image(pot, 2.0, 0.2, probe[..., None], [1.0], 0.2, 100.0)
    assert np.max(np.abs(cbed - ref["cbed"])) < 1e-13
    assert fcal == pytest.approx(1.0 / (32 * 0.2))


def test_golden_stem4d_matches_reference_capture():
    ref = np.load(GOLDEN)
    pot, probe = _plan01_inputs()
    pos = np.array([[12.0, 12.0], [12.0, 20.0], [20.0, 12.0], [20.0, 20.0]])
    s4d = o.stem_4d(pot, 2.0, 0.2, probe[..., None], [1.0], 0.2, 100.0, pos, 0.2)
    assert np.max(np.abs(s4d["data"] - ref["stem4d"])) < 1e-13
    np.testing.assert_allclose(s4d["scan_positions"], pos * 0.2)


def test_golden_grad_pot_is_parseval_zero():
    """sum(cbed) is constant in V by Parseval: the reference's grad_pot is ~0."""
    ref = np.load(GOLDEN)
    assert np.max(np.abs(ref["grad_pot"])) < 1e-16
    pot, probe = _plan01_inputs()
    slab = np.stack([o.transmission_func(pot[..., s], 100.0) for s in range(2)])
    prop = o.propagation_func(32, 32, 2.0, 100.0, 0.2)
    g_t, _ = o.multislice_vjp(slab, probe, [(0, 0)], [(0.0, 0.0)], prop, np.ones((1, 32, 32)))
    g_v = o.interaction_sigma(100.0) * np.imag(g_t * np.conj(slab))
    assert np.max(np.abs(g_v)) < 1e-12


@pytest.mark.parametrize("kv,expected", [(100.0, 0.92440e-3), (200.0, 0.72884e-3), (300.0, 0.65262e-3)])
def test_interaction_sigma_published(kv, expected):
    """tests/test_ptyrodactyl/test_types/test_constants.py:173-183."""
    assert o.interaction_sigma(kv) == pytest.approx(expected, rel=1e-4)


def test_wavelength_published():
    """tests/test_ptyrodactyl/test_types/test_constants.py:239."""
    assert o.wavelength_ang(200.0) == pytest.approx(0.02507934, rel=1e-6)


def _aperture_count(n, calib, kv, mrad):
    q = o.probe_fourier_grid(n, n, calib * 100)
    return int(np.sum(q <= (mrad / 1000.0) / o.wavelength_ang(kv)))


def test_vacuum_gives_fftshifted_aperture():
    """V = 0: the pattern is the aperture indicator, whatever S, dz, chi and position."""
    n, calib, kv, mrad = 32, 0.2, 200.0, 25.0
    probe = o.make_probe(kv, mrad, 40.0, 1e-3, 0.0, (n, n), calib * 100)
    s4d = o.stem_4d(np.zeros((n, n, 3)), 1.5, calib, probe[..., None], [1.0], calib, kv,
                    [[0.0, 0.0], [1.37, -2.1]], calib)
    q = o.probe_fourier_grid(n, n, calib * 100)
    ind = np.fft.fftshift((q <= (mrad / 1000.0) / o.wavelength_ang(kv)).astype(float))
    assert np.max(np.abs(s4d["data"] - ind[None])) < 1e-12


def test_unitarity_sum_equals_aperture_count():
    rng = np.random.default_rng(0)
    n, calib, kv, mrad = 32, 0.25, 300.0, 20.0
    pot = rng.uniform(0, 20, (n, n, 4))
    probe = o.make_probe(kv, mrad, 30.0, 0.0, 0.0, (n, n), calib * 100)
    cbed, _ = o.cbed_image(pot, 1.0, calib, probe[..., None], [1.0], calib, kv)
    assert cbed.sum() == pytest.approx(_aperture_count(n, calib, kv, mrad), rel=1e-12)


def test_single_slice_reduces_to_fft_of_product():
    rng = np.random.default_rng(1)
    n = 16
    pot = rng.uniform(0, 5, (n, n, 1))
    probe = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    cbed, _ = o.cbed_image(pot, 3.0, 0.3, probe[..., None], [1.0], 0.3, 200.0)
    t = o.transmission_func(pot[..., 0], 200.0)
    assert np.allclose(cbed, np.abs(np.fft.fftshift(np.fft.fft2(probe * t))) ** 2, atol=1e-12)


def _dft_matrix(n, sign):
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n)


def test_bruteforce_dft_multislice():
    """O(N^4) DFT-matrix multislice with the propagator built from its own formula."""
    rng = np.random.default_rng(2)
    n, s_count, dz, kv, calib = 8, 3, 1.7, 120.0, 0.4
    pot = rng.uniform(0, 10, (n, n, s_count))
    probe = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    lam = o.wavelength_ang(kv)
    q = np.array([k if k < n / 2 else k - n for k in range(n)]) / (n * calib)
    prop = np.exp(-1j * np.pi * lam * dz * (q[:, None] ** 2 + q[None, :] ** 2))
    f = _dft_matrix(n, -1)
    fi = _dft_matrix(n, 1) / n
    sigma = o.interaction_sigma(kv)
    wave = probe.copy()
    for s in range(s_count):
        wave = wave * np.exp(1j * sigma * pot[..., s])
        if s != s_count - 1:
            wave = fi @ ((f @ wave @ f.T) * prop) @ fi.T
    amp = f @ wave @ f.T
    expected = np.abs(np.roll(amp, (n // 2, n // 2), axis=(0, 1))) ** 2
    cbed, _ = o.cbed_image(pot, dz, calib, probe[..., None], [1.0], calib, kv)
    assert np.allclose(cbed, expected, atol=1e-10)


def test_translation_equivariance():
    """Rolling V by k pixels equals shifting the probe position by -k*calib Angstrom."""
    rng = np.random.default_rng(3)
    n, calib, kv = 16, 0.2, 200.0
    pot = rng.uniform(0, 10, (n, n, 2))
    probe = o.make_probe(kv, 25.0, 10.0, 0.0, 0.0, (n, n), calib * 100)
    k = (3, -5)
    a = o.stem_4d(np.roll(pot, k, axis=(0, 1)), 1.0, calib, probe[..., None], [1.0], calib, kv,
                  [[0.5, 0.3]], calib)["data"]
    b = o.stem_4d(pot, 1.0, calib, probe[..., None], [1.0], calib, kv,
                  [[0.5 - k[0] * calib, 0.3 - k[1] * calib]], calib)["data"]
    assert np.allclose(a, b, atol=1e-11)


def test_window_on_tiled_slab_equals_shifted_stem():
    rng = np.random.default_rng(4)
    n, s_count = 16, 3
    t_win = np.exp(1j * rng.uniform(0, 0.5, (s_count, n, n)))
    slab = np.tile(t_win, (1, 3, 2))
    probe = o.make_probe(200.0, 25.0, 5.0, 0.0, 0.0, (n, n), 20.0)
    prop = o.propagation_func(n, n, 1.0, 200.0, 0.2)
    origins = np.array([[5, 7], [n + 2, 0]])
    shifts = np.array([[0.25, -0.4], [1.0, 2.5]])
    a = o.window_stem(slab, probe, origins, shifts, prop)
    b = o.window_stem(t_win, probe, np.zeros_like(origins), shifts + origins, prop)
    assert np.allclose(a, b, atol=1e-11)
    c = o.window_stem_batched(slab, probe, origins, shifts, prop)
    assert np.allclose(a, c, atol=1e-11)


def test_window_stem_equals_reference_stem4d_on_full_grid():
    rng = np.random.default_rng(5)
    n, calib, kv = 16, 0.3, 150.0
    pot = rng.uniform(0, 10, (n, n, 2))
    probe = o.make_probe(kv, 22.0, 0.0, 0.0, 0.0, (n, n), calib * 100)
    pos = np.array([[0.7, 1.9], [-2.0, 0.33]])
    ref = o.stem_4d(pot, 1.2, calib, probe[..., None], [1.0], calib, kv, pos, calib)["data"]
    slab = np.stack([o.transmission_func(pot[..., s], kv) for s in range(2)])
    prop = o.propagation_func(n, n, 1.2, kv, calib)
    got = o.window_stem(slab, probe, np.zeros((2, 2), int), pos / calib, prop)
    assert np.allclose(ref, got, atol=1e-12)


def _small_problem(seed=6, n=8, s_count=3, p_count=3):
    rng = np.random.default_rng(seed)
    slab = np.exp(1j * rng.uniform(0, 1.0, (s_count, 12, 10))) * rng.uniform(0.9, 1.0, (s_count, 12, 10))
    probe = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    origins = rng.integers(0, 10, (p_count, 2))
    shifts = rng.uniform(-2, 2, (p_count, 2))
    prop = o.propagation_func(n, n, 1.3, 200.0, 0.3)
    c = rng.standard_normal((p_count, n, n))
    return slab, probe, origins, shifts, prop, c


def test_vjp_matches_central_differences():
    slab, probe, origins, shifts, prop, c = _small_problem()
    g_slab, g_probe = o.multislice_vjp(slab, probe, origins, shifts, prop, c)

    def loss(sl, pr):
        return float(np.sum(c * o.window_stem(sl, pr, origins, shifts, prop)))

    h = 1e-6
    rng = np.random.default_rng(7)
    for _ in range(6):
        idx = tuple(rng.integers(0, d) for d in slab.shape)
        for direction in (1.0, 1j):
            sp, sm = slab.copy(), slab.copy()
            sp[idx] += h * direction
            sm[idx] -= h * direction
            fd = (loss(sp, probe) - loss(sm, probe)) / (2 * h)
            expected = g_slab[idx].real if direction == 1.0 else g_slab[idx].imag
            assert fd == pytest.approx(expected, rel=1e-5, abs=1e-7)
    for _ in range(4):
        idx = tuple(rng.integers(0, d) for d in probe.shape)
        for direction in (1.0, 1j):
            pp, pm = probe.copy(), probe.copy()
            pp[idx] += h * direction
            pm[idx] -= h * direction
            fd = (loss(slab, pp) - loss(slab, pm)) / (2 * h)
            expected = g_probe[idx].real if direction == 1.0 else g_probe[idx].imag
            assert fd == pytest.approx(expected, rel=1e-5, abs=1e-7)


def test_window_stem_perturbed_equals_forward_of_perturbed_slab():
    """The shared-prefix variants equal full forwards of the perturbed slab (pixels in the first, a middle and
    the last slice, so the prefix is empty, partial and complete)."""
    slab, probe, origins, shifts, prop, _ = _small_problem(seed=12)
    for index in ((0, 3, 4), (1, 11, 0), (2, 5, 9)):
        deltas = (1e-3, -1e-3, 2e-3j, 0.25 - 0.5j)
        got = o.window_stem_perturbed(slab, probe, origins, shifts, prop, index, deltas)
        for j, d in enumerate(deltas):
            sl = slab.astype(np.complex128).copy()
            sl[index] += d
            want = o.window_stem_batched(sl, probe, origins, shifts, prop)
            assert np.allclose(got[j], want, rtol=1e-12, atol=1e-14)


def test_vjp_matches_torch_autograd_twin():
    torch = pytest.importorskip("torch")
    slab, probe, origins, shifts, prop, c = _small_problem(seed=8)
    g_slab, g_probe = o.multislice_vjp(slab, probe, origins, shifts, prop, c)
    n = probe.shape[0]
    ts = torch.tensor(slab, requires_grad=True)
    tp = torch.tensor(probe, requires_grad=True)
    tprop = torch.tensor(prop)
    k = torch.tensor(np.fft.fftfreq(n) * n)
    total = 0
    for p in range(len(origins)):
        ry = torch.exp(-2j * np.pi * k * shifts[p][0] / n)
        rx = torch.exp(-2j * np.pi * k * shifts[p][1] / n)
        wave = torch.fft.ifft2(torch.fft.fft2(tp) * ry[:, None] * rx[None, :])
        rows = (origins[p][0] + np.arange(n)) % slab.shape[1]
        cols = (origins[p][1] + np.arange(n)) % slab.shape[2]
        for s in range(slab.shape[0]):
            wave = wave * ts[s][rows][:, cols]
            if s != slab.shape[0] - 1:
                wave = torch.fft.ifft2(torch.fft.fft2(wave) * tprop)
        inten = torch.abs(torch.fft.fftshift(torch.fft.fft2(wave))) ** 2
        total = total + (torch.tensor(c[p]) * inten).sum()
    total.backward()
    assert np.allclose(ts.grad.numpy(), g_slab, atol=1e-9)
    assert np.allclose(tp.grad.numpy(), g_probe, atol=1e-9)


def test_shift_vjp_matches_central_differences():
    slab, probe, origins, shifts, prop, c = _small_problem(seed=11)
    _, _, g_shift = o.multislice_vjp(slab, probe, origins, shifts, prop, c, want_shifts=True)

    def loss(sh):
        return float(np.sum(c * o.window_stem(slab, probe, origins, sh, prop)))

    h = 1e-6
    for p in range(len(origins)):
        for ax in range(2):
            sp, sm = shifts.copy(), shifts.copy()
            sp[p, ax] += h
            sm[p, ax] -= h
            fd = (loss(sp) - loss(sm)) / (2 * h)
            assert fd == pytest.approx(g_shift[p, ax], rel=1e-5, abs=1e-7)


def test_shift_vjp_matches_torch_autograd_twin():
    torch = pytest.importorskip("torch")
    slab, probe, origins, shifts, prop, c = _small_problem(seed=12)
    _, _, g_shift = o.multislice_vjp(slab, probe, origins, shifts, prop, c, want_shifts=True)
    n = probe.shape[0]
    tsh = torch.tensor(shifts, requires_grad=True)
    tprop = torch.tensor(prop)
    k = torch.tensor(np.fft.fftfreq(n) * n)
    pk = torch.fft.fft2(torch.tensor(probe))
    total = 0
    for p in range(len(origins)):
        ry = torch.exp(-2j * np.pi * k * tsh[p, 0] / n)
        rx = torch.exp(-2j * np.pi * k * tsh[p, 1] / n)
        wave = torch.fft.ifft2(pk * ry[:, None] * rx[None, :])
        win = torch.tensor(o.window_of(slab, origins[p], n))
        for s in range(slab.shape[0]):
            wave = wave * win[s]
            if s != slab.shape[0] - 1:
                wave = torch.fft.ifft2(torch.fft.fft2(wave) * tprop)
        inten = torch.abs(torch.fft.fftshift(torch.fft.fft2(wave))) ** 2
        total = total + (torch.tensor(c[p]) * inten).sum()
    total.backward()
    assert np.allclose(tsh.grad.numpy(), g_shift, rtol=1e-9, atol=1e-9)


def test_shift_vjp_vanishes_in_vacuum():
    """With T = 1 every pattern is |F(P psi0)|^2 and the shift is a Fourier phase: the intensity does not depend
    on it, so the shift gradient is zero while the probe gradient is not."""
    slab, probe, origins, shifts, prop, c = _small_problem(seed=13)
    vac = np.ones_like(slab)
    _, g_probe, g_shift = o.multislice_vjp(vac, probe, origins, shifts, prop, c, want_shifts=True)
    assert np.abs(g_probe).max() > 1e-3
    assert np.abs(g_shift).max() < 1e-9 * np.abs(g_probe).max() * probe.size


def test_jvp_vjp_adjoint_identity():
    slab, probe, origins, shifts, prop, c = _small_problem(seed=9)
    rng = np.random.default_rng(10)
    d_slab = rng.standard_normal(slab.shape) + 1j * rng.standard_normal(slab.shape)
    d_probe = rng.standard_normal(probe.shape) + 1j * rng.standard_normal(probe.shape)
    d_i = o.multislice_jvp(slab, probe, origins, shifts, prop, d_slab, d_probe)
    g_slab, g_probe = o.multislice_vjp(slab, probe, origins, shifts, prop, c)
    lhs = np.sum(c * d_i)
    rhs = np.sum(np.real(np.conj(g_slab) * d_slab)) + np.sum(np.real(np.conj(g_probe) * d_probe))
    assert lhs == pytest.approx(rhs, rel=1e-10)


def test_aberration12_reduces_to_reference_aberration():
    n, calib, kv = 32, 0.15, 300.0
    lam = o.wavelength_ang(kv)
    q = o.fftfreq(n, calib)
    qy, qx = np.meshgrid(q, q, indexing="ij")
    c1, c3_ang = 45.0, 1e-3
    coeffs = np.zeros(12)
    coeffs[0] = c1
    coeffs[5] = c3_ang * 1e7
    chi12 = o.aberration12(qy, qx, lam, coeffs)
    chi_ref = o.aberration(np.sqrt(qx**2 + qy**2), lam, c1, c3_ang, 0.0)
    assert np.allclose(chi12, chi_ref, rtol=1e-12, atol=1e-12)


def test_aberration12_astigmatism_symmetry():
    """A1 (two-fold) is invariant under a 180-degree rotation and flips sign under 90 degrees."""
    lam = o.wavelength_ang(300.0)
    coeffs = np.zeros(12)
    coeffs[1] = 20.0
    qy, qx = 0.3, 0.7
    a = o.aberration12(qy, qx, lam, coeffs)
    assert o.aberration12(-qy, -qx, lam, coeffs) == pytest.approx(a)
    assert o.aberration12(qx, -qy, lam, coeffs) == pytest.approx(-a)


# (index, radial order k in lambda |q|, fold m, prefactor of the Krivanek form chi = (2 pi / lambda) Re[...])
ABERRATION_TERMS = {
    "C1": (0, 2, 0, 1 / 2), "A1a": (1, 2, 2, 1 / 2), "A1b": (2, 2, 2, 1 / 2), "B2a": (3, 3, 1, 1.0),
    "B2b": (4, 3, 1, 1.0), "C3": (5, 4, 0, 1 / 4), "A2a": (6, 3, 3, 1 / 3), "A2b": (7, 3, 3, 1 / 3),
    "S3a": (8, 4, 2, 1.0), "S3b": (9, 4, 2, 1.0), "A3a": (10, 4, 4, 1 / 4), "A3b": (11, 4, 4, 1 / 4),
}


def _rotate(qy, qx, phi):
    """Counter-clockwise rotation of (qx, qy) by phi (omega = lambda (qx + i qy) -> omega e^{i phi})."""
    return qx * np.sin(phi) + qy * np.cos(phi), qx * np.cos(phi) - qy * np.sin(phi)


@pytest.mark.parametrize("name", list(ABERRATION_TERMS))
def test_aberration12_term_order_symmetry_prefactor_orientation(name):
    """Each coefficient alone, by the properties its name fixes: chi scales as (lambda |q|)^k (k = 2 for C1, A1;
    3 for B2, A2; 4 for C3, S3, A3); it is m-fold symmetric (C1, C3 round; B2 1-, A1 and S3 2-, A2 3-, A3
    4-fold) and changes sign under a rotation by pi/m; its maximum over the azimuth at a radius is
    (2 pi / lambda) c (lambda |q|)^k times the prefactor (1/2 C1, A1; 1/3 A2; 1/4 C3, A3; 1 B2, S3); the real
    part (suffix a) peaks on the +qx axis and the imaginary part (suffix b) at the azimuth pi / (2 m)."""
    idx, k, m, pref = ABERRATION_TERMS[name]
    lam = o.wavelength_ang(300.0)
    c = 37.0
    coeffs = np.zeros(12)
    coeffs[idx] = c
    rng = np.random.default_rng(idx)
    qy, qx = rng.uniform(-0.8, 0.8, 50), rng.uniform(-0.8, 0.8, 50)
    chi = o.aberration12(qy, qx, lam, coeffs)
    assert np.allclose(o.aberration12(1.7 * qy, 1.7 * qx, lam, coeffs), 1.7 ** k * chi, rtol=1e-12, atol=1e-12)
    if m == 0:
        assert np.allclose(o.aberration12(*_rotate(qy, qx, 0.7), lam, coeffs), chi, rtol=1e-12, atol=1e-12)
    else:
        assert np.allclose(o.aberration12(*_rotate(qy, qx, 2 * np.pi / m), lam, coeffs), chi, rtol=1e-10, atol=1e-12)
        assert np.allclose(o.aberration12(*_rotate(qy, qx, np.pi / m), lam, coeffs), -chi, rtol=1e-10, atol=1e-12)
    # azimuthal profile on a circle |q| = 0.5
    theta = np.linspace(0.0, 2 * np.pi, 7201)[:-1]
    prof = o.aberration12(0.5 * np.sin(theta), 0.5 * np.cos(theta), lam, coeffs)
    peak = (2 * np.pi / lam) * c * (lam * 0.5) ** k * pref
    assert prof.max() == pytest.approx(peak, rel=1e-12)
    if m:
        want_theta = 0.0 if name.endswith("a") else np.pi / (2 * m)
        assert o.aberration12(0.5 * np.sin(want_theta), 0.5 * np.cos(want_theta), lam, coeffs) == \
            pytest.approx(peak, rel=1e-12)


def test_lm_step_rules_closed_forms():
    """The reference's closed forms (test_solvers.py:131-226) through the oracle's step functions, on
    f(p) = 2 p with data 2 from p = 0: the GN step (lambda = 0) lands on 1; the LM step on 4 / (4 + lambda) with
    residual norm 2 lambda / (4 + lambda); on this quadratic the gain ratio is exactly 2 (the reference divides a
    reduction of |r|^2 by the model's reduction of |r|^2 / 2) and the damping drops to lambda / 3; the damping
    rule's thresholds are strict."""
    jac = np.array([[2.0]])
    r0 = np.array([2.0 * 0.0 - 2.0])
    step, _, _ = o.lm_step_dense(jac, r0, 0.0)
    assert 0.0 - step[0] == pytest.approx(1.0, abs=1e-15)
    for lam in (1e-3, 0.25, 3.0):
        step, jtj, g = o.lm_step_dense(jac, r0, lam)
        cand = 0.0 - step[0]
        assert cand == pytest.approx(4.0 / (4.0 + lam), rel=1e-15)
        r1 = 2.0 * cand - 2.0
        assert abs(r1) == pytest.approx(2.0 * lam / (4.0 + lam), rel=1e-12)
        rho = o.lm_gain_ratio(float(r0 @ r0), r1 * r1, step, g, jtj)
        assert rho == pytest.approx(2.0, rel=1e-9)
        assert o.lm_update_damping(rho, lam) == pytest.approx(lam / 3.0)
    assert o.lm_update_damping(0.5, 0.2) == 0.2
    assert o.lm_update_damping(0.1, 0.2) == pytest.approx(0.4)
    assert o.lm_update_damping(0.75, 0.2) == 0.2 and o.lm_update_damping(0.25, 0.2) == 0.2
    assert o.lm_update_damping(-1.0, 0.2) == pytest.approx(0.4)


def _small_aberration_problem(offset):
    n, calib, kv, aperture, dz = 32, 0.3, 300.0, 20.0, 1.5
    rng = np.random.default_rng(31)
    slab = np.exp(1j * rng.uniform(0.0, 0.6, (2, 40, 40)))
    origins = rng.integers(0, 40, (5, 2))
    shifts = np.zeros((5, 2))
    true = np.array([40.0, 8.0, -5.0, 60.0, 30.0, 2e3, 150.0, -90.0, 400.0, 300.0, 900.0, -700.0])
    pk = o.probe_k_from_coeffs(n, calib, kv, aperture, true)
    prop = o.propagation_func(n, n, dz, kv, calib)
    amp = o.window_amplitudes(slab, pk[None], origins, shifts, prop)[:, 0]
    data = np.fft.fftshift(np.abs(amp) ** 2, axes=(-2, -1))
    init = true * (1.0 + offset * np.linspace(-1.0, 1.0, 12))
    return dict(slab=slab, origins=origins, shifts=shifts, data=data, init=init, true=true, calib=calib, kv=kv,
                aperture=aperture, dz=dz)


def test_lm_aberration_fit_freezes_when_converged():
    """Started at the coefficients that produced the data, the residual is 0 below the tolerance: the loop
    takes no step and returns the start (test_solvers.py:198-226)."""
    pb = _small_aberration_problem(0.0)
    coeffs, hist = o.lm_aberration_fit(pb["slab"], pb["origins"], pb["shifts"], pb["data"], pb["init"], pb["calib"],
                                       pb["kv"], pb["aperture"], pb["dz"], iterations=5, tolerance=1e-8)
    assert len(hist) == 1 and hist[0] < 1e-8
    assert np.allclose(coeffs, pb["init"], rtol=1e-15, atol=0)


def test_lm_aberration_fit_recovers_known_coefficients():
    """From +-3 % offsets the FP64 loop returns the twelve coefficients that produced noise-free data (a wrong
    step sign or gain-ratio rule diverges or stalls), with a non-increasing residual history."""
    pb = _small_aberration_problem(0.03)
    coeffs, hist = o.lm_aberration_fit(pb["slab"], pb["origins"], pb["shifts"], pb["data"], pb["init"], pb["calib"],
                                       pb["kv"], pb["aperture"], pb["dz"], iterations=25, tolerance=1e-12)
    assert all(b <= a * (1 + 1e-12) for a, b in zip(hist, hist[1:]))
    assert hist[-1] < 1e-6 * hist[0]
    assert np.allclose(coeffs, pb["true"], rtol=1e-5)


def test_threefry_known_answer():
    """Random123 known-answer vector for Threefry-2x32-20 (key=0, ctr=0)."""
    x0, x1 = o.threefry2x32((0, 0), np.array([0], np.uint32), np.array([0], np.uint32))
    assert (int(x0[0]), int(x1[0])) == (0x6B200159, 0x99BA4EFE)


def _plan01_atoms():
    """Atom inputs of tests/test_data/plan01_ug_capture.py:71-87 in the reference."""
    pot, probe = _plan01_inputs()
    coords = np.array([[2.0, 2.5, 0.25], [4.5, 5.0, 2.25], [3.0, 4.0, 1.25]])
    return pot, probe, coords, np.zeros(3, int), np.array([[0.0, 2.0], [2.0, 4.0]]), pot[..., 0][None]


def test_golden_stem4d_sharded_matches_reference_capture():
    ref = np.load(GOLDEN)
    pot, probe, coords, types, bounds, kern = _plan01_atoms()
    pos = np.array([[12.0, 12.0], [12.0, 20.0], [20.0, 12.0], [20.0, 20.0]]) * 0.2
    got = o.stem4d_from_atoms(probe[..., None], [1.0], coords, types, bounds, kern, 100.0, 0.2, pos)
    assert np.max(np.abs(got - ref["stem4d_sharded"])) < 1e-12


# ---------------------------------------------------------------------------
# Non-square and odd grids: the windowed oracle on h x w windows
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("shape", [(15, 32), (7, 5), (12, 9)])
def test_window_stem_equals_reference_stem4d_on_non_square_grids(shape):
    """The windowed restatement on an h x w window = the reference-level stem_4d (shape-generic numpy) at the
    same positions; odd edges exercise the fftfreq sign convention of the shift ramp."""
    rng = np.random.default_rng(21)
    h, w = shape
    calib, kv = 0.25, 200.0
    pot = rng.uniform(0, 10, (h, w, 3))
    probe = o.make_probe(kv, 24.0, 4.0, 0.0, 0.0, (h, w), calib * 100)
    pos = np.array([[0.7, 1.9], [-2.0, 0.33], [3.1, -0.45]])
    ref = o.stem_4d(pot, 1.1, calib, probe[..., None], [1.0], calib, kv, pos, calib)["data"]
    slab = np.stack([o.transmission_func(pot[..., s], kv) for s in range(3)])
    prop = o.propagation_func(h, w, 1.1, kv, calib)
    got = o.window_stem(slab, probe, np.zeros((3, 2), int), pos / calib, prop)
    assert got.shape == (3, h, w)
    assert np.allclose(ref, got, atol=1e-12)


def test_bruteforce_dft_multislice_non_square():
    """DFT-matrix multislice on a 6 x 9 grid, propagator from its own formula, shift applied as a DFT phase."""
    rng = np.random.default_rng(22)
    h, w, s_count, dz, kv, calib = 6, 9, 3, 1.3, 160.0, 0.35
    slab = np.exp(1j * rng.uniform(0, 1.0, (s_count, h, w)))
    probe = rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))
    shift = (0.6, -1.3)
    lam = o.wavelength_ang(kv)
    ky = np.array([k if k < (h + 1) // 2 else k - h for k in range(h)])
    kx = np.array([k if k < (w + 1) // 2 else k - w for k in range(w)])
    prop = np.exp(-1j * np.pi * lam * dz * ((ky / (h * calib))[:, None] ** 2 + (kx / (w * calib))[None, :] ** 2))
    fy, fx = _dft_matrix(h, -1), _dft_matrix(w, -1)
    fyi, fxi = _dft_matrix(h, 1) / h, _dft_matrix(w, 1) / w
    ramp = np.exp(-2j * np.pi * (ky[:, None] * shift[0] / h + kx[None, :] * shift[1] / w))
    wave = fyi @ ((fy @ probe @ fx.T) * ramp) @ fxi.T
    for s in range(s_count):
        wave = wave * slab[s]
        if s != s_count - 1:
            wave = fyi @ ((fy @ wave @ fx.T) * prop) @ fxi.T
    amp = fy @ wave @ fx.T
    expected = np.abs(np.roll(amp, (h // 2, w // 2), axis=(0, 1))) ** 2
    got = o.window_stem(slab, probe, np.zeros((1, 2), int), np.array([shift]),
                        o.propagation_func(h, w, dz, kv, calib))[0]
    assert np.allclose(got, expected, atol=1e-10)


def _small_problem_hw(seed, h, w, s_count=3, p_count=3):
    rng = np.random.default_rng(seed)
    slab = np.exp(1j * rng.uniform(0, 1.0, (s_count, h + 3, w + 4))) * rng.uniform(0.9, 1.0, (s_count, h + 3, w + 4))
    probe = rng.standard_normal((h, w)) + 1j * rng.standard_normal((h, w))
    origins = rng.integers(0, 12, (p_count, 2))
    shifts = rng.uniform(-2, 2, (p_count, 2))
    prop = o.propagation_func(h, w, 1.3, 200.0, 0.3)
    c = rng.standard_normal((p_count, h, w))
    return slab, probe, origins, shifts, prop, c


def test_vjp_non_square_matches_central_differences():
    slab, probe, origins, shifts, prop, c = _small_problem_hw(23, 5, 7)
    g_slab, g_probe, g_shift = o.multislice_vjp(slab, probe, origins, shifts, prop, c, want_shifts=True)

    def loss(sl, pr, sh):
        return float(np.sum(c * o.window_stem(sl, pr, origins, sh, prop)))

    eps = 1e-6
    rng = np.random.default_rng(24)
    for _ in range(6):
        idx = tuple(rng.integers(0, d) for d in slab.shape)
        for direction in (1.0, 1j):
            sp, sm = slab.copy(), slab.copy()
            sp[idx] += eps * direction
            sm[idx] -= eps * direction
            fd = (loss(sp, probe, shifts) - loss(sm, probe, shifts)) / (2 * eps)
            expected = g_slab[idx].real if direction == 1.0 else g_slab[idx].imag
            assert fd == pytest.approx(expected, rel=1e-5, abs=1e-7)
    for _ in range(4):
        idx = tuple(rng.integers(0, d) for d in probe.shape)
        for direction in (1.0, 1j):
            pp, pm = probe.copy(), probe.copy()
            pp[idx] += eps * direction
            pm[idx] -= eps * direction
            fd = (loss(slab, pp, shifts) - loss(slab, pm, shifts)) / (2 * eps)
            expected = g_probe[idx].real if direction == 1.0 else g_probe[idx].imag
            assert fd == pytest.approx(expected, rel=1e-5, abs=1e-7)
    for p in range(len(origins)):
        for ax in range(2):
            sp, sm = shifts.copy(), shifts.copy()
            sp[p, ax] += eps
            sm[p, ax] -= eps
            fd = (loss(slab, probe, sp) - loss(slab, probe, sm)) / (2 * eps)
            assert fd == pytest.approx(g_shift[p, ax], rel=1e-5, abs=1e-7)
