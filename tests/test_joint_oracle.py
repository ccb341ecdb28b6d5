"""CPU pins of the joint aberration + object LM oracle (``oracle/joint_oracle.py``)."""

import numpy as np
import pytest

from oracle import joint_oracle as jo
from oracle import ms_oracle as o


def _problem(seed=3, n=16, s_count=3, p_count=5, hw=(20, 24)):
    rng = np.random.default_rng(seed)
    obj = np.exp(1j * rng.uniform(0, 0.8, hw)) * rng.uniform(0.9, 1.0, hw)
    probe = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    origins = np.stack([rng.integers(0, hw[0], p_count), rng.integers(0, hw[1], p_count)], 1)
    shifts = rng.uniform(-2, 2, (p_count, 2))
    prop = o.propagation_func(n, n, 1.3, 200.0, 0.3)
    return obj, s_count, probe, origins, shifts, prop, rng


def test_batched_forward_jvp_vjp_equal_the_per_position_oracle():
    obj, s_count, probe, origins, shifts, prop, rng = _problem()
    slab = np.repeat(obj[None], s_count, axis=0)
    pk = np.fft.fft2(probe)
    got = jo.patterns(obj, s_count, pk, origins, shifts, prop, workers=1)
    want = o.window_stem(slab, probe, origins, shifts, prop)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    d_obj = rng.standard_normal(obj.shape) + 1j * rng.standard_normal(obj.shape)
    d_slab = np.repeat(d_obj[None], s_count, axis=0)
    got = jo.object_jvp(obj, s_count, pk, origins, shifts, prop, d_obj, workers=1)
    want = o.multislice_jvp(slab, probe, origins, shifts, prop, d_slab, np.zeros_like(probe))
    assert np.allclose(got, want, rtol=1e-11, atol=1e-11 * np.abs(want).max())
    g = rng.standard_normal(got.shape)
    got = jo.object_vjp(obj, s_count, pk, origins, shifts, prop, g, workers=1)
    want_slab, _ = o.multislice_vjp(slab, probe, origins, shifts, prop, g)
    want = want_slab.sum(axis=0)
    assert np.allclose(got, want, rtol=1e-11, atol=1e-11 * np.abs(want).max())


def test_object_jvp_matches_central_differences_and_vjp():
    obj, s_count, probe, origins, shifts, prop, rng = _problem(seed=4)
    pk = np.fft.fft2(probe)
    d_obj = rng.standard_normal(obj.shape) + 1j * rng.standard_normal(obj.shape)
    h = 1e-6
    fd = (jo.patterns(obj + h * d_obj, s_count, pk, origins, shifts, prop, 1)
          - jo.patterns(obj - h * d_obj, s_count, pk, origins, shifts, prop, 1)) / (2 * h)
    jv = jo.object_jvp(obj, s_count, pk, origins, shifts, prop, d_obj, 1)
    assert np.linalg.norm(jv - fd) < 1e-6 * np.linalg.norm(fd)
    u = rng.standard_normal(jv.shape)
    lhs = np.sum(u * jv)
    rhs = np.real(np.vdot(jo.object_vjp(obj, s_count, pk, origins, shifts, prop, u, 1), d_obj))
    assert lhs == pytest.approx(rhs, rel=1e-11)


def test_cg_fixed_solves_a_damped_system():
    rng = np.random.default_rng(8)
    a = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    m = a.conj().T @ a + 0.5 * np.eye(6)
    b = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    x = jo.cg_fixed(lambda v: m @ v, b, 12)
    assert np.allclose(x, np.linalg.solve(m, b), rtol=1e-8)
    # one step: x = (b.b / b.Ab) b, the reference's first CG iterate
    x1 = jo.cg_fixed(lambda v: m @ v, b, 1)
    assert np.allclose(x1, np.vdot(b, b).real / np.vdot(b, m @ b).real * b)


def test_joint_fit_reduces_the_residual_monotonically():
    n, calib, kv, aperture, dz, s_count = 32, 0.3, 300.0, 20.0, 1.5, 2
    rng = np.random.default_rng(31)
    obj_true = np.exp(1j * rng.uniform(0.0, 0.6, (48, 48)))
    lin = np.arange(0, 17, 4)
    iy, ix = np.meshgrid(lin, lin, indexing="ij")
    origins = np.stack([iy.ravel(), ix.ravel()], 1)
    shifts = np.zeros((origins.shape[0], 2))
    true = np.array([40.0, 8.0, -5.0, 60.0, 30.0, 2e3, 150.0, -90.0, 400.0, 300.0, 900.0, -700.0])
    pk = o.probe_k_from_coeffs(n, calib, kv, aperture, true)
    prop = o.propagation_func(n, n, dz, kv, calib)
    data = jo.patterns(obj_true, s_count, pk, origins, shifts, prop, 1)
    init = true * (1.0 + 0.05 * np.linspace(-1.0, 1.0, 12))
    obj0 = obj_true * (1.0 + 0.03 * rng.uniform(-1, 1, obj_true.shape))
    coeffs, obj, hist = jo.joint_fit(obj0, s_count, origins, shifts, data, init, calib, kv, aperture, dz,
                                     iterations=8, cg_iterations=4, workers=1)
    norms = [h[1] for h in hist]
    assert all(b <= a * (1 + 1e-12) for a, b in zip(norms, norms[1:]))
    assert all(h[1] <= h[0] * (1 + 1e-12) for h in hist)
    assert norms[-1] < 0.15 * norms[0]
    # threads over position chunks give the same run
    c2, obj2, hist2 = jo.joint_fit(obj0, s_count, origins, shifts, data, init, calib, kv, aperture, dz,
                                   iterations=2, cg_iterations=4, workers=3)
    assert np.allclose(np.array(hist2), np.array(hist[:3]), rtol=1e-10)
