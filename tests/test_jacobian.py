"""CPU tests of the GN/LM/CG solvers against the reference's closed forms
(tests/test_ptyrodactyl/test_jacobian/test_solvers.py:53-226) and of the
aberration basis against the oracle."""

import numpy as np
import pytest
import torch

from oracle import ms_oracle as o
from ptyrodactyl_amd import jacobian as jac


def test_tree_dot_real_hermitian():
    a = {"c": torch.tensor([1.0 + 2.0j, -3.0 + 0.5j]), "r": torch.tensor([2.0, -1.0])}
    b = {"c": torch.tensor([0.5 - 1.0j, 2.0 + 4.0j]), "r": torch.tensor([-3.0, 5.0])}
    expected = np.real(np.vdot(a["c"].numpy(), b["c"].numpy()) + np.vdot(a["r"].numpy(), b["r"].numpy()))
    assert float(jac.tree_dot(a, b)) == pytest.approx(expected)
    assert float(jac.tree_dot(a, a)) > 0.0


def test_cg_one_step_diagonal():
    rhs = torch.tensor([2.0, 4.0], dtype=torch.float64)
    sol, it = jac.conjugate_gradient(lambda v: 2.0 * v, rhs, torch.zeros_like(rhs), max_iterations=1,
                                     tolerance=0.0)
    assert it == 1
    assert torch.allclose(sol, rhs / 2.0)


def test_cg_complex_hermitian():
    d = torch.tensor([2.0, 5.0], dtype=torch.float64)
    rhs = torch.tensor([2.0 + 4.0j, 10.0 - 5.0j], dtype=torch.complex128)
    sol, it = jac.conjugate_gradient(lambda v: d * v, rhs, torch.zeros_like(rhs), max_iterations=4,
                                     tolerance=1e-12)
    assert it <= 4
    assert torch.allclose(sol, rhs / d)


def test_gauss_newton_one_step():
    params, state = jac.gauss_newton_solve(lambda p: 2.0 * p, torch.tensor([0.0], dtype=torch.float64),
                                           torch.tensor([2.0], dtype=torch.float64), max_iterations=4,
                                           tolerance=1e-10, cg_max_iterations=1, cg_tolerance=0.0)
    assert torch.allclose(params, torch.tensor([1.0], dtype=torch.float64), atol=1e-12)
    assert float(state.residual_norm) < 1e-10
    assert state.iteration == 1


def test_levenberg_marquardt_first_step():
    lam = 1e-3
    params, state = jac.levenberg_marquardt_solve(lambda p: 2.0 * p, torch.tensor([0.0], dtype=torch.float64),
                                                  torch.tensor([2.0], dtype=torch.float64), max_iterations=4,
                                                  tolerance=1e-3, damping_init=lam, cg_max_iterations=1,
                                                  cg_tolerance=0.0)
    expected = 4.0 / (4.0 + lam)
    assert float(params[0]) == pytest.approx(expected, abs=1e-12)
    assert float(state.residual_norm) == pytest.approx(abs(2.0 * expected - 2.0), abs=1e-12)
    assert float(state.damping) == pytest.approx(lam / 3.0)
    assert state.iteration == 1


def test_converged_states_freeze():
    p0 = torch.tensor([1.0], dtype=torch.float64)
    data = torch.tensor([2.0], dtype=torch.float64)
    gp, gs = jac.gauss_newton_solve(lambda p: 2.0 * p, p0, data, max_iterations=4, tolerance=1e-10,
                                    cg_max_iterations=1)
    lp, ls = jac.levenberg_marquardt_solve(lambda p: 2.0 * p, p0, data, max_iterations=4, tolerance=1e-10,
                                           damping_init=0.25, cg_max_iterations=1)
    assert torch.equal(gp, p0) and gs.iteration == 0 and float(gs.residual_norm) == 0.0
    assert torch.equal(lp, p0) and ls.iteration == 0 and float(ls.damping) == 0.25


def test_lm_on_complex_parameters_uses_torch_gradient_convention():
    """f(z) = |z|^2 fitted to 4: the LM must converge on a complex leaf (conjugation handled)."""
    params, state = jac.levenberg_marquardt_solve(lambda z: (z * z.conj()).real, torch.tensor([1.0 + 1.0j]),
                                                  torch.tensor([4.0], dtype=torch.float64), max_iterations=30,
                                                  tolerance=1e-10)
    assert abs(float(torch.abs(params[0])) - 2.0) < 1e-6


def test_aberration_basis_matches_oracle():
    n, calib, kv = 64, 0.15, 300.0
    b = jac.aberration_basis(n, calib, kv).numpy()
    ref = o.aberration_basis12(n, calib, kv)
    assert np.allclose(b, ref, rtol=1e-12, atol=1e-9 * np.abs(ref).max())


def test_probe_k_closed_form_matches_fft_of_probe():
    n, calib, kv, mrad = 64, 0.15, 300.0, 20.0
    coeffs = np.array([50.0, 12.0, -8.0, 150.0, -90.0, 1.0e4, 400.0, 250.0, -1200.0, 800.0, 3000.0, -2000.0])
    basis = jac.aberration_basis(n, calib, kv)
    chi = torch.einsum("j,jyx->yx", torch.tensor(coeffs), basis)
    pk = jac.probe_k_from_chi(jac.aperture_mask(n, calib, kv, mrad), chi).numpy()
    ref = o.probe_k_from_coeffs(n, calib, kv, mrad, coeffs)
    assert np.allclose(pk, ref, atol=1e-10)


def test_lanczos_diagonal_operator_closed_forms():
    """test_solvers.py:229-269: diag(1, 2, 4) from ones(3): alpha_0 = 7/3, beta_0 = sqrt(14)/3 and the 3-step
    tridiagonal matrix has the operator's spectrum."""
    diagonal = torch.tensor([1.0, 2.0, 4.0], dtype=torch.float64)
    alpha, beta = jac.lanczos_tridiagonal(lambda v: diagonal * v, torch.ones(3, dtype=torch.float64), 3)
    assert alpha.shape == (3,) and beta.shape == (2,)
    assert float(alpha[0]) == pytest.approx(7.0 / 3.0, rel=1e-14)
    assert float(beta[0]) == pytest.approx(14.0 ** 0.5 / 3.0, rel=1e-14)
    t = torch.diag(alpha) + torch.diag(beta, 1) + torch.diag(beta, -1)
    assert torch.allclose(torch.linalg.eigvalsh(t), diagonal, atol=1e-10)


def test_lanczos_extreme_eigenvalues_of_jtj():
    """On a J^T J operator of a random 40x12 Jacobian the 12-step tridiagonal matrix reproduces its spectrum."""
    g = torch.Generator().manual_seed(3)
    jm = torch.randn(40, 12, generator=g, dtype=torch.float64)
    alpha, beta = jac.lanczos_tridiagonal(lambda v: jm.T @ (jm @ v), torch.randn(12, generator=g,
                                                                                   dtype=torch.float64), 12)
    t = torch.diag(alpha) + torch.diag(beta, 1) + torch.diag(beta, -1)
    want = torch.linalg.eigvalsh(jm.T @ jm)
    got = torch.linalg.eigvalsh(t)
    assert float(got[-1]) == pytest.approx(float(want[-1]), rel=1e-8)
    assert float(got[0]) == pytest.approx(float(want[0]), rel=1e-6)
