"""Pins of the reconstructor oracle (``oracle/recon_oracle.py``), CPU only.

* The reference's fixed-seed reconstructor scalars
  (tests/test_ptyrodactyl/test_multislice/test_multislice_recon.py:178-285)
  regenerated from its inputs (:35-61: jax.random.split / normal of
  PRNGKey(11), restated here) through the oracle loop.  They exercise Optax
  adam, the MSE loss, the conjugated-gradient convention, the potential,
  probe-mode and position gradients, mode weights, snapshot slot 0 and the
  separate position learning rate.  The reference compares bit patterns of
  its own runs; here XLA's FFT and erfinv roundings differ from numpy's, so
  the bound is 1e-12 relative (+ 2e-15 absolute); the variants' pinned values
  differ from each other by 1e-7 relative.
* Optax rules on constant gradients, where the published rules have closed
  forms: adam's bias correction makes every step -lr g / (|g| + eps);
  adagrad's k-th step is -lr g / sqrt(k |g|^2 + eps); rmsprop's is
  -lr g / sqrt((1 - 0.9^k) |g|^2 + eps).
* Loss gradients and the stem_4d gradients against central differences.
"""

import numpy as np
import pytest

from oracle import recon_oracle as r

ATOL_PINNED = 2e-15
RTOL_PINNED = 1e-12


def _tiny_inputs():
    keys = r.jax_split(r.jax_key(11), 5)
    shape = (4, 4)
    pot = (r.jax_normal_f64(keys[0], shape) * 0.01).astype(np.complex128)
    beam = (r.jax_normal_f64(keys[1], shape) + 1j * r.jax_normal_f64(keys[2], shape)).astype(np.complex128)
    positions = np.array([[0.0, 0.0], [1.0, 0.0]])
    return np.ones((2, 4, 4)), positions, pot, beam


def _run(pot, modes, weights, positions, data, lr, pos_lr, update_positions):
    return r.reconstruct(data, positions, 1.0, 80.0, pot, modes, weights, 1.0, 1.0, 1, 2, lr, pos_lr, "mse", "adam",
                         update_positions)


def _close(actual, expected):
    assert abs(complex(actual) - complex(expected)) <= ATOL_PINNED + RTOL_PINNED * abs(complex(expected)), \
        (actual, expected)


def test_single_slice_ptychography_pinned_scalars():
    data, positions, pot, beam = _tiny_inputs()
    out = _run(pot, beam[..., None], [1.0], positions, data, 0.001, 0.001, False)
    assert out[3].shape == (4, 4, 2) and out[4].shape == (4, 4, 1, 2)
    _close(out[0][0, 0], 0.02866302842954916 + 0j)
    _close(out[1][0, 0, 0], -1.4996974690426799 - 1.4814616332191217j)
    _close(out[3][0, 0, 0], 0.027663112105862656 + 0j)


def test_single_slice_poscorrected_pinned_scalars():
    data, positions, pot, beam = _tiny_inputs()
    out = _run(pot, beam[..., None], [1.0], positions, data, 0.001, 0.001, True)
    assert out[2].shape == (2, 2) and out[5].shape == (2, 2, 2)
    _close(out[0][0, 0], 0.02866303668627853 + 0j)
    _close(out[1][0, 0, 0], -1.4996974690390232 - 1.4814616332263657j)
    _close(out[2][0, 0], 0.002001176434824531)


def test_single_slice_multi_modal_pinned_scalars():
    data, positions, pot, beam = _tiny_inputs()
    modes = np.stack([beam, beam * (0.5 + 0.1j)], axis=-1)
    out = _run(pot, modes, [0.7, 0.3], positions, data, 0.001, 0.001, True)
    assert out[4].shape == (4, 4, 2, 2)
    _close(out[0][0, 0], 0.0286630285066058 + 0j)
    _close(out[1][0, 0, 0], -1.4996968965393802 - 1.481462648068354j)
    _close(out[2][0, 0], 0.0020011719351587594)


def test_multi_slice_multi_modal_pinned_scalars():
    data, positions, pot, beam = _tiny_inputs()
    out = _run(pot, beam[..., None], [1.0], positions, data, 0.001, 0.01, True)
    _close(out[0][0, 0], 0.028663098749764054 + 0j)
    _close(out[1][0, 0, 0], -1.4996974690074887 - 1.4814616332889121j)
    _close(out[2][0, 0], 0.02000915293273863)


@pytest.mark.parametrize("complex_params", [False, True])
def test_optax_rules_on_constant_gradients(complex_params):
    rng = np.random.default_rng(3)
    g = rng.standard_normal(64) * np.logspace(-3, 2, 64)
    if complex_params:
        g = g + 1j * rng.standard_normal(64) * np.logspace(2, -3, 64)
    a2 = np.abs(g) ** 2
    lr = 0.01
    for name in ("adam", "adagrad", "rmsprop"):
        rule = r.OptaxRule(name, lr)
        state = rule.init(g)
        for k in range(1, 6):
            u, state = rule.update(g, state)
            if name == "adam":
                expect = -lr * g / (np.abs(g) + 1e-8)
            elif name == "adagrad":
                expect = -lr * g / np.sqrt(k * a2 + 1e-8)
            else:
                expect = -lr * g / np.sqrt((1 - 0.9 ** k) * a2 + 1e-8)
            np.testing.assert_allclose(u, expect, rtol=1e-12, err_msg=f"{name} step {k}")


def test_optax_rejects_unknown_name():
    with pytest.raises(ValueError, match="Unknown optimizer"):
        r.OptaxRule("sgd", 0.1)


@pytest.mark.parametrize("mode", ["mae", "mse", "rmse"])
def test_loss_gradient_matches_central_differences(mode):
    rng = np.random.default_rng(4)
    model = rng.standard_normal((2, 3, 3))
    data = rng.standard_normal((2, 3, 3))
    lt = r.LossType(mode)
    g = r.loss_grad(model, data, lt)
    h = 1e-6
    for idx in [(0, 0, 0), (1, 2, 1), (0, 1, 2)]:
        mp, mm = model.copy(), model.copy()
        mp[idx] += h
        mm[idx] -= h
        fd = (r.reduce_loss(mp, data, lt) - r.reduce_loss(mm, data, lt)) / (2 * h)
        assert fd == pytest.approx(g[idx], rel=1e-6, abs=1e-9)


def test_mae_gradient_at_zero_residual_is_plus_one():
    """jax.grad(abs)(0) = 1: the reference descends an MAE loss from a zero residual."""
    g = r.loss_grad(np.zeros((1, 2, 2)), np.zeros((1, 2, 2)), r.LossType.MAE)
    assert np.all(g == 0.25)


def test_stem4d_value_and_grad_matches_central_differences():
    rng = np.random.default_rng(5)
    n, p_count = 8, 3
    pot = rng.uniform(-1, 1, (n, n)) * 20
    modes = (rng.standard_normal((n, n, 2)) + 1j * rng.standard_normal((n, n, 2)))
    weights = np.array([0.6, 0.4])
    positions = rng.uniform(-1, 1, (p_count, 2))
    data = rng.uniform(0, 5, (p_count, n, n))
    args = (0.5, 0.5, 0.5, 1.2, 150.0, data, r.LossType.MSE)

    def loss(pt, md, ps):
        return r.stem4d_value_and_grad(pt, md, weights, ps, *args)[0]

    _, g_pot, g_modes, g_pos = r.stem4d_value_and_grad(pot, modes, weights, positions, *args)
    h = 1e-3  # the phase moves by sigma h ~ 1e-6 rad
    for idx in [(0, 0), (3, 5), (7, 2)]:
        pp, pm = pot.copy(), pot.copy()
        pp[idx] += h
        pm[idx] -= h
        assert (loss(pp, modes, positions) - loss(pm, modes, positions)) / (2 * h) == pytest.approx(
            g_pot[idx], rel=1e-5, abs=1e-9)
    h = 1e-5
    for idx in [(1, 1, 0), (4, 6, 1)]:
        for direction, part in ((1.0, np.real), (1j, np.imag)):
            mp, mm = modes.copy(), modes.copy()
            mp[idx] += h * direction
            mm[idx] -= h * direction
            fd = (loss(pot, mp, positions) - loss(pot, mm, positions)) / (2 * h)
            assert fd == pytest.approx(part(g_modes[idx]), rel=1e-5, abs=1e-9)
    for p in range(p_count):
        for ax in range(2):
            sp, sm = positions.copy(), positions.copy()
            sp[p, ax] += h
            sm[p, ax] -= h
            fd = (loss(pot, modes, sp) - loss(pot, modes, sm)) / (2 * h)
            assert fd == pytest.approx(g_pos[p, ax], rel=1e-5, abs=1e-9)


def test_promote_to_complex_keeps_width():
    assert r.promote_to_complex(np.zeros(2, np.float32)).dtype == np.complex64
    assert r.promote_to_complex(np.zeros(2, np.complex64)).dtype == np.complex64
    assert r.promote_to_complex(np.zeros(2, np.float64)).dtype == np.complex128
