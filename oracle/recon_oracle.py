"""CPU oracle for the reconstructors (Optax loops over the 4D-STEM gradient).

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` may import this module, and only
as the checker.  The product reconstructors in ``ptyrodactyl_amd.recon``
never call into it.

A float64 / complex128 numpy restatement of
``src/ptyrodactyl/multislice/multislice_recon.py`` (paths relative to the
reference repository root):

* ``reduce_loss`` / ``loss_grad``: ``_reduce_loss`` ``:127-163`` and the
  derivative ``jax.grad`` takes of it (``|x|'`` at 0 is +1 in JAX's abs rule).
* ``OptaxRule``: optax 0.2.3 (pinned in the reference's ``uv.lock``) adam,
  adagrad and rmsprop as configured in ``OPTIMIZERS`` ``:70-83``, including
  Optax's complex handling (second moments of ``abs_sq(g)``).
* ``stem4d_value_and_grad``: the loss of ``stem_4d`` and its gradient
  w.r.t. the real potential, the probe modes and the positions, in the
  convention the reference passes to Optax (``conj(jax.grad)``,
  ``_apply_optimizer_step`` ``:166-210``); built on ``ms_oracle``'s
  forward and vector-Jacobian product.
* ``reconstruct``: the loop shared by the four reconstructors
  (``:242-667, 670-1220, 1223-1736, 1739-2227``): one update, snapshot slot
  0, then iterations 1..N-1 with a snapshot at ``ii // save_every`` when
  ``ii % save_every == 0`` (updates out of range dropped, as JAX's scatter
  does).
* ``jax_split`` / ``jax_normal_f64``: ``jax.random.split`` and
  ``jax.random.normal`` (jax 0.9.2, partitionable Threefry) so that the
  reference's fixed-seed inputs of
  ``tests/test_ptyrodactyl/test_multislice/test_multislice_recon.py:35-61``
  can be regenerated.

Parity pins (``tests/test_recon_oracle.py``): the pinned reconstructor
scalars of ``test_multislice_recon.py:178-285``; closed forms of the first
Optax steps; the loss gradients against central differences; the
descent test with a monkeypatched forward (``test_multislice_recon.py:64-129``).
"""

from __future__ import annotations

from enum import Enum

import numpy as np

from . import ms_oracle as o

# ---------------------------------------------------------------------------
# jax.random restatement (key splitting and normal variates)
# ---------------------------------------------------------------------------


def jax_key(seed: int):
    """``jax.random.PRNGKey(seed)``: (seed >> 32, seed & 0xffffffff)."""
    return (np.uint32((seed >> 32) & 0xFFFFFFFF), np.uint32(seed & 0xFFFFFFFF))


def jax_split(key, num: int):
    """``jax.random.split(key, num)`` with partitionable Threefry: key i = threefry(key, (0, i))."""
    b0, b1 = o.threefry2x32(key, np.zeros(num, np.uint32), np.arange(num, dtype=np.uint32))
    return [(b0[i], b1[i]) for i in range(num)]


def _bits64(key, count: int):
    idx = np.arange(count, dtype=np.uint64)
    hi = (idx >> np.uint64(32)).astype(np.uint32)
    lo = (idx & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    b0, b1 = o.threefry2x32(key, hi, lo)
    return (b0.astype(np.uint64) << np.uint64(32)) | b1.astype(np.uint64)


def jax_uniform_key_f64(key, shape, minval=0.0, maxval=1.0):
    """``jax.random.uniform(key, shape, float64, minval, maxval)``."""
    count = int(np.prod(shape))
    one = np.array(1.0, np.float64).view(np.uint64)
    floats = ((_bits64(key, count) >> np.uint64(12)) | one).view(np.float64) - 1.0
    return np.maximum(minval, (floats * (maxval - minval) + minval).reshape(shape))


def jax_normal_f64(key, shape):
    """``jax.random.normal(key, shape, float64)`` = sqrt(2) erfinv(U(nextafter(-1, 0), 1))."""
    from scipy.special import erfinv
    lo = np.nextafter(-1.0, 0.0)
    return np.sqrt(2.0) * erfinv(jax_uniform_key_f64(key, shape, lo, 1.0))


# ---------------------------------------------------------------------------
# Losses (multislice_recon.py:127-163)
# ---------------------------------------------------------------------------


class LossType(str, Enum):
    MAE = "mae"
    MSE = "mse"
    RMSE = "rmse"


def reduce_loss(model, data, mode: LossType) -> float:
    r = np.asarray(model, np.float64) - np.asarray(data, np.float64)
    if mode is LossType.MAE:
        return float(np.mean(np.abs(r)))
    if mode is LossType.MSE:
        return float(np.mean(r * r))
    return float(np.sqrt(np.mean(r * r)))


def loss_grad(model, data, mode: LossType):
    """dL/dmodel of ``reduce_loss``; d|r|/dr = +1 at r = 0 as in JAX's abs derivative."""
    r = np.asarray(model, np.float64) - np.asarray(data, np.float64)
    count = r.size
    if mode is LossType.MAE:
        return np.where(r >= 0, 1.0, -1.0) / count
    if mode is LossType.MSE:
        return 2.0 * r / count
    rmse = np.sqrt(np.mean(r * r))
    return r / (count * rmse)


# ---------------------------------------------------------------------------
# Optax 0.2.3 rules (multislice_recon.py:70-83)
# ---------------------------------------------------------------------------


class OptaxRule:
    """adam / adagrad / rmsprop as the reference configures them, on real or complex arrays."""

    def __init__(self, name: str, learning_rate: float):
        if name not in ("adam", "adagrad", "rmsprop"):
            raise ValueError(f"Unknown optimizer: {name}")
        self.name = name
        self.lr = float(learning_rate)

    def init(self, params):
        p = np.asarray(params)
        return {"mu": np.zeros_like(p), "nu": np.zeros(p.shape, np.float64), "count": 0}

    def update(self, grads, state):
        """Returns (updates, new_state); ``grads`` is what Optax receives (conj of jax.grad)."""
        g = np.asarray(grads)
        g2 = (np.conj(g) * g).real if np.iscomplexobj(g) else g * g
        count = state["count"] + 1
        if self.name == "adam":
            b1, b2, eps = 0.9, 0.999, 1e-8
            mu = (1 - b1) * g + b1 * state["mu"]
            nu = (1 - b2) * g2 + b2 * state["nu"]
            mu_hat = mu / (1 - b1 ** count)
            nu_hat = nu / (1 - b2 ** count)
            u = mu_hat / (np.sqrt(nu_hat) + eps)
            new = {"mu": mu, "nu": nu, "count": count}
        elif self.name == "adagrad":
            eps = 1e-8
            nu = g2 + state["nu"]
            inv = np.where(nu > 0, 1.0 / np.sqrt(nu + eps), 0.0)
            u = inv * g
            new = {"mu": state["mu"], "nu": nu, "count": count}
        else:
            decay, eps = 0.9, 1e-8
            nu = (1 - decay) * g2 + decay * state["nu"]
            u = g / np.sqrt(nu + eps)
            new = {"mu": state["mu"], "nu": nu, "count": count}
        return -self.lr * u, new


def promote_to_complex(values):
    """``_promote_to_complex`` (``:213-239``): float32/complex64 -> complex64, else complex128."""
    v = np.asarray(values)
    return v.astype(np.result_type(v.dtype, np.complex64))


# ---------------------------------------------------------------------------
# stem_4d loss and gradient (the jax.value_and_grad of the reconstructors)
# ---------------------------------------------------------------------------


def stem4d_value_and_grad(pot, modes, weights, positions, det_calib, pot_calib, beam_calib, slice_thickness,
                          voltage_kv, data, mode: LossType):
    """Loss of ``stem_4d`` on one real potential slice and its gradients.

    pot (H, W) real; modes (H, W, M) complex; positions (P, 2) in the units
    ``stem_4d`` hands to ``shift_beam_fourier`` (Angstrom, shift = pos /
    det_calib pixels).  Returns (loss, g_pot (H, W) real, g_modes (H, W, M)
    complex, g_pos (P, 2)) with complex gradients as dL/dRe + i dL/dIm.
    """
    pot = np.asarray(pot, np.float64)
    modes = np.asarray(modes, np.complex128)
    weights = np.asarray(weights, np.float64)
    weights = weights / weights.sum()
    positions = np.atleast_2d(np.asarray(positions, np.float64))
    n = pot.shape[0]
    calib = min(float(pot_calib), float(beam_calib))
    prop = o.propagation_func(n, pot.shape[1], slice_thickness, voltage_kv, calib)
    sigma = o.interaction_sigma(voltage_kv)
    trans = np.exp(1j * sigma * pot)[None]
    origins = np.zeros((positions.shape[0], 2), np.int64)
    shifts = positions / float(det_calib)
    patterns = 0
    for m in range(modes.shape[2]):
        patterns = patterns + weights[m] * o.window_stem(trans, modes[..., m], origins, shifts, prop)
    loss = reduce_loss(patterns, data, mode)
    g_i = loss_grad(patterns, data, mode)
    g_t = 0
    g_modes = np.zeros_like(modes)
    g_shift = 0
    for m in range(modes.shape[2]):
        gt, gp, gs = o.multislice_vjp(trans, modes[..., m], origins, shifts, prop, weights[m] * g_i,
                                      want_shifts=True)
        g_t = g_t + gt
        g_modes[..., m] = gp
        g_shift = g_shift + gs
    # T = exp(i sigma V): dL/dV = Re(conj(g_T) i sigma T)
    g_pot = np.real(np.conj(g_t[0]) * 1j * sigma * trans[0])
    return loss, g_pot, g_modes, g_shift / float(det_calib)


def reconstruct(data, positions, det_calib, voltage_kv, pot0, modes0, weights, beam_calib, slice_thickness,
                save_every, num_iterations, learning_rate, pos_learning_rate, loss_type, optimizer_name,
                update_positions: bool):
    """The shared Optax loop; returns (pot, modes, positions, snap_pot, snap_modes, snap_pos).

    pot0 is the complex potential parameter (the forward uses its real
    part); modes0 (H, W, M) complex; snapshots stack along a trailing axis.
    """
    mode = LossType(loss_type)
    param_opt = OptaxRule(optimizer_name, learning_rate)
    pos_opt = OptaxRule(optimizer_name, pos_learning_rate)
    pot = promote_to_complex(pot0)
    modes = promote_to_complex(modes0)
    pos = np.asarray(positions, np.float64).copy()
    st_pot, st_modes, st_pos = param_opt.init(pot), param_opt.init(modes), pos_opt.init(pos)
    count = num_iterations // save_every
    snap_pot = np.zeros(pot.shape + (count,), pot.dtype)
    snap_modes = np.zeros(modes.shape + (count,), modes.dtype)
    snap_pos = np.zeros(pos.shape + (count,), pos.dtype)

    def step():
        nonlocal pot, modes, pos, st_pot, st_modes, st_pos
        _, g_pot, g_modes, g_pos = stem4d_value_and_grad(pot.real, modes, weights, pos, det_calib, det_calib,
                                                         beam_calib, slice_thickness, voltage_kv, data, mode)
        u, st_pot = param_opt.update(g_pot.astype(pot.dtype), st_pot)
        pot = (pot + u).astype(pot.dtype)
        u, st_modes = param_opt.update(g_modes.astype(modes.dtype), st_modes)
        modes = (modes + u).astype(modes.dtype)
        if update_positions:
            u, st_pos = pos_opt.update(g_pos, st_pos)
            pos = pos + u

    def save(slot):
        if 0 <= slot < count:
            snap_pot[..., slot] = pot
            snap_modes[..., slot] = modes
            snap_pos[..., slot] = pos

    if num_iterations > 0:
        step()
        save(0)
        for ii in range(1, num_iterations):
            step()
            if ii % save_every == 0:
                save(ii // save_every)
    return pot, modes, pos, snap_pot, snap_modes, snap_pos
