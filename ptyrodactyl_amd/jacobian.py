"""Gauss-Newton / Levenberg-Marquardt solvers.

Two layers:

* Generic matrix-free solvers with the reference's signatures and update
  rules (``src/ptyrodactyl/jacobian/solvers.py:62-666``,
  ``operators.py:29-185``, ``_treemath.py:62-99``): ``conjugate_gradient``,
  ``gauss_newton_step/solve``, ``levenberg_marquardt_step/solve``, with
  parameters given as a tensor, a list/tuple or a dict of tensors.  The
  inner product is Re sum conj(a) b; the reference's conj(jax vjp(r)) is
  exactly torch's vjp(r) (torch returns dL/dRe + i dL/dIm).  JVPs and
  VJPs come from ``torch.func`` (on the GPU path they run through the HIP
  forward and adjoint kernels).

* :func:`fit_probe_aberrations`: the mixed-precision LM of the aberration
  reconstruction.  The 12 Jacobian columns are obtained as complex64
  directional derivatives from ONE multislice forward with 1 + 12 probe
  "modes" (the multislice is linear in the probe), the damped normal
  equations are formed in FP64 by ``pty_normal_equations`` and solved in
  FP64 by ``pty_damped_solve``, all on the GPU.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Optional

import torch

from . import _native, engine, physics

# ---------------------------------------------------------------------------
# Tree helpers (``_treemath.py:62-186``)
# ---------------------------------------------------------------------------


def _map(fn, *trees):
    t0 = trees[0]
    if isinstance(t0, dict):
        return {k: _map(fn, *[t[k] for t in trees]) for k in t0}
    if isinstance(t0, (list, tuple)):
        return type(t0)(_map(fn, *items) for items in zip(*trees))
    return fn(*trees)


def _leaves(tree):
    if isinstance(tree, dict):
        return [leaf for k in tree for leaf in _leaves(tree[k])]
    if isinstance(tree, (list, tuple)):
        return [leaf for item in tree for leaf in _leaves(item)]
    return [tree]


def tree_dot(a, b) -> torch.Tensor:
    """Real Hermitian inner product Re sum conj(a) b over all leaves, accumulated in float64 / complex128 on the
    leaves' device (one host read per call)."""
    total = None
    for x, y in zip(_leaves(a), _leaves(b)):
        if x.is_complex() or y.is_complex():
            v = torch.sum(torch.conj(x.to(torch.complex128)) * y.to(torch.complex128)).real
        else:
            v = torch.sum(x.to(torch.float64) * y.to(torch.float64))
        total = v if total is None else total + v.to(total.device)
    return total.cpu() if total is not None else torch.zeros((), dtype=torch.float64)


def tree_add(a, b):
    return _map(lambda x, y: x + y, a, b)


def tree_sub(a, b):
    return _map(lambda x, y: x - y, a, b)


def tree_scale(s, a):
    return _map(lambda x: s * x, a)


def tree_conj(a):
    return _map(lambda x: torch.conj(x) if x.is_complex() else x, a)


def tree_zeros_like(a):
    return _map(torch.zeros_like, a)


# ---------------------------------------------------------------------------
# States (``types/jacobian_types.py:404-485``)
# ---------------------------------------------------------------------------


@dataclass
class CGState:
    x: object
    r: object
    p: object
    r_dot_r: torch.Tensor
    iteration: int


@dataclass
class GNState:
    params: object
    residual_norm: torch.Tensor
    iteration: int


@dataclass
class LMState:
    params: object
    residual_norm: torch.Tensor
    damping: torch.Tensor
    iteration: int


@dataclass
class LanczosState:
    v_prev: torch.Tensor
    v_curr: torch.Tensor
    alpha: torch.Tensor
    beta: torch.Tensor
    iteration: int


# ---------------------------------------------------------------------------
# Operators (``operators.py:29-185``)
# ---------------------------------------------------------------------------


def jvp_operator(forward_fn: Callable, params) -> Callable:
    """v -> J v at ``params``."""
    def apply(v):
        return torch.func.jvp(forward_fn, (params,), (v,))[1]
    return apply


def vjp_operator(forward_fn: Callable, params) -> Callable:
    """u -> J^T u at ``params`` (torch convention; conjugate for JAX's)."""
    _, fn = torch.func.vjp(forward_fn, params)

    def apply(u):
        return fn(u)[0]
    return apply


def jtj_operator(forward_fn: Callable, params) -> Callable:
    """v -> J^T J v."""
    jv = jvp_operator(forward_fn, params)
    vj = vjp_operator(forward_fn, params)

    def apply(v):
        return vj(jv(v))
    return apply


# ---------------------------------------------------------------------------
# Conjugate gradient, Gauss-Newton, Levenberg-Marquardt (``solvers.py``)
# ---------------------------------------------------------------------------


def conjugate_gradient(linear_operator: Callable, rhs, x0, max_iterations: int = 100, tolerance: float = 1e-6):
    """Solve A x = b for symmetric PSD A (``solvers.py:62-180``).

    Same recurrence and freezing rule as the reference: alpha = rr/(pAp + 1e-12),
    beta = rr_new/(rr + 1e-12), and the state freezes once rr < tol^2.
    Returns (solution, iterations).
    """
    r = tree_sub(rhs, linear_operator(x0))
    state = CGState(x=x0, r=r, p=r, r_dot_r=tree_dot(r, r), iteration=0)
    tol2 = tolerance**2
    for _ in range(max_iterations):
        if float(state.r_dot_r) < tol2:
            break
        ap = linear_operator(state.p)
        p_ap = tree_dot(state.p, ap)
        alpha = float(state.r_dot_r / (p_ap + 1e-12))
        x_new = tree_add(state.x, tree_scale(alpha, state.p))
        r_new = tree_sub(state.r, tree_scale(alpha, ap))
        rr_new = tree_dot(r_new, r_new)
        beta = float(rr_new / (state.r_dot_r + 1e-12))
        p_new = tree_add(r_new, tree_scale(beta, state.p))
        state = CGState(x=x_new, r=r_new, p=p_new, r_dot_r=rr_new, iteration=state.iteration + 1)
    return state.x, state.iteration


def lanczos_tridiagonal(linear_operator: Callable, initial_vector: torch.Tensor, num_iterations: int):
    """k-step Lanczos tridiagonalisation of a symmetric operator on real vectors (``solvers.py:669-777``).

    Same three-term recurrence as the reference, without reorthogonalisation:
    w = A v_i, alpha_i = <w, v_i>, w -= alpha_i v_i + beta_{i-1} v_{i-1}, beta_i = |w|,
    v_{i+1} = w / (beta_i + 1e-12), from v_0 = initial_vector / |initial_vector|.  Returns
    (alpha (k,), beta (k-1,)): the diagonal and off-diagonal of the tridiagonal matrix, whose eigenvalues
    approximate the extreme eigenvalues of the operator (used on J^T J for conditioning and nullspace
    estimates).
    """
    k = int(num_iterations)
    v0 = initial_vector / torch.linalg.norm(initial_vector)
    state = LanczosState(torch.zeros_like(v0), v0, torch.zeros(k, dtype=v0.dtype, device=v0.device),
                         torch.zeros(k, dtype=v0.dtype, device=v0.device), 0)
    for _ in range(k):
        w = linear_operator(state.v_curr)
        alpha_i = torch.dot(w, state.v_curr)
        beta_prev = state.beta[state.iteration - 1] if state.iteration > 0 else torch.zeros((), dtype=v0.dtype,
                                                                                             device=v0.device)
        w_orth = w - alpha_i * state.v_curr - beta_prev * state.v_prev
        beta_i = torch.linalg.norm(w_orth)
        alpha = state.alpha.clone()
        beta = state.beta.clone()
        alpha[state.iteration] = alpha_i
        beta[state.iteration] = beta_i
        state = LanczosState(state.v_curr, w_orth / (beta_i + 1e-12), alpha, beta, state.iteration + 1)
    return state.alpha, state.beta[:-1]


def _residual_sum_squares(forward_fn, params, data) -> torch.Tensor:
    r = forward_fn(params) - data
    return torch.sum(r.to(torch.float64) ** 2).cpu()


def gauss_newton_step(forward_fn: Callable, params, data, cg_max_iterations: int = 50, cg_tolerance: float = 1e-6):
    """One GN step: delta = CG(J^T J, J^T r), theta <- theta - delta (``solvers.py:183-264``)."""
    residual = forward_fn(params) - data
    # torch's VJP already returns what the reference obtains as conj(jax vjp)
    gradient = vjp_operator(forward_fn, params)(residual)
    jtj = jtj_operator(forward_fn, params)
    step, _ = conjugate_gradient(jtj, gradient, tree_zeros_like(params), cg_max_iterations, cg_tolerance)
    new_params = tree_sub(params, step)
    return new_params, torch.sqrt(_residual_sum_squares(forward_fn, new_params, data))


def gauss_newton_solve(forward_fn: Callable, params_init, data, max_iterations: int = 20, tolerance: float = 1e-8,
                       cg_max_iterations: int = 50, cg_tolerance: float = 1e-6):
    """Iterated GN with the reference's freezing rule (``solvers.py:267-384``)."""
    state = GNState(params_init, torch.sqrt(_residual_sum_squares(forward_fn, params_init, data)), 0)
    for _ in range(max_iterations):
        if float(state.residual_norm) < tolerance:
            break
        p, norm = gauss_newton_step(forward_fn, state.params, data, cg_max_iterations, cg_tolerance)
        state = GNState(p, norm, state.iteration + 1)
    return state.params, state


def _lm_damping_update(gain: float, damping: float) -> float:
    """lambda/3 if rho > 0.75, 2 lambda if rho < 0.25 (``solvers.py:491-503``)."""
    if gain > 0.75:
        return damping / 3.0
    if gain < 0.25:
        return damping * 2.0
    return damping


def levenberg_marquardt_step(forward_fn: Callable, params, data, damping, cg_max_iterations: int = 50,
                             cg_tolerance: float = 1e-6):
    """One LM step with the reference's gain ratio and acceptance (``solvers.py:387-521``).

    rho = (|r|^2 - |r_new|^2) / (<s, g> - 1/2 <s, JtJ s> + 1e-12); the
    candidate is accepted when rho > 0.  Returns (params, residual_norm, damping).
    """
    lam = float(damping)
    residual = forward_fn(params) - data
    rss = torch.sum(residual.to(torch.float64) ** 2).cpu()
    gradient = vjp_operator(forward_fn, params)(residual)  # = conj(jax vjp), as the reference uses
    jtj = jtj_operator(forward_fn, params)

    def damped(v):
        return tree_add(jtj(v), tree_scale(lam, v))

    step, _ = conjugate_gradient(damped, gradient, tree_zeros_like(params), cg_max_iterations, cg_tolerance)
    candidate = tree_sub(params, step)
    rss_new = _residual_sum_squares(forward_fn, candidate, data)
    predicted = tree_dot(step, gradient) - 0.5 * tree_dot(step, jtj(step))
    gain = float((rss - rss_new) / (predicted + 1e-12))
    new_damping = _lm_damping_update(gain, lam)
    if gain > 0.0:
        return candidate, torch.sqrt(rss_new), torch.tensor(new_damping, dtype=torch.float64)
    return params, torch.sqrt(rss), torch.tensor(new_damping, dtype=torch.float64)


def levenberg_marquardt_solve(forward_fn: Callable, params_init, data, max_iterations: int = 50,
                              tolerance: float = 1e-8, damping_init: float = 1e-3, cg_max_iterations: int = 50,
                              cg_tolerance: float = 1e-6):
    """Iterated LM with the reference's freezing rule (``solvers.py:524-666``)."""
    state = LMState(params_init, torch.sqrt(_residual_sum_squares(forward_fn, params_init, data)),
                    torch.tensor(float(damping_init), dtype=torch.float64), 0)
    for _ in range(max_iterations):
        if float(state.residual_norm) < tolerance:
            break
        p, norm, lam = levenberg_marquardt_step(forward_fn, state.params, data, state.damping, cg_max_iterations,
                                                cg_tolerance)
        state = LMState(p, norm, lam, state.iteration + 1)
    return state.params, state


# ---------------------------------------------------------------------------
# Twelve-coefficient probe aberrations and the FP64 LM on the GPU
# ---------------------------------------------------------------------------

ABERRATION_NAMES = ("C1", "A1a", "A1b", "B2a", "B2b", "C3", "A2a", "A2b", "S3a", "S3b", "A3a", "A3b")


def aberration_basis(n: int, calib_ang: float, voltage_kv: float, device=None) -> torch.Tensor:
    """d chi / d c_j for the 12 axial coefficients (Angstrom) on the fftfreq grid, float64 (12, n, n).

    chi = (2 pi / lambda) Re[C1 |w|^2/2 + A1 conj(w)^2/2 + B2 w conj(w)^2 + A2 conj(w)^3/3
                              + C3 |w|^4/4 + S3 w conj(w)^3 + A3 conj(w)^4/4],  w = lambda (qx + i qy),
    complex coefficients X = Xa + i Xb contribute Re(X tau) = Xa Re tau - Xb Im tau.  With only C1 and
    C3 set this equals the reference ``aberration`` (``simulations.py:434-498``, C3 there in units of 1e7 A).
    """
    lam = physics.relativistic_wavelength_ang(voltage_kv)
    q = physics.fftfreq(n, calib_ang, device=device)
    qy, qx = torch.meshgrid(q, q, indexing="ij")
    w = lam * torch.complex(qx, qy)
    wc = torch.conj(w)
    ww = (w * wc).real
    scale = 2 * math.pi / lam
    terms = {
        "A1": wc**2 / 2, "B2": w * wc**2, "A2": wc**3 / 3, "S3": w * wc**3, "A3": wc**4 / 4,
    }
    out = [ww / 2]
    for name in ("A1", "B2"):
        out += [terms[name].real, -terms[name].imag]
    out.append(ww**2 / 4)
    for name in ("A2", "S3", "A3"):
        out += [terms[name].real, -terms[name].imag]
    return scale * torch.stack(out)


def aperture_mask(n: int, calib_ang: float, voltage_kv: float, aperture_mrad: float, device=None) -> torch.Tensor:
    lam = physics.relativistic_wavelength_ang(voltage_kv)
    q = physics.fftfreq(n, calib_ang, device=device)
    qy, qx = torch.meshgrid(q, q, indexing="ij")
    return (torch.sqrt(qx**2 + qy**2) <= (aperture_mrad / 1000.0) / lam).to(torch.float64)


def probe_k_from_chi(aperture: torch.Tensor, chi: torch.Tensor) -> torch.Tensor:
    """fft2 of the make_probe probe ifftshift(ifft2(A exp(-i chi))) = A exp(-i chi) (-1)^(ky + kx)."""
    n = aperture.shape[-1]
    k = torch.arange(n, device=aperture.device)
    sign = 1.0 - 2.0 * ((k[:, None] + k[None, :]) % 2).to(torch.float64)
    return torch.polar(aperture * sign, -chi)


@dataclass
class AberrationFit:
    coefficients: torch.Tensor      # (12,) float64, Angstrom
    residual_norm: float
    damping: float
    iterations: int
    history: list


class _AberrationModel:
    """The pieces of the aberration LM shared by the coefficient-only and the joint fit: the scaled basis
    (a unit step of a scaled coefficient moves chi by at most 1 rad inside the aperture), the FP64
    normal-equation and solve buffers, and the coefficient step."""

    def __init__(self, n, p_count, calib_ang, voltage_kv, aperture_mrad, slice_thickness, device):
        self.n, self.p_count, self.dev = int(n), int(p_count), device
        basis = aberration_basis(n, calib_ang, voltage_kv, device=device)
        self.aperture = aperture_mask(n, calib_ang, voltage_kv, aperture_mrad, device=device)
        self.unit = 1.0 / torch.clamp(torch.amax(torch.abs(basis) * self.aperture, dim=(1, 2)), min=1e-300)
        self.sbasis = basis * self.unit[:, None, None]     # d chi / d theta_j
        self.q = self.sbasis.shape[0]
        self.prop = physics.propagator_factors(n, slice_thickness, voltage_kv, calib_ang, device=device)
        self.ones = torch.ones(self.q + 1, dtype=torch.float32, device=device)
        self.one = torch.ones(1, dtype=torch.float32, device=device)
        self.lib = _native.load()
        self.ws_bytes = int(self.lib.pty_normal_equations_workspace_bytes(p_count, n, self.q))
        self.ne_ws = torch.empty(max(self.ws_bytes, 8) // 8, dtype=torch.float64, device=device)
        self.jtj = torch.empty((self.q, self.q), dtype=torch.float64, device=device)
        self.jtr = torch.empty(self.q, dtype=torch.float64, device=device)
        self.rss_buf = torch.empty(1, dtype=torch.float64, device=device)
        self.step = torch.empty(self.q, dtype=torch.float64, device=device)
        self.model = torch.empty(1, dtype=torch.float64, device=device)
        self.status = torch.empty(1, dtype=torch.int32, device=device)
        self.stream = _native.stream_handle(device)

    def probe_k(self, theta) -> torch.Tensor:
        return probe_k_from_chi(self.aperture, torch.einsum("j,jyx->yx", theta, self.sbasis))

    def rss(self, slab_t, theta, origins, shifts_px, data32) -> torch.Tensor:
        pk = self.probe_k(theta).to(torch.complex64)[None].contiguous()
        pat, _ = engine.multislice_forward(slab_t, pk, self.one, origins, shifts_px, self.prop)
        return torch.sum((pat.to(torch.float64) - data32.to(torch.float64)) ** 2)

    def coefficient_step(self, slab_t, theta, damping, origins, shifts_px, data32):
        """One LM step on the coefficients: 13-mode complex64 forward (the probe and its 12 directional
        derivatives) for the Jacobian columns J_j = 2 Re(conj(A) dA_j), FP64 J^T J / J^T r / |r|^2
        (``pty_normal_equations``) and the FP64 damped solve (``pty_damped_solve``) on the GPU, the
        reference's gain ratio, damping rule and acceptance.  Returns (theta, rss, damping)."""
        lib = self.lib
        pk0 = self.probe_k(theta)
        modes = torch.cat([pk0[None], (-1j) * self.sbasis * pk0[None]], dim=0).to(torch.complex64).contiguous()
        _, amp = engine.multislice_forward(slab_t, modes, self.ones, origins, shifts_px, self.prop,
                                           want_patterns=False, want_amplitude=True)
        _native.check(lib.pty_normal_equations(_native.ptr(amp), _native.ptr(data32), self.p_count, self.n, self.q,
                                               _native.ptr(self.ne_ws), self.ws_bytes, _native.ptr(self.jtj),
                                               _native.ptr(self.jtr), _native.ptr(self.rss_buf), self.stream),
                      "pty_normal_equations")
        _native.check(lib.pty_damped_solve(_native.ptr(self.jtj), _native.ptr(self.jtr), float(damping), self.q,
                                           _native.ptr(self.step), _native.ptr(self.model), _native.ptr(self.status),
                                           self.stream), "pty_damped_solve")
        candidate = theta - self.step
        rss_new = self.rss(slab_t, candidate, origins, shifts_px, data32)
        rss = self.rss_buf[0].clone()
        gain = float((rss - rss_new) / (self.model[0] + 1e-12))
        damping = _lm_damping_update(gain, float(damping))
        if gain > 0.0:
            return candidate, rss_new, damping
        return theta, rss, damping


def fit_probe_aberrations(slab_t: torch.Tensor, origins: torch.Tensor, shifts_px: torch.Tensor,
                          data: torch.Tensor, coeffs_init, *, calib_ang: float, voltage_kv: float,
                          aperture_mrad: float, slice_thickness: float, iterations: int = 30,
                          damping_init: float = 1e-3, tolerance: float = 1e-8) -> AberrationFit:
    """Levenberg-Marquardt fit of the 12 probe aberration coefficients to 4D-STEM intensities.

    Per iteration: one complex64 multislice forward with 13 probe modes (the
    probe and its 12 directional derivatives -i dchi/dc_j probe_k) gives the
    Jacobian columns J_j = 2 Re(conj(A) dA_j); ``pty_normal_equations`` forms
    J^T J, J^T r and |r|^2 in FP64 and ``pty_damped_solve`` solves
    (J^T J + lambda I) s = J^T r in FP64 on the GPU.  Parameters are scaled
    so a unit step changes chi by at most 1 rad inside the aperture; the
    gain ratio, damping schedule and acceptance follow
    ``levenberg_marquardt_step`` (``solvers.py:387-521``).
    """
    dev = slab_t.device
    model = _AberrationModel(int(data.shape[-1]), int(origins.shape[0]), calib_ang, voltage_kv, aperture_mrad,
                             slice_thickness, dev)
    theta = torch.as_tensor(coeffs_init, dtype=torch.float64, device=dev) / model.unit
    data32 = data.to(torch.float32).contiguous()
    damping = float(damping_init)
    res_norm = float(torch.sqrt(model.rss(slab_t, theta, origins, shifts_px, data32)))
    history = [res_norm]
    it = 0
    for _ in range(iterations):
        if res_norm < tolerance:
            break
        theta, rss, damping = model.coefficient_step(slab_t, theta, damping, origins, shifts_px, data32)
        res_norm = float(torch.sqrt(rss))
        it += 1
        history.append(res_norm)
    return AberrationFit(coefficients=(theta * model.unit).detach(), residual_norm=res_norm, damping=damping,
                         iterations=it, history=history)


@dataclass
class JointFit:
    coefficients: torch.Tensor      # (12,) float64, Angstrom
    object: torch.Tensor            # (H, W) complex128, the object slice every slice repeats
    residual_norm: float
    damping: tuple                  # (coefficients, object)
    iterations: int
    history: list                   # [(|r| after the coefficient step, |r| after the object step)]


def _cg_fixed(apply_a, rhs, iterations: int):
    """The reference's conjugate gradient (``solvers.py:62-180``) from x0 = 0 for a fixed number of iterations,
    complex128 vectors on the device and FP64 scalars."""
    x = torch.zeros_like(rhs)
    r = rhs.clone()
    p = r.clone()
    rr = tree_dot(r, r)
    for _ in range(int(iterations)):
        ap = apply_a(p)
        alpha = float(rr / (tree_dot(p, ap) + 1e-12))
        x = x + alpha * p
        r = r - alpha * ap
        rr_new = tree_dot(r, r)
        beta = float(rr_new / (rr + 1e-12))
        p = r + beta * p
        rr = rr_new
    return x


def fit_aberrations_and_object(object_init: torch.Tensor, num_slices: int, origins: torch.Tensor,
                               shifts_px: torch.Tensor, data: torch.Tensor, coeffs_init, *, calib_ang: float,
                               voltage_kv: float, aperture_mrad: float, slice_thickness: float,
                               iterations: int = 30, damping_coeffs: float = 1e-3, damping_object: float = 1e-3,
                               cg_iterations: int = 3, ckpt_every: int = 4) -> JointFit:
    """Joint reconstruction of the 12 probe aberrations and an object slice that every one of ``num_slices``
    slices repeats, from 4D-STEM intensities (block-coordinate Levenberg-Marquardt).

    Every iteration takes one LM step on the coefficients with the object fixed (complex64 Jacobian columns,
    FP64 normal equations and damped solve on the GPU, as :func:`fit_probe_aberrations`), then one LM step on
    the object with the coefficients fixed, matrix-free as the reference's ``levenberg_marquardt_step``
    (``src/ptyrodactyl/jacobian/solvers.py:387-521``): g = J^T r from the HIP adjoint, ``cg_iterations`` steps
    of the reference's conjugate gradient on (J^T J + lambda I) s = g with J^T J v = adjoint(JVP(v)) (``pty_jvp``
    then ``pty_adjoint``, the object tangent moving every slice), candidate O - s, the reference's gain ratio,
    damping rule and acceptance.  CG vectors are complex128 and every inner product and residual norm FP64
    on the GPU; the multislice, its JVP and its adjoint run in complex64.
    """
    dev = object_init.device
    n = int(data.shape[-1])
    model = _AberrationModel(n, int(origins.shape[0]), calib_ang, voltage_kv, aperture_mrad, slice_thickness, dev)
    theta = torch.as_tensor(coeffs_init, dtype=torch.float64, device=dev) / model.unit
    obj = object_init.to(torch.complex128).clone()
    data32 = data.to(torch.float32).contiguous()
    lam_c, lam_o = float(damping_coeffs), float(damping_object)
    k = int(ckpt_every)
    s_count = int(num_slices)

    def slab_of(o_):
        return o_.to(torch.complex64)[None].expand(s_count, -1, -1).contiguous()

    slab = slab_of(obj)
    rss = model.rss(slab, theta, origins, shifts_px, data32)
    history = [(float(torch.sqrt(rss)), float(torch.sqrt(rss)))]
    for _ in range(iterations):
        theta, rss, lam_c = model.coefficient_step(slab, theta, lam_c, origins, shifts_px, data32)
        after_c = float(torch.sqrt(rss))
        pk = model.probe_k(theta).to(torch.complex64)[None].contiguous()
        pat, ckpt = engine.multislice_forward_checkpointed(slab, pk, model.one, origins, shifts_px, model.prop, k)
        resid = (pat - data32).contiguous()

        def vjp(u):
            g_slab, _ = engine.multislice_adjoint(slab, pk, model.one, origins, shifts_px, model.prop, u, ckpt, k,
                                                  want_probe=False)
            return g_slab.to(torch.complex128).sum(dim=0)

        def jtj(v):
            d_slab = v.to(torch.complex64)[None].expand(s_count, -1, -1).contiguous()
            _, _, d_pat, _ = engine.multislice_jvp(slab, pk, model.one, origins, shifts_px, model.prop, d_slab)
            return vjp(d_pat)

        g = vjp(resid)
        step = _cg_fixed(lambda v: jtj(v) + lam_o * v, g, cg_iterations)
        cand = obj - step
        cand_slab = slab_of(cand)
        rss_new = model.rss(cand_slab, theta, origins, shifts_px, data32)
        predicted = tree_dot(step, g) - 0.5 * tree_dot(step, jtj(step))
        gain = float((rss.cpu() - rss_new.cpu()) / (predicted + 1e-12))
        lam_o = _lm_damping_update(gain, lam_o)
        if gain > 0.0:
            obj, slab, rss = cand, cand_slab, rss_new
        history.append((after_c, float(torch.sqrt(rss))))
    return JointFit(coefficients=(theta * model.unit).detach(), object=obj.detach(),
                    residual_norm=float(torch.sqrt(rss)), damping=(lam_c, lam_o), iterations=iterations,
                    history=history)
