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
    n = int(data.shape[-1])
    p_count = int(origins.shape[0])
    basis = aberration_basis(n, calib_ang, voltage_kv, device=dev)
    aperture = aperture_mask(n, calib_ang, voltage_kv, aperture_mrad, device=dev)
    unit = 1.0 / torch.clamp(torch.amax(torch.abs(basis) * aperture, dim=(1, 2)), min=1e-300)
    sbasis = basis * unit[:, None, None]                   # d chi / d theta_j
    q_count = sbasis.shape[0]
    theta = torch.as_tensor(coeffs_init, dtype=torch.float64, device=dev) / unit
    prop = physics.propagator_factors(n, slice_thickness, voltage_kv, calib_ang, device=dev)
    ones = torch.ones(q_count + 1, dtype=torch.float32, device=dev)
    one = torch.ones(1, dtype=torch.float32, device=dev)
    data32 = data.to(torch.float32).contiguous()
    lib = _native.load()
    ws_bytes = int(lib.pty_normal_equations_workspace_bytes(p_count, n, q_count))
    ne_ws = torch.empty(max(ws_bytes, 8) // 8, dtype=torch.float64, device=dev)
    jtj = torch.empty((q_count, q_count), dtype=torch.float64, device=dev)
    jtr = torch.empty(q_count, dtype=torch.float64, device=dev)
    rss = torch.empty(1, dtype=torch.float64, device=dev)
    step = torch.empty(q_count, dtype=torch.float64, device=dev)
    model = torch.empty(1, dtype=torch.float64, device=dev)
    status = torch.empty(1, dtype=torch.int32, device=dev)
    stream = _native.stream_handle(dev)

    def chi_of(th):
        return torch.einsum("j,jyx->yx", th, sbasis)

    def rss_of(th) -> torch.Tensor:
        pk = probe_k_from_chi(aperture, chi_of(th)).to(torch.complex64)[None].contiguous()
        pat, _ = engine.multislice_forward(slab_t, pk, one, origins, shifts_px, prop)
        return torch.sum((pat.to(torch.float64) - data32.to(torch.float64)) ** 2)

    damping = float(damping_init)
    res_norm = float(torch.sqrt(rss_of(theta)))
    history = [res_norm]
    it = 0
    for _ in range(iterations):
        if res_norm < tolerance:
            break
        pk0 = probe_k_from_chi(aperture, chi_of(theta))
        modes = torch.cat([pk0[None], (-1j) * sbasis * pk0[None]], dim=0).to(torch.complex64).contiguous()
        _, amp = engine.multislice_forward(slab_t, modes, ones, origins, shifts_px, prop, want_patterns=False,
                                           want_amplitude=True)
        _native.check(lib.pty_normal_equations(_native.ptr(amp), _native.ptr(data32), p_count, n, q_count,
                                               _native.ptr(ne_ws), ws_bytes, _native.ptr(jtj), _native.ptr(jtr),
                                               _native.ptr(rss), stream), "pty_normal_equations")
        _native.check(lib.pty_damped_solve(_native.ptr(jtj), _native.ptr(jtr), damping, q_count, _native.ptr(step),
                                           _native.ptr(model), _native.ptr(status), stream), "pty_damped_solve")
        candidate = theta - step
        rss_new = rss_of(candidate)
        gain = float((rss[0] - rss_new) / (model[0] + 1e-12))
        damping = _lm_damping_update(gain, damping)
        if gain > 0.0:
            theta = candidate
            res_norm = float(torch.sqrt(rss_new))
        else:
            res_norm = float(torch.sqrt(rss[0]))
        it += 1
        history.append(res_norm)
    return AberrationFit(coefficients=(theta * unit).detach(), residual_norm=res_norm, damping=damping,
                         iterations=it, history=history)
