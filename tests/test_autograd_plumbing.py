"""CPU checks of the differentiable multislice Functions' plumbing (no GPU).

The HIP entry points behind ``ptyrodactyl_amd.autograd`` are replaced by a complex128 torch restatement
of the windowed multislice and its torch-autograd derivatives.  The Functions must then reproduce, through
``torch.autograd``, ``torch.func.vjp`` and ``torch.func.jvp``, the derivatives torch computes on the
restatement directly: this checks which tangent and cotangent goes to which kernel argument, the complex
conventions, the shift and weight derivatives and the checkpoint hand-off, independently of the kernels
(whose values the GPU tests compare with the oracle).  The Functions hand float32 weights and dL/dI to the
kernels, as the C-ABI takes them, so agreement is to float32 rounding (relative 1e-6); a misrouted tangent
or a conjugation error is an O(1) difference.
"""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from ptyrodactyl_amd import autograd, engine  # noqa: E402


def _ramp(n, shift):
    k = torch.fft.fftfreq(n, dtype=torch.float64) * n
    return torch.exp(-2j * math.pi * k * shift / n)


def _twin(slab, pk, weights, origins, shifts, prop, amplitudes=False):
    """Windowed multislice in torch: patterns (P, n, n) or amplitudes (P, M, n, n), fftshifted."""
    prop_y, prop_x = prop
    n = pk.shape[-1]
    hs, ws = slab.shape[-2:]
    p_out = []
    for p in range(origins.shape[0]):
        oy, ox = int(origins[p, 0]), int(origins[p, 1])
        rows = (oy + torch.arange(n)) % hs
        cols = (ox + torch.arange(n)) % ws
        ramp = _ramp(n, shifts[p, 0])[:, None] * _ramp(n, shifts[p, 1])[None, :]
        m_out = []
        for m in range(pk.shape[0]):
            wave = torch.fft.ifft2(pk[m] * ramp)
            for s in range(slab.shape[0]):
                wave = wave * slab[s][rows][:, cols]
                if s != slab.shape[0] - 1:
                    # prop_y / prop_x carry 1/n each, as the kernels' factors do
                    wave = torch.fft.ifft2(torch.fft.fft2(wave) * (prop_y[:, None] * prop_x[None, :]) * n * n)
            m_out.append(torch.fft.fftshift(torch.fft.fft2(wave)))
        amp = torch.stack(m_out)
        p_out.append(amp if amplitudes else (weights[:, None, None] * (amp.real ** 2 + amp.imag ** 2)).sum(0))
    return torch.stack(p_out)


def _plain(*tensors):
    """The kernels read raw device memory: every operand must be a plain tensor with storage, never one that a
    torch.func transform wraps (reading its data pointer raises)."""
    for t in tensors:
        if isinstance(t, torch.Tensor):
            t.data_ptr()
        elif isinstance(t, (tuple, list)):
            _plain(*t)


class _MockEngine:
    JVP_N = engine.JVP_N

    @staticmethod
    def multislice_forward(slab, pk, w, origins, shifts, prop, want_patterns=True, want_amplitude=False, **kw):
        _plain(slab, pk, w, origins, shifts, prop)
        with torch.no_grad():
            pat = _twin(slab, pk, w, origins, shifts, prop) if want_patterns else None
            amp = _twin(slab, pk, w, origins, shifts, prop, True) if want_amplitude else None
        return pat, amp

    @staticmethod
    def multislice_forward_checkpointed(slab, pk, w, origins, shifts, prop, k, **kw):
        pat, _ = _MockEngine.multislice_forward(slab, pk, w, origins, shifts, prop)
        return pat, torch.zeros((1,), dtype=torch.complex64)

    @staticmethod
    def multislice_amplitude_checkpointed(slab, pk, w, origins, shifts, prop, k, **kw):
        _, amp = _MockEngine.multislice_forward(slab, pk, w, origins, shifts, prop, want_patterns=False,
                                                want_amplitude=True)
        return amp, torch.zeros((1,), dtype=torch.complex64)

    @staticmethod
    def multislice_adjoint(slab, pk, w, origins, shifts, prop, grad_patterns, ckpt, k, want_probe=True,
                           grad_amplitude=None, grad_shifts=None, **kw):
        _plain(slab, pk, w, origins, shifts, prop, grad_patterns, ckpt, grad_amplitude, grad_shifts)
        assert ckpt is not None and ckpt.numel() == 1  # the forward's checkpoints are handed back
        with torch.enable_grad():
            xs = slab.detach().clone().requires_grad_(True)
            xp = pk.detach().clone().requires_grad_(True)
            xh = shifts.detach().clone().requires_grad_(True)
            if grad_amplitude is not None:
                out = _twin(xs, xp, w, origins, xh, prop, True)
                seed = grad_amplitude
            else:
                out = _twin(xs, xp, w, origins, xh, prop)
                seed = grad_patterns
            g_s, g_p, g_h = torch.autograd.grad(out, (xs, xp, xh), seed.to(out.dtype))
        if grad_shifts is not None:
            grad_shifts += g_h
        return g_s, (g_p if want_probe else None)

    @staticmethod
    def multislice_jvp(slab, pk, w, origins, shifts, prop, d_slab=None, d_probe_k=None, d_shifts=None,
                       want_patterns=False, want_amplitude=False, want_d_patterns=True, want_d_amplitude=False,
                       **kw):
        _plain(slab, pk, w, origins, shifts, prop, d_slab, d_probe_k, d_shifts)
        zeros = torch.zeros_like
        tangents = (zeros(slab) if d_slab is None else d_slab, zeros(pk) if d_probe_k is None else d_probe_k,
                    zeros(shifts) if d_shifts is None else d_shifts)
        _, d_pat = torch.func.jvp(lambda a, b, c: _twin(a, b, w, origins, c, prop), (slab, pk, shifts), tangents)
        _, d_amp = torch.func.jvp(lambda a, b, c: _twin(a, b, w, origins, c, prop, True), (slab, pk, shifts),
                                  tangents)
        amp = _twin(slab, pk, w, origins, shifts, prop, True) if want_amplitude else None
        return None, amp, (d_pat if want_d_patterns else None), (d_amp if want_d_amplitude else None)


@pytest.fixture
def problem(monkeypatch):
    for name in ("multislice_forward", "multislice_forward_checkpointed", "multislice_amplitude_checkpointed",
                 "multislice_adjoint", "multislice_jvp"):
        monkeypatch.setattr(autograd.engine, name, getattr(_MockEngine, name))
    monkeypatch.setattr(autograd, "_jvp_supported", lambda probe_k: True)  # the 8 x 8 grid of the restatement
    rng = np.random.default_rng(5)
    n, s_count, p_count, modes = 8, 3, 3, 2
    slab = torch.tensor(np.exp(1j * rng.uniform(0, 1, (s_count, 11, 10))) * rng.uniform(0.9, 1, (s_count, 11, 10)))
    pk = torch.tensor(rng.standard_normal((modes, n, n)) + 1j * rng.standard_normal((modes, n, n)))
    weights = torch.tensor([0.75, 0.25], dtype=torch.float64)  # exact in float32, as the kernels take them
    origins = torch.tensor(rng.integers(0, 10, (p_count, 2)), dtype=torch.int32)
    shifts = torch.tensor(rng.uniform(-2, 2, (p_count, 2)))
    q = torch.fft.fftfreq(n, dtype=torch.float64)
    prop = (torch.exp(-1j * 0.4 * q ** 2) / n, torch.exp(-1j * 0.7 * q ** 2) / n)
    c = torch.tensor(rng.standard_normal((p_count, n, n)))
    tangents = tuple(torch.tensor(rng.standard_normal(t.shape) + (1j * rng.standard_normal(t.shape)
                                                                  if t.is_complex() else 0))
                     for t in (slab, pk, weights, shifts))
    return slab, pk, weights, origins, shifts, prop, c, tangents


def test_reverse_mode_matches_torch_on_the_restatement(problem):
    slab, pk, weights, origins, shifts, prop, c, _ = problem
    leaves = [t.clone().requires_grad_(True) for t in (slab, pk, weights, shifts)]
    loss = (c * autograd.multislice_patterns(leaves[0], leaves[1], leaves[2], origins, leaves[3], prop)).sum()
    got = torch.autograd.grad(loss, leaves)
    ref_leaves = [t.clone().requires_grad_(True) for t in (slab, pk, weights, shifts)]
    ref_loss = (c * _twin(ref_leaves[0], ref_leaves[1], ref_leaves[2], origins, ref_leaves[3], prop)).sum()
    want = torch.autograd.grad(ref_loss, ref_leaves)
    for g, r in zip(got, want):
        assert torch.allclose(g.to(r.dtype), r, rtol=1e-6, atol=1e-6 * float(r.abs().max()))


def test_func_vjp_and_jvp_match_torch_on_the_restatement(problem):
    slab, pk, weights, origins, shifts, prop, c, tangents = problem

    def f(a, b, w, h):
        return autograd.multislice_patterns(a, b, w, origins, h, prop)

    def f_ref(a, b, w, h):
        return _twin(a, b, w, origins, h, prop)

    out, vjp_fn = torch.func.vjp(f, slab, pk, weights, shifts)
    ref_out, ref_vjp_fn = torch.func.vjp(f_ref, slab, pk, weights, shifts)
    assert torch.allclose(out.to(ref_out.dtype), ref_out, rtol=1e-6)
    for g, r in zip(vjp_fn(c), ref_vjp_fn(c)):
        assert torch.allclose(g.to(r.dtype), r, rtol=1e-6, atol=1e-6 * float(r.abs().max()))
    _, d_out = torch.func.jvp(f, (slab, pk, weights, shifts), tangents)
    _, d_ref = torch.func.jvp(f_ref, (slab, pk, weights, shifts), tangents)
    assert torch.allclose(d_out.to(d_ref.dtype), d_ref, rtol=1e-6, atol=1e-6 * float(d_ref.abs().max()))


def test_amplitude_function_vjp_and_jvp(problem):
    slab, pk, _, origins, shifts, prop, _, tangents = problem
    ones = torch.ones(pk.shape[0], dtype=torch.float64)

    def f(a, b, h):
        return autograd.multislice_amplitudes(a, b, origins, h, prop)

    def f_ref(a, b, h):
        return _twin(a, b, ones, origins, h, prop, True)

    _, d_out = torch.func.jvp(f, (slab, pk, shifts), (tangents[0], tangents[1], tangents[3]))
    _, d_ref = torch.func.jvp(f_ref, (slab, pk, shifts), (tangents[0], tangents[1], tangents[3]))
    assert torch.allclose(d_out, d_ref, rtol=1e-6, atol=1e-6 * float(d_ref.abs().max()))
    rng = np.random.default_rng(2)
    u = torch.tensor(rng.standard_normal(d_ref.shape) + 1j * rng.standard_normal(d_ref.shape))
    _, vjp_fn = torch.func.vjp(f, slab, pk, shifts)
    _, ref_vjp_fn = torch.func.vjp(f_ref, slab, pk, shifts)
    for g, r in zip(vjp_fn(u), ref_vjp_fn(u)):
        assert torch.allclose(g.to(r.dtype), r, rtol=1e-6, atol=1e-6 * float(r.abs().max()))


def test_differentiable_detects_transformed_operands():
    x = torch.ones(3)
    assert not autograd.differentiable(x)
    assert autograd.differentiable(x.clone().requires_grad_(True))
    seen = []
    torch.func.jvp(lambda t: seen.append(autograd.differentiable(t)) or t * 2, (x,), (x,))
    torch.func.vjp(lambda t: seen.append(autograd.differentiable(t)) or t * 2, x)
    assert seen == [True, True]
    with torch.no_grad():
        assert not autograd.differentiable(x.clone().requires_grad_(True))


def test_gauss_newton_and_lm_run_through_the_multislice_functions(problem):
    """The reference's matrix-free GN / LM (jacobian_operators through torch.func) take a forward built on the
    differentiable multislice: from a perturbed slab they reduce the residual to the data of the true slab."""
    from ptyrodactyl_amd import jacobian
    slab, pk, weights, origins, shifts, prop, _, _ = problem
    data = _twin(slab, pk, weights, origins, shifts, prop)
    rng = np.random.default_rng(9)
    start = slab * torch.tensor(1.0 + 0.02 * rng.standard_normal(slab.shape))

    def forward_fn(params):
        return autograd.multislice_patterns(params["slab"], pk, weights, origins, shifts, prop)

    r0 = float(torch.linalg.norm(forward_fn({"slab": start}) - data))
    p_gn, st_gn = jacobian.gauss_newton_solve(forward_fn, {"slab": start}, data, max_iterations=3,
                                              cg_max_iterations=20)
    p_lm, st_lm = jacobian.levenberg_marquardt_solve(forward_fn, {"slab": start}, data, max_iterations=3,
                                                     damping_init=1e-3, cg_max_iterations=20)
    assert st_gn.iteration == 3 and st_lm.iteration == 3
    assert float(st_gn.residual_norm) < 0.1 * r0 and float(st_lm.residual_norm) < 0.1 * r0
    assert float(torch.linalg.norm(forward_fn(p_lm) - data)) == pytest.approx(float(st_lm.residual_norm), rel=1e-6)
