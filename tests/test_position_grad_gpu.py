"""GPU parity of the scan-position (probe shift) gradient.

The position-correcting reconstructors of the reference differentiate the
4D-STEM loss w.r.t. the scan positions with jax.grad
(src/ptyrodactyl/multislice/multislice_recon.py:849-885, 1915-1951).  Here
the HIP adjoint stores each wave's probe contribution and a fixed-order FP64
reduction turns it into dL/d(shift_y, shift_x) per position.  Compared with
the complex128 oracle (``multislice_vjp(..., want_shifts=True)``, itself
pinned by central differences and a torch autograd twin) at relative L2
1e-5, the bound the slab and probe gradients meet (measured: about 2e-6).
"""

import numpy as np
import pytest

from oracle import ms_oracle as o

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL_SHIFT = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _case(n, s_count, p_count, slab_hw, seed, modes=1):
    rng = np.random.default_rng(seed)
    slab = (rng.uniform(0.9, 1.0, (s_count,) + slab_hw) * np.exp(1j * rng.uniform(0, 0.6, (s_count,) + slab_hw)))
    slab = slab.astype(np.complex64)
    probes = np.stack([o.make_probe(300.0, 20.0 + 4 * m, 25.0 * (m + 1), 0.0, 0.0, (n, n), 10.0)
                       for m in range(modes)])
    origins = np.stack([rng.integers(0, slab_hw[0], p_count), rng.integers(0, slab_hw[1], p_count)], 1)
    shifts = rng.uniform(-2.5, 2.5, (p_count, 2))
    g = rng.standard_normal((p_count, n, n))
    return slab, probes, origins.astype(np.int32), shifts, g


def _gpu_shift_grad(slab, probes, origins, shifts, g, n, k, workspace_bytes=None):
    from ptyrodactyl_amd import _native, engine, physics
    dev = torch.device("cuda", 0)
    m_count = probes.shape[0]
    slab_t = torch.tensor(slab, device=dev)
    pk = torch.fft.fft2(torch.tensor(probes, device=dev)).to(torch.complex64).contiguous()
    wt = torch.full((m_count,), 1.0 / m_count, dtype=torch.float32, device=dev)
    ot = torch.tensor(origins, dtype=torch.int32, device=dev)
    st = torch.tensor(shifts, device=dev)
    prop = physics.propagator_factors(n, 1.5, 300.0, 0.1, device=dev)
    _, ckpt = engine.multislice_forward_checkpointed(slab_t, pk, wt, ot, st, prop, k)
    gt = torch.tensor(g, dtype=torch.float32, device=dev)
    g_shift = torch.zeros((origins.shape[0], 2), dtype=torch.float64, device=dev)
    if workspace_bytes is None:
        engine.multislice_adjoint(slab_t, pk, wt, ot, st, prop, gt, ckpt, k, want_probe=False, grad_shifts=g_shift)
    else:
        lib = _native.load()
        s_count, p_count = slab.shape[0], origins.shape[0]
        ws = torch.empty(int(workspace_bytes), dtype=torch.uint8, device=dev)
        g_slab = torch.zeros_like(slab_t)
        nc = (s_count - 1) // k
        status = lib.pty_adjoint(n, s_count, p_count, m_count, _native.ptr(slab_t), slab.shape[1], slab.shape[2],
                                 _native.ptr(pk), _native.ptr(wt), _native.ptr(ot), _native.ptr(st),
                                 _native.ptr(prop), _native.ptr(prop), _native.ptr(gt),
                                 _native.ptr(ckpt) if nc else 0, k, _native.ptr(g_slab), 0, _native.ptr(g_shift),
                                 _native.ptr(ws), int(workspace_bytes), _native.stream_handle(dev))
        _native.check(status, "pty_adjoint")
    torch.cuda.synchronize()
    return g_shift.cpu().numpy()


def _oracle_shift_grad(slab, probes, origins, shifts, g, n):
    prop = o.propagation_func(n, n, 1.5, 300.0, 0.1)
    m_count = probes.shape[0]
    total = 0
    for m in range(m_count):
        _, _, gs = o.multislice_vjp(slab.astype(np.complex128), probes[m], origins, shifts, prop, g / m_count,
                                    want_shifts=True)
        total = total + gs
    return total


@pytest.mark.parametrize("n,s_count,k,p_count,slab_hw,modes", [
    (32, 1, 1, 5, (40, 40), 1),
    (64, 4, 2, 6, (70, 64), 2),
    (128, 5, 3, 4, (140, 128), 1),
    (256, 4, 2, 3, (256, 280), 2),
    (512, 2, 1, 2, (512, 512), 1),
])
def test_shift_gradient_matches_oracle(n, s_count, k, p_count, slab_hw, modes):
    print(f"shift gradient n={n} S={s_count} k={k} P={p_count} M={modes}", flush=True)
    slab, probes, origins, shifts, g = _case(n, s_count, p_count, slab_hw, seed=5 * n + modes, modes=modes)
    got = _gpu_shift_grad(slab, probes, origins, shifts, g, n, k)
    ref = _oracle_shift_grad(slab, probes, origins, shifts, g, n)
    err = rel_l2(got, ref)
    print(f"  rel-L2 {err:.2e}  |ref| {np.abs(ref).max():.3e}", flush=True)
    assert err < TOL_SHIFT


def test_shift_gradient_is_bitwise_reproducible():
    n = 128
    slab, probes, origins, shifts, g = _case(n, 3, 9, (150, 150), seed=77, modes=2)
    a = _gpu_shift_grad(slab, probes, origins, shifts, g, n, 2)
    b = _gpu_shift_grad(slab, probes, origins, shifts, g, n, 2)
    assert np.array_equal(a, b)


def test_shift_gradient_needs_the_gathered_workspace():
    from ptyrodactyl_amd import _native
    n, k = 64, 2
    slab, probes, origins, shifts, g = _case(n, 3, 4, (64, 64), seed=78)
    full = int(_native.load().pty_adjoint_workspace_bytes(n, 3, k, 4, 1))
    got = _gpu_shift_grad(slab, probes, origins, shifts, g, n, k, workspace_bytes=full)
    np.testing.assert_allclose(got, _oracle_shift_grad(slab, probes, origins, shifts, g, n), rtol=1e-3, atol=1e-6)
    with pytest.raises(ValueError, match="position gradients"):
        _gpu_shift_grad(slab, probes, origins, shifts, g, n, k, workspace_bytes=full - 1)


def test_autograd_shift_gradient_patterns_and_amplitudes():
    """multislice_patterns and multislice_amplitudes pass the shift gradient to a float64 leaf."""
    from ptyrodactyl_amd import autograd, physics
    n = 64
    slab, probes, origins, shifts, c = _case(n, 3, 5, (80, 64), seed=79)
    dev = torch.device("cuda", 0)
    slab_t = torch.tensor(slab, device=dev)
    pk = torch.fft.fft2(torch.tensor(probes, device=dev)).to(torch.complex64).contiguous()
    ot = torch.tensor(origins, device=dev)
    prop = physics.propagator_factors(n, 1.5, 300.0, 0.1, device=dev)
    sh = torch.tensor(shifts, device=dev, requires_grad=True)
    pat = autograd.multislice_patterns(slab_t, pk, torch.ones(1, device=dev), ot, sh, prop, ckpt_every=2)
    (pat * torch.tensor(c, dtype=torch.float32, device=dev)).sum().backward()
    ref = _oracle_shift_grad(slab, probes, origins, shifts, c, n)
    assert sh.grad.dtype == torch.float64
    assert rel_l2(sh.grad.cpu().numpy(), ref) < TOL_SHIFT

    # amplitude seed: L = Re sum conj(C) A  =>  dL/dA = C; oracle by central differences in the shifts
    rng = np.random.default_rng(80)
    cc = rng.standard_normal((5, 1, n, n)) + 1j * rng.standard_normal((5, 1, n, n))
    sh2 = torch.tensor(shifts, device=dev, requires_grad=True)
    amp = autograd.multislice_amplitudes(slab_t, pk, ot, sh2, prop, ckpt_every=2)
    torch.real(torch.sum(torch.conj(torch.tensor(cc, dtype=torch.complex64, device=dev)) * amp)).backward()
    prop_np = o.propagation_func(n, n, 1.5, 300.0, 0.1)

    def loss(p, s):
        _, a = o.window_stem(slab.astype(np.complex128), probes[0], origins[p:p + 1], s[None], prop_np,
                             return_amplitude=True)
        return float(np.real(np.sum(np.conj(cc[p, 0]) * a[0])))

    h = 1e-4
    fd = np.zeros((5, 2))
    for p in range(5):
        for ax in range(2):
            sp, sm = shifts[p].copy(), shifts[p].copy()
            sp[ax] += h
            sm[ax] -= h
            fd[p, ax] = (loss(p, sp) - loss(p, sm)) / (2 * h)
    err = rel_l2(sh2.grad.cpu().numpy(), fd)
    print(f"amplitude shift gradient vs oracle FD: rel-L2 {err:.2e}", flush=True)
    assert err < 1e-3
