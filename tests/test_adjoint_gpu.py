"""GPU parity of the HIP adjoint (gradient) kernels.

Gradients are compared with the complex128 oracle's reverse mode
(``multislice_vjp``, itself pinned by finite differences and an independent
torch autograd twin) at relative L2 <= 1e-5, and with central finite
differences of the oracle loss (h = 1e-3).  With the full workspace the
slab and probe gradients are gathered in a fixed order and must be bitwise
reproducible; the atomic fallback (small workspace) is compared with a
tolerance.
"""

import numpy as np
import pytest

from oracle import ms_oracle as o

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _case(n, s_count, p_count, slab_hw, seed, modes=1):
    rng = np.random.default_rng(seed)
    slab = (rng.uniform(0.95, 1.0, (s_count,) + slab_hw) * np.exp(1j * rng.uniform(0, 0.4, (s_count,) + slab_hw)))
    slab = slab.astype(np.complex64)
    probes = np.stack([o.make_probe(300.0, 20.0 + 4 * m, 25.0 * (m + 1), 0.0, 0.0, (n, n), 10.0)
                       for m in range(modes)])
    origins = np.stack([rng.integers(0, slab_hw[0], p_count), rng.integers(0, slab_hw[1], p_count)], 1)
    shifts = rng.uniform(-2.5, 2.5, (p_count, 2))
    g = rng.standard_normal((p_count, n, n))
    return slab, probes, origins.astype(np.int32), shifts, g


def _gpu_grads(slab, probes, origins, shifts, g, n, k, weights=None, dz=1.5, kv=300.0, calib=0.1):
    from ptyrodactyl_amd import engine, physics
    dev = torch.device("cuda", 0)
    m_count = probes.shape[0]
    w = np.full(m_count, 1.0 / m_count) if weights is None else np.asarray(weights) / np.sum(weights)
    slab_t = torch.tensor(slab, device=dev)
    pk = torch.fft.fft2(torch.tensor(probes, device=dev)).to(torch.complex64).contiguous()
    wt = torch.tensor(w, dtype=torch.float32, device=dev)
    ot = torch.tensor(origins, dtype=torch.int32, device=dev)
    st = torch.tensor(shifts, device=dev)
    prop = physics.propagator_factors(n, dz, kv, calib, device=dev)
    pat, ckpt = engine.multislice_forward_checkpointed(slab_t, pk, wt, ot, st, prop, k)
    gt = torch.tensor(g, dtype=torch.float32, device=dev)
    g_slab, g_pk = engine.multislice_adjoint(slab_t, pk, wt, ot, st, prop, gt, ckpt, k)
    # probe_k = fft2(probe)  =>  dL/dprobe = F^H dL/dprobe_k = n^2 ifft2(.)
    g_probe = torch.fft.ifft2(g_pk.to(torch.complex128)) * (n * n)
    torch.cuda.synchronize()
    return pat.cpu().numpy(), g_slab.cpu().numpy(), g_probe.cpu().numpy(), w


def _oracle_grads(slab, probes, origins, shifts, g, w, n, dz=1.5, kv=300.0, calib=0.1):
    prop = o.propagation_func(n, n, dz, kv, calib)
    gs = 0
    gps = []
    for m in range(probes.shape[0]):
        a, b = o.multislice_vjp(slab.astype(np.complex128), probes[m], origins, shifts, prop, w[m] * g)
        gs = gs + a
        gps.append(b)
    return gs, np.stack(gps)


@pytest.mark.parametrize("n,s_count,k,p_count,slab_hw,modes", [
    (32, 1, 1, 3, (40, 40), 1),
    (64, 5, 2, 4, (70, 64), 1),
    (64, 6, 4, 3, (64, 96), 2),
    (128, 7, 3, 3, (140, 128), 1),
    (128, 4, 4, 2, (128, 128), 1),
    (256, 5, 2, 3, (256, 280), 1),
    (256, 3, 8, 2, (300, 256), 2),
    (512, 2, 1, 2, (512, 512), 1),
    (256, 6, 1, 3, (280, 256), 2),
    (256, 1, 1, 2, (256, 270), 1),
    (256, 7, 3, 2, (260, 300), 2),
])
def test_adjoint_matches_oracle_vjp(n, s_count, k, p_count, slab_hw, modes):
    print(f"adjoint n={n} S={s_count} k={k} P={p_count} M={modes}", flush=True)
    slab, probes, origins, shifts, g = _case(n, s_count, p_count, slab_hw, seed=n * 7 + s_count, modes=modes)
    _, g_slab, g_probe, w = _gpu_grads(slab, probes, origins, shifts, g, n, k)
    ref_slab, ref_probe = _oracle_grads(slab, probes, origins, shifts, g, w, n)
    e_slab, e_probe = rel_l2(g_slab, ref_slab), rel_l2(g_probe, ref_probe)
    print(f"  rel-L2 slab {e_slab:.2e} probe {e_probe:.2e}", flush=True)
    assert e_slab < TOL and e_probe < TOL


def test_checkpoint_interval_does_not_change_gradient():
    n = 128
    slab, probes, origins, shifts, g = _case(n, 8, 3, (128, 160), seed=21)
    _, a, pa, _ = _gpu_grads(slab, probes, origins, shifts, g, n, 1)
    _, b, pb, _ = _gpu_grads(slab, probes, origins, shifts, g, n, 8)
    assert rel_l2(a, b) < 1e-6 and rel_l2(pa, pb) < 1e-6


@pytest.mark.parametrize("n", [256, 512])
def test_checkpoint_store_at_every_slice(n):
    """Checkpoints are side writes of the streaming passes: the patterns of a forward with a checkpoint at every
    slice or every 4th equal the plain forward bitwise, and slot 3 (the wave entering slice 4) equals slot 0 of
    the forward that checkpoints every 4th slice."""
    from ptyrodactyl_amd import engine
    print(f"checkpoint store n={n}", flush=True)
    s_count = 9
    slab, probes, origins, shifts, _ = _case(n, s_count, 5, (n + 40, n + 24), seed=n + 3, modes=1)
    dev = torch.device("cuda", 0)
    sl = torch.tensor(slab, device=dev)
    pk = torch.fft.fft2(torch.tensor(probes, device=dev)).to(torch.complex64).contiguous()
    w = torch.ones(1, dtype=torch.float32, device=dev)
    org = torch.tensor(origins, dtype=torch.int32, device=dev)
    sh = torch.tensor(shifts, device=dev)
    from ptyrodactyl_amd import physics
    prop = physics.propagator_factors(n, 1.2, 200.0, 0.2, device=dev)
    plain, _ = engine.multislice_forward(sl, pk, w, org, sh, prop)
    p1, ck1 = engine.multislice_forward_checkpointed(sl, pk, w, org, sh, prop, 1)
    p4, ck4 = engine.multislice_forward_checkpointed(sl, pk, w, org, sh, prop, 4)
    assert ck1.shape[1] == s_count - 1 and ck4.shape[1] == 2
    assert torch.equal(p1, plain) and torch.equal(p4, plain)
    assert torch.equal(ck1[:, 3], ck4[:, 0]) and torch.equal(ck1[:, 7], ck4[:, 1])


def test_autograd_function_matches_oracle():
    from ptyrodactyl_amd import autograd, physics
    n = 64
    slab, probes, origins, shifts, c = _case(n, 4, 5, (80, 64), seed=33)
    dev = torch.device("cuda", 0)
    slab_t = torch.tensor(slab, device=dev, requires_grad=True)
    probe = torch.tensor(probes[0], device=dev, requires_grad=True)
    pk = torch.fft.fft2(probe).to(torch.complex64)[None].contiguous()
    pat = autograd.multislice_patterns(slab_t, pk, torch.ones(1, device=dev), torch.tensor(origins, device=dev),
                                       torch.tensor(shifts, device=dev),
                                       physics.propagator_factors(n, 1.5, 300.0, 0.1, device=dev), ckpt_every=2)
    loss = (pat * torch.tensor(c, dtype=torch.float32, device=dev)).sum()
    loss.backward()
    ref_slab, ref_probe = _oracle_grads(slab, probes, origins, shifts, c, [1.0], n)
    assert rel_l2(slab_t.grad.cpu().numpy(), ref_slab) < TOL
    assert rel_l2(probe.grad.cpu().numpy(), ref_probe[0]) < TOL


def test_gradient_matches_central_differences():
    """GPU gradient against central FD (h = 1e-3) of the complex128 oracle loss, real and imaginary."""
    n = 64
    slab, probes, origins, shifts, c = _case(n, 3, 4, (64, 64), seed=44)
    _, g_slab, _, _ = _gpu_grads(slab, probes, origins, shifts, c, n, 2)
    prop = o.propagation_func(n, n, 1.5, 300.0, 0.1)
    base = slab.astype(np.complex128)

    def loss(sl):
        return float(np.sum(c * o.window_stem(sl, probes[0], origins, shifts, prop)))

    rng = np.random.default_rng(45)
    h = 1e-3
    for _ in range(8):
        idx = (rng.integers(0, 3), rng.integers(0, 64), rng.integers(0, 64))
        for direction in (1.0, 1j):
            sp, sm = base.copy(), base.copy()
            sp[idx] += h * direction
            sm[idx] -= h * direction
            fd = (loss(sp) - loss(sm)) / (2 * h)
            got = g_slab[idx].real if direction == 1.0 else g_slab[idx].imag
            assert got == pytest.approx(fd, rel=1e-3, abs=1e-4 * np.abs(g_slab).max())


def test_config3_full_gradient():
    """Gradient workload at full size: 16 x 512^2 object, 128^2 probe, 32 x 32 jittered scan,
    sqrt-intensity loss against Poisson data, checkpoints every 4 slices.  GPU gradients w.r.t. all
    slices and the probe against the oracle's reverse mode end to end (the loss evaluated on the GPU's
    complex64 patterns against the oracle's complex128 ones), at relative L2 <= 1e-5, and against
    central differences (h = 1e-3) of the complex128 oracle loss on 64 seeded slab pixels, real and
    imaginary.  eps = 1e-6 keeps the loss derivative finite where I = 0; it is about 1e-4 of the intensity
    one electron contributes at the 1e5-electron dose (each pattern sums to ~1200)."""
    from ptyrodactyl_amd import autograd, physics, synth
    w = synth.config3()
    n = w.n
    prop = o.propagation_func(n, n, w.slice_thickness, w.voltage_kv, w.calib)
    clean = o.window_stem_batched(w.object_true, w.probe, w.origins, w.shifts_px, prop, workers=16)
    measured = synth.poisson_measure(clean)
    eps = 1e-6
    dev = torch.device("cuda", 0)
    obj = torch.tensor(w.object_t, device=dev, requires_grad=True)
    probe = torch.tensor(w.probe, device=dev, requires_grad=True)
    pk = torch.fft.fft2(probe).to(torch.complex64)[None].contiguous()
    print("config3 forward+adjoint on GPU", flush=True)
    pat = autograd.multislice_patterns(obj, pk, torch.ones(1, device=dev), torch.tensor(w.origins, device=dev),
                                       torch.tensor(w.shifts_px, device=dev),
                                       physics.propagator_factors(n, w.slice_thickness, w.voltage_kv, w.calib,
                                                                  device=dev), ckpt_every=w.ckpt_every)
    meas_t = torch.tensor(measured, dtype=torch.float32, device=dev)
    loss = ((torch.sqrt(pat + eps) - torch.sqrt(meas_t)) ** 2).sum()
    loss.backward()
    g_obj, g_probe = obj.grad.cpu().numpy(), probe.grad.cpu().numpy()
    # oracle reverse mode with the same upstream gradient dL/dI = (sqrt(I+eps) - sqrt(M)) / sqrt(I+eps)
    inten = o.window_stem_batched(w.object_t.astype(np.complex128), w.probe, w.origins, w.shifts_px, prop,
                                  workers=16)
    root = np.sqrt(inten + eps)
    gi = (root - np.sqrt(measured)) / root
    print("config3 oracle reverse mode", flush=True)
    r_obj, r_probe = o.multislice_vjp(w.object_t.astype(np.complex128), w.probe, w.origins, w.shifts_px, prop, gi)
    e_pat = rel_l2(pat.detach().double().cpu().numpy(), inten)
    e_obj, e_probe = rel_l2(g_obj, r_obj), rel_l2(g_probe, r_probe)
    print(f"config3 patterns rel-L2 {e_pat:.2e}; end-to-end gradient rel-L2: slices {e_obj:.2e}  probe {e_probe:.2e}",
          flush=True)
    # adjoint alone: the oracle's reverse mode fed the GPU run's own dL/dI
    pat64 = pat.detach().double().cpu().numpy()
    root_g = np.sqrt(pat64 + eps)
    gi_gpu = (root_g - np.sqrt(meas_t.double().cpu().numpy())) / root_g
    a_obj, a_probe = o.multislice_vjp(w.object_t.astype(np.complex128), w.probe, w.origins, w.shifts_px, prop, gi_gpu)
    f_obj, f_probe = rel_l2(g_obj, a_obj), rel_l2(g_probe, a_probe)
    print(f"config3 adjoint rel-L2 (same dL/dI): slices {f_obj:.2e}  probe {f_probe:.2e}", flush=True)
    assert f_obj < TOL and f_probe < TOL
    assert e_obj < TOL and e_probe < TOL

    # central differences of the complex128 oracle loss on 64 seeded slab pixels (real and imaginary)
    rng = np.random.default_rng(64)
    h = 1e-3
    base = w.object_t.astype(np.complex128)
    worst = 0.0
    scale = np.abs(g_obj).max()
    for i in range(64):
        s, y, x = rng.integers(0, w.num_slices), rng.integers(64, 448), rng.integers(64, 448)
        covers = np.nonzero(((y - w.origins[:, 0]) % 512 < n) & ((x - w.origins[:, 1]) % 512 < n))[0]
        pats = o.window_stem_perturbed(base, w.probe, w.origins[covers], w.shifts_px[covers], prop, (s, y, x),
                                       (h, -h, 1j * h, -1j * h), workers=16)
        losses = np.sum((np.sqrt(pats + eps) - np.sqrt(measured[covers])[None]) ** 2, axis=(1, 2, 3))
        fd_re = (losses[0] - losses[1]) / (2 * h)
        fd_im = (losses[2] - losses[3]) / (2 * h)
        for got, fd in ((g_obj[s, y, x].real, fd_re), (g_obj[s, y, x].imag, fd_im)):
            worst = max(worst, abs(got - fd) / (abs(fd) + 1e-3 * scale))
    print(f"config3 FD check: worst relative deviation {worst:.2e} over 128 directional derivatives "
          f"(64 pixels)", flush=True)
    assert worst < 1e-2


def _adjoint_through_abi(slab, probes, origins, shifts, g, n, k, workspace_bytes=None):
    """pty_forward_checkpointed + pty_adjoint through the C-ABI; workspace_bytes=None takes the full size."""
    from ptyrodactyl_amd import _native, engine, physics
    dev = torch.device("cuda", 0)
    slab_t = torch.tensor(slab, device=dev)
    pk = torch.fft.fft2(torch.tensor(probes, device=dev)).to(torch.complex64).contiguous()
    m_count = probes.shape[0]
    wt = torch.full((m_count,), 1.0 / m_count, dtype=torch.float32, device=dev)
    ot = torch.tensor(origins, dtype=torch.int32, device=dev)
    st = torch.tensor(shifts, device=dev)
    prop = physics.propagator_factors(n, 1.5, 300.0, 0.1, device=dev)
    _, ckpt = engine.multislice_forward_checkpointed(slab_t, pk, wt, ot, st, prop, k)
    gt = torch.tensor(g, dtype=torch.float32, device=dev)
    lib = _native.load()
    s_count, p_count = slab.shape[0], origins.shape[0]
    full = int(lib.pty_adjoint_workspace_bytes(n, s_count, k, p_count, m_count))
    nbytes = full if workspace_bytes is None else int(workspace_bytes)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    g_slab = torch.zeros_like(slab_t)
    g_pk = torch.zeros_like(pk)
    nc = (s_count - 1) // k
    status = lib.pty_adjoint(n, s_count, p_count, m_count, _native.ptr(slab_t), slab.shape[1], slab.shape[2],
                             _native.ptr(pk), _native.ptr(wt), _native.ptr(ot), _native.ptr(st), _native.ptr(prop),
                             _native.ptr(prop), _native.ptr(gt), _native.ptr(ckpt) if nc else 0, k,
                             _native.ptr(g_slab), _native.ptr(g_pk), 0, _native.ptr(ws), nbytes,
                             _native.stream_handle(dev))
    _native.check(status, "pty_adjoint")
    torch.cuda.synchronize()
    return g_slab, g_pk, full


@pytest.mark.parametrize("n,s_count,p_count,slab_hw,modes", [(128, 6, 37, (300, 280), 1), (64, 5, 300, (200, 260), 2),
                                                          (256, 5, 21, (300, 300), 1)])
def test_gathered_gradients_are_bitwise_reproducible(n, s_count, p_count, slab_hw, modes):
    print(f"gathered adjoint n={n} S={s_count} P={p_count} M={modes}", flush=True)
    slab, probes, origins, shifts, g = _case(n, s_count, p_count, slab_hw, seed=n + p_count, modes=modes)
    a_slab, a_pk, _ = _adjoint_through_abi(slab, probes, origins, shifts, g, n, 2)
    b_slab, b_pk, _ = _adjoint_through_abi(slab, probes, origins, shifts, g, n, 2)
    assert torch.equal(a_slab, b_slab) and torch.equal(a_pk, b_pk)


@pytest.mark.parametrize("n", [128, 256])
def test_atomic_fallback_matches_gathered(n):
    """A workspace below the gathered size runs the atomic path; both sum the same contributions."""
    from ptyrodactyl_amd import _native
    slab, probes, origins, shifts, g = _case(n, 4, 19, (300, 290), seed=3 * n)
    full_slab, full_pk, full = _adjoint_through_abi(slab, probes, origins, shifts, g, n, 2)
    k = 2
    small = (k + 3) * n * n * 8  # scratch for one wave in flight, no room for the contributions
    assert small < full
    at_slab, at_pk, _ = _adjoint_through_abi(slab, probes, origins, shifts, g, n, k, workspace_bytes=small)
    assert rel_l2(at_slab.cpu().numpy(), full_slab.cpu().numpy()) < 1e-6
    assert rel_l2(at_pk.cpu().numpy(), full_pk.cpu().numpy()) < 1e-6
    assert _native.load().pty_abi_version() == _native.ABI_VERSION
