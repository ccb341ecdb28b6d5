"""GPU parity of the streaming forward and adjoint at 1024^2, 2048^2 and 4096^2 windows (three-pass row / column FFTs)
against the complex128 oracle, vacuum closed forms at full size, and the size limits of the paths that stop at
512 (batched) / 1024 (fused detectors).

Tolerance: relative L2 error of the patterns <= 1e-5 (complex64 kernels against the complex128 oracle)."""

import numpy as np
import pytest

from oracle import ms_oracle as o

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
TOL = 1e-5


def rel_l2(a, b):
    """Relative L2 error; complex arrays keep their imaginary parts."""
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a.astype(np.result_type(a, np.float64)) - b) / np.linalg.norm(b))


def _case(n, s_count, p_count, slab_hw, seed, modes):
    rng = np.random.default_rng(seed)
    slab = (rng.uniform(0.95, 1.0, (s_count,) + slab_hw)
            * np.exp(1j * rng.uniform(0, 0.4, (s_count,) + slab_hw))).astype(np.complex64)
    probes = np.stack([o.make_probe(300.0, 20.0 + 3 * m, 30.0 * m, 0.0, 0.0, (n, n), 10.0) for m in range(modes)])
    origins = np.stack([rng.integers(0, slab_hw[0], p_count), rng.integers(0, slab_hw[1], p_count)], 1)
    return slab, probes, origins.astype(np.int32), rng.uniform(-3.0, 3.0, (p_count, 2))


def _hip(slab, probes, origins, shifts, n, **kw):
    from ptyrodactyl_amd import engine, physics
    dev = torch.device("cuda", 0)
    m = probes.shape[0]
    pk = torch.fft.fft2(torch.tensor(probes, device=dev)).to(torch.complex64).contiguous()
    pat, _ = engine.multislice_forward(torch.tensor(slab, device=dev), pk,
                                      torch.full((m,), 1.0 / m, dtype=torch.float32, device=dev),
                                      torch.tensor(origins, device=dev), torch.tensor(shifts, device=dev),
                                      physics.propagator_factors(n, 1.5, 300.0, 0.1, device=dev), **kw)
    torch.cuda.synchronize()
    return pat.cpu().numpy()


@pytest.mark.parametrize("n, s_count, p_count, slab_hw, modes", [(1024, 4, 3, (1100, 1060), 2),
                                                                 (2048, 3, 2, (2048, 2100), 1),
                                                                 (4096, 2, 1, (4096, 4096), 1)])
def test_large_windows_match_oracle(n, s_count, p_count, slab_hw, modes):
    print(f"n={n} S={s_count} P={p_count} M={modes} slab={slab_hw}", flush=True)
    slab, probes, origins, shifts = _case(n, s_count, p_count, slab_hw, n, modes)
    got = _hip(slab, probes, origins, shifts, n)
    prop = o.propagation_func(n, n, 1.5, 300.0, 0.1)
    want = sum(o.window_stem(slab.astype(np.complex128), probes[m], origins, shifts, prop) for m in range(modes))
    want = want / modes
    err = rel_l2(got, want)
    print(f"n={n}: rel-L2 {err:.2e}", flush=True)
    assert err < TOL


@pytest.mark.parametrize("n", [2048, 4096])
def test_large_window_vacuum_is_the_aperture_indicator(n):
    """V = 0: every pattern is the fftshifted aperture indicator, independent of slices and shifts."""
    probe = o.make_probe(300.0, 20.0, 40.0, 0.0, 0.0, (n, n), 10.0)
    slab = np.ones((3, n, n), np.complex64)
    got = _hip(slab, probe[None], np.zeros((2, 2), np.int32), np.asarray([[0.0, 0.0], [1.3, -2.7]]), n)
    ind = np.fft.fftshift((np.abs(np.fft.fft2(probe)) > 0.5).astype(np.float64))
    for p in range(2):
        assert np.abs(got[p] - ind).max() < 1e-4
        assert abs(got[p].sum() - ind.sum()) < 1e-4 * ind.sum()


@pytest.mark.parametrize("n, s_count, k, p_count, slab_hw", [(1024, 3, 2, 2, (1024, 1060)), (2048, 2, 1, 1, (2048, 2048)),
                                                              (4096, 2, 2, 1, (4096, 4096))])
def test_large_window_adjoint_matches_oracle_vjp(n, s_count, k, p_count, slab_hw):
    """Checkpointed forward + gathered adjoint at 1024^2 - 4096^2 against the oracle's VJP (slab and probe)."""
    from test_adjoint_gpu import _case as adj_case, _gpu_grads, _oracle_grads
    print(f"adjoint n={n} S={s_count} k={k} P={p_count}", flush=True)
    slab, probes, origins, shifts, g = adj_case(n, s_count, p_count, slab_hw, seed=n + s_count)
    _, g_slab, g_probe, w = _gpu_grads(slab, probes, origins, shifts, g, n, k)
    ref_slab, ref_probe = _oracle_grads(slab, probes, origins, shifts, g, w, n)
    e_slab, e_probe = rel_l2(g_slab, ref_slab), rel_l2(g_probe, ref_probe)
    print(f"  n={n}: rel-L2 slab {e_slab:.2e} probe {e_probe:.2e}", flush=True)
    assert e_slab < TOL and e_probe < TOL


def test_paths_limited_to_smaller_windows_raise():
    from ptyrodactyl_amd import _native, engine, physics
    dev = torch.device("cuda", 0)
    w = torch.ones(1, dtype=torch.float32, device=dev)
    org = torch.zeros((1, 2), dtype=torch.int32, device=dev)
    sh = torch.zeros((1, 2), dtype=torch.float64, device=dev)
    lib = _native.load()
    assert lib.pty_forward_batched(1024, 1, 1, None, None, None, None, None, None, None, 1, None, 0,
                                   None) == _native.PTY_ERR_INVALID
    assert b"batched paths" in lib.pty_last_error()
    big = torch.ones((1, 2048, 2048), dtype=torch.complex64, device=dev)
    with pytest.raises(ValueError, match="virtual detectors"):
        engine.multislice_forward_detectors(big, big, w, org, sh, physics.propagator_factors(2048, 1.0, 300.0, 0.1,
                                                                                             device=dev),
                                            torch.ones((1, 2048, 2048), dtype=torch.uint8, device=dev))


def test_drop_in_stem_4d_and_its_potential_gradient_at_1024():
    """The drop-in stem_4d on a 1024^2 grid (positions in A, as the reference applies them) against the oracle's
    stem_4d, and its potential gradient (HIP adjoint through autograd) against central differences of the oracle
    along a random direction."""
    import ptyrodactyl_amd as pa
    n, calib, dz, kv = 1024, 0.05, 1.5, 200.0
    rng = np.random.default_rng(9)
    pot = rng.uniform(0.0, 5.0, (n, n, 2))
    positions = np.asarray([[10.0, 12.0], [25.3, 17.9]])
    pot_t = torch.tensor(pot, requires_grad=True)
    slices = pa.create_potential_slices(pot_t, dz, calib)
    mic = pa.create_microscope_config(voltage_kv=kv, aperture_mrad=20.0, defocus_ang=30.0, probe_shape=(n, n))
    det = pa.create_detector_config(real_space_calib_ang=calib, probe_calibration_pm=calib * 100.0,
                                    scan_positions_px=torch.tensor(positions))
    probe = pa.make_probe(mic, det)
    pm = pa.create_probe_modes(probe[..., None], torch.tensor([1.0]), calib)
    s4d = pa.stem_4d(slices, pm, mic, det)
    modes = probe.detach().cpu().numpy()[..., None]

    def oracle(v):
        return o.stem_4d(v, dz, calib, modes, [1.0], calib, kv, positions, calib)["data"]

    want = oracle(pot)
    err = rel_l2(s4d.data.detach().cpu().numpy(), want)
    print(f"drop-in stem_4d n=1024: rel-L2 {err:.2e}", flush=True)
    assert err < TOL
    g = rng.standard_normal(want.shape)
    (s4d.data * torch.tensor(g, device=s4d.data.device, dtype=s4d.data.dtype)).sum().backward()
    d = rng.standard_normal(pot.shape)
    h = 1e-3
    fd = (np.sum(g * oracle(pot + h * d)) - np.sum(g * oracle(pot - h * d))) / (2 * h)
    got = float(np.sum(pot_t.grad.cpu().numpy() * d))
    print(f"  directional derivative: HIP adjoint {got:.6e} oracle FD {fd:.6e}", flush=True)
    assert abs(got - fd) < 1e-4 * abs(fd)
