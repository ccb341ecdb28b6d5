"""Gradients through the drop-in API (cbed_amplitude, cbed_image, stem_4d) on the GPU.

The reference differentiates these functions with jax.grad (its reconstructors and
tests/test_ptyrodactyl/test_multislice/test_simulations.py:319-430).  Here the potential and
probe modes reach the HIP kernels through torch (exp(i sigma V), fft2) and the HIP adjoint
supplies the rest.  Checked against central differences of the complex128 oracle; tolerances
cover complex64 arithmetic and the FD truncation error.
"""

import numpy as np
import pytest

from oracle import ms_oracle as o

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _phase_case(n=32):
    base = (np.arange(1, n * n + 1, dtype=np.float64).reshape(n, n) / (n * n) + 0.125j).astype(np.complex128)
    return base, np.zeros((n, n, 1))


def test_amplitude_phase_gradient_survives_intensity_seam():
    """Port of test_simulations.py:376-430 at the kernels' smallest grid: the amplitude keeps a global probe
    phase that the intensity removes; both gradients are finite and they differ.  The amplitude gradient
    equals central differences of the oracle."""
    import ptyrodactyl_amd as pa
    base, pot = _phase_case()
    dev = torch.device("cuda", 0)
    pots = pa.create_potential_slices(torch.tensor(pot, device=dev), 1.0, 0.5)
    mic = pa.create_microscope_config(80.0, 1.0)

    def modes_for(phase):
        return pa.ProbeModes((torch.tensor(base, device=dev) * torch.exp(1j * phase))[..., None],
                             torch.ones(1, dtype=torch.float64, device=dev), 0.5)

    phase = torch.tensor(0.37, dtype=torch.float64, device=dev, requires_grad=True)
    amp_loss = torch.real(torch.sum(pa.cbed_amplitude(pots, modes_for(phase), mic)))
    (g_amp,) = torch.autograd.grad(amp_loss, phase)
    phase2 = torch.tensor(0.37, dtype=torch.float64, device=dev, requires_grad=True)
    int_loss = torch.sum(pa.cbed_image(pots, modes_for(phase2), mic).data_array)
    (g_int,) = torch.autograd.grad(int_loss, phase2)
    assert torch.isfinite(g_amp) and torch.isfinite(g_int)
    assert abs(float(g_amp) - float(g_int)) > 1e-8

    def oracle_amp_loss(ph):
        return float(np.real(o.cbed_amplitude(pot, 1.0, 0.5, (base * np.exp(1j * ph))[..., None], 0.5, 80.0).sum()))

    h = 1e-4
    fd = (oracle_amp_loss(0.37 + h) - oracle_amp_loss(0.37 - h)) / (2 * h)
    print(f"amplitude phase gradient: GPU {float(g_amp):.6e}  oracle FD {fd:.6e}", flush=True)
    assert abs(float(g_amp) - fd) <= 1e-4 * abs(fd) + 1e-6
    # the intensity of a probe with a global phase does not depend on it: what remains is float32 cancellation
    # noise on the scale of the loss itself
    assert abs(float(g_int)) < 1e-6 * abs(float(int_loss.detach()))


def test_stem4d_gradient_wrt_potential_matches_oracle_fd():
    import ptyrodactyl_amd as pa
    n, s_count = 32, 3
    rng = np.random.default_rng(5)
    pot = rng.uniform(0, 30, (n, n, s_count))
    dev = torch.device("cuda", 0)
    mic = pa.create_microscope_config(voltage_kv=200.0, aperture_mrad=25.0, defocus_ang=10.0, probe_shape=(n, n))
    det0 = pa.create_detector_config(real_space_calib_ang=0.25)
    probe = pa.make_probe(mic, det0)
    beam = pa.create_probe_modes(probe[..., None].to(dev), torch.tensor([1.0]), 0.25)
    positions = torch.tensor([[1.0, 2.0], [3.5, 0.5], [2.0, 2.5]], dtype=torch.float64)
    det = pa.create_detector_config(real_space_calib_ang=0.25, scan_positions_px=positions)
    wts = rng.standard_normal((3, n, n))
    v = torch.tensor(pot, device=dev, requires_grad=True)
    out = pa.stem_4d(pa.create_potential_slices(v, 1.5, 0.25), beam, mic, det)
    loss = torch.sum(out.data.double() * torch.tensor(wts, device=dev))
    (g_v,) = torch.autograd.grad(loss, v)
    g_v = g_v.cpu().numpy()

    probe_np = probe.numpy()

    def oracle_loss(p):
        pats = o.stem_4d(p, 1.5, 0.25, probe_np[..., None], np.array([1.0]), 0.25, 200.0, positions.numpy(), 0.25)["data"]
        return float(np.sum(pats * wts))

    h = 1e-3
    worst = 0.0
    for _ in range(8):
        y, x, s = rng.integers(0, n), rng.integers(0, n), rng.integers(0, s_count)
        pp, pm = pot.copy(), pot.copy()
        pp[y, x, s] += h
        pm[y, x, s] -= h
        fd = (oracle_loss(pp) - oracle_loss(pm)) / (2 * h)
        worst = max(worst, abs(g_v[y, x, s] - fd) / (abs(fd) + 1e-3 * np.abs(g_v).max()))
    print(f"stem_4d dL/dV vs oracle FD: worst relative deviation {worst:.2e}", flush=True)
    assert worst < 1e-2


def test_cbed_image_gradient_wrt_mode_weights():
    """I = sum_m w_m |A_m|^2, so dL/dw_m = sum g |A_m|^2 (oracle amplitudes)."""
    import ptyrodactyl_amd as pa
    n = 32
    rng = np.random.default_rng(9)
    pot = rng.uniform(0, 20, (n, n, 2))
    dev = torch.device("cuda", 0)
    probes = np.stack([o.make_probe(200.0, 20.0 + 5 * m, 15.0 * m, 0.0, 0.0, (n, n), 25.0) for m in range(2)], axis=2)
    w = torch.tensor([0.7, 0.3], dtype=torch.float64, device=dev, requires_grad=True)
    beam = pa.ProbeModes(torch.tensor(probes, device=dev), w, 0.25)
    mic = pa.create_microscope_config(voltage_kv=200.0, aperture_mrad=20.0, probe_shape=(n, n))
    g = rng.standard_normal((n, n))
    img = pa.cbed_image(pa.create_potential_slices(torch.tensor(pot, device=dev), 1.0, 0.25), beam, mic)
    loss = torch.sum(img.data_array.double() * torch.tensor(g, device=dev))
    (g_w,) = torch.autograd.grad(loss, w)
    amps = o.cbed_amplitude(pot, 1.0, 0.25, probes, 0.25, 200.0)
    ref = np.array([np.sum(g * np.abs(amps[..., m]) ** 2) for m in range(2)])
    print(f"dL/dw GPU {g_w.cpu().numpy()}  oracle {ref}", flush=True)
    np.testing.assert_allclose(g_w.cpu().numpy(), ref, rtol=1e-4)


def test_stem4d_gradient_wrt_scan_positions_matches_oracle_fd():
    """The position-correcting reconstructors differentiate stem_4d w.r.t. the scan positions
    (multislice_recon.py:849-885): positions reach the kernels as shifts of positions / calibration."""
    import ptyrodactyl_amd as pa
    n, s_count = 32, 3
    rng = np.random.default_rng(15)
    pot = rng.uniform(0, 30, (n, n, s_count))
    dev = torch.device("cuda", 0)
    mic = pa.create_microscope_config(voltage_kv=200.0, aperture_mrad=25.0, defocus_ang=10.0, probe_shape=(n, n))
    probe = pa.make_probe(mic, pa.create_detector_config(real_space_calib_ang=0.25))
    beam = pa.create_probe_modes(probe[..., None].to(dev), torch.tensor([1.0]), 0.25)
    pos0 = np.array([[1.0, 2.0], [3.5, 0.5], [2.0, 2.5], [0.3, 5.1]])
    positions = torch.tensor(pos0, dtype=torch.float64, device=dev, requires_grad=True)
    det = pa.create_detector_config(real_space_calib_ang=0.25, scan_positions_px=positions)
    wts = rng.standard_normal((4, n, n))
    out = pa.stem_4d(pa.create_potential_slices(torch.tensor(pot, device=dev), 1.5, 0.25), beam, mic, det)
    loss = torch.sum(out.data.double() * torch.tensor(wts, device=dev))
    (g_pos,) = torch.autograd.grad(loss, positions)
    g_pos = g_pos.cpu().numpy()
    probe_np = probe.numpy()

    def oracle_loss(pos):
        pats = o.stem_4d(pot, 1.5, 0.25, probe_np[..., None], np.array([1.0]), 0.25, 200.0, pos, 0.25)["data"]
        return float(np.sum(pats * wts))

    h = 1e-4
    fd = np.zeros_like(pos0)
    for p in range(pos0.shape[0]):
        for ax in range(2):
            pp, pm = pos0.copy(), pos0.copy()
            pp[p, ax] += h
            pm[p, ax] -= h
            fd[p, ax] = (oracle_loss(pp) - oracle_loss(pm)) / (2 * h)
    err = float(np.linalg.norm(g_pos - fd) / np.linalg.norm(fd))
    print(f"stem_4d dL/dpositions vs oracle FD: rel-L2 {err:.2e}", flush=True)
    assert err < 1e-3


def test_cbed_image_gradient_wrt_probe_aberrations():
    """Defocus and C3 given as tensors reach the probe through make_probe's aberration phase, so cbed_image is
    differentiable w.r.t. them (the reference differentiates the carrier scalars, test_simulations.py:319-374);
    compared with central differences of the oracle."""
    import ptyrodactyl_amd as pa
    n = 32
    rng = np.random.default_rng(21)
    pot = rng.uniform(0, 25, (n, n, 2))
    g = rng.standard_normal((n, n))
    dev = torch.device("cuda", 0)
    defocus = torch.tensor(30.0, dtype=torch.float64, requires_grad=True)
    c3 = torch.tensor(2e-3, dtype=torch.float64, requires_grad=True)
    mic = pa.create_microscope_config(voltage_kv=200.0, aperture_mrad=25.0, defocus_ang=defocus, c3_ang=c3,
                                      probe_shape=(n, n))
    det = pa.create_detector_config(real_space_calib_ang=0.25)
    probe = pa.make_probe(mic, det)
    beam = pa.create_probe_modes(probe[..., None].to(dev), torch.tensor([1.0]), 0.25)
    img = pa.cbed_image(pa.create_potential_slices(torch.tensor(pot, device=dev), 1.5, 0.25), beam, mic)
    g_def, g_c3 = torch.autograd.grad(torch.sum(img.data_array.double() * torch.tensor(g, device=dev)), (defocus, c3))

    def oracle_loss(df, cs):
        pr = o.make_probe(200.0, 25.0, df, cs, 0.0, (n, n), 25.0)
        return float(np.sum(g * o.cbed_image(pot, 1.5, 0.25, pr[..., None], [1.0], 0.25, 200.0)[0]))

    fd_def = (oracle_loss(30.01, 2e-3) - oracle_loss(29.99, 2e-3)) / 0.02
    fd_c3 = (oracle_loss(30.0, 2e-3 + 1e-6) - oracle_loss(30.0, 2e-3 - 1e-6)) / 2e-6
    print(f"d/ddefocus GPU {float(g_def):.6e} FD {fd_def:.6e}; d/dC3 GPU {float(g_c3):.6e} FD {fd_c3:.6e}",
          flush=True)
    assert abs(float(g_def) - fd_def) <= 1e-3 * abs(fd_def) + 1e-9
    assert abs(float(g_c3) - fd_c3) <= 1e-3 * abs(fd_c3) + 1e-9
