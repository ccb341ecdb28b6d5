"""GPU parity of the HIP multislice forward against the complex128 oracle.

Tolerance: relative L2 error of the patterns <= 1e-5 (complex64 kernels
against the complex128 oracle), the bar the forward path is held to.
"""

import os
from pathlib import Path

import numpy as np
import pytest

from oracle import ms_oracle as o

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLDEN = Path(__file__).resolve().parent / "golden" / "plan01_ug_reference.npz"
TOL = 1e-5


def rel_l2(a, b):
    """Relative L2 error; complex arrays keep their imaginary parts."""
    a, b = np.asarray(a), np.asarray(b)
    a = a.astype(np.result_type(a, np.float64))
    b = b.astype(np.result_type(b, np.float64))
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _random_case(n, s_count, p_count, slab_hw, seed, modes=1):
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0, 0.4, (s_count,) + slab_hw)
    amp = rng.uniform(0.95, 1.0, (s_count,) + slab_hw)
    slab = (amp * np.exp(1j * phase)).astype(np.complex64)
    probes = [o.make_probe(300.0, 20.0 + 3 * m, 30.0 * m, 0.0, 0.0, (n, n), 10.0) for m in range(modes)]
    origins = np.stack([rng.integers(0, slab_hw[0], p_count), rng.integers(0, slab_hw[1], p_count)], 1)
    shifts = rng.uniform(-3.0, 3.0, (p_count, 2))
    prop = o.propagation_func(n, n, 1.5, 300.0, 0.1)
    return slab, np.stack(probes), origins.astype(np.int32), shifts, prop


def _oracle_patterns(slab, probes, origins, shifts, prop, weights):
    total = 0
    for m in range(probes.shape[0]):
        total = total + weights[m] * o.window_stem(slab.astype(np.complex128), probes[m], origins, shifts, prop)
    return total


def _run_hip(slab, probes, origins, shifts, n, dz=1.5, kv=300.0, calib=0.1, weights=None, amplitude=False):
    from ptyrodactyl_amd import engine, physics
    dev = torch.device("cuda", 0)
    m_count = probes.shape[0]
    w = np.full(m_count, 1.0 / m_count) if weights is None else np.asarray(weights) / np.sum(weights)
    pk = torch.fft.fft2(torch.tensor(probes, device=dev)).to(torch.complex64).contiguous()
    pat, amp = engine.multislice_forward(
        torch.tensor(slab, device=dev), pk, torch.tensor(w, dtype=torch.float32, device=dev),
        torch.tensor(origins, dtype=torch.int32, device=dev), torch.tensor(shifts, device=dev),
        physics.propagator_factors(n, dz, kv, calib, device=dev), want_amplitude=amplitude)
    torch.cuda.synchronize()
    return pat.cpu().numpy(), (amp.cpu().numpy() if amp is not None else None), w


def test_library_loads_on_gpu():
    from ptyrodactyl_amd import _native
    lib = _native.load()
    assert lib.pty_abi_version() == _native.ABI_VERSION


@pytest.mark.parametrize("n,s_count,p_count,slab_hw", [
    (32, 3, 5, (40, 36)),
    (64, 4, 7, (64, 80)),
    (128, 5, 6, (160, 128)),
    (256, 3, 9, (256, 300)),
    (256, 4, 5, (300, 260)),
    (512, 2, 3, (512, 512)),
    (512, 3, 2, (520, 512)),
])
def test_forward_matches_oracle(n, s_count, p_count, slab_hw):
    slab, probes, origins, shifts, prop = _random_case(n, s_count, p_count, slab_hw, seed=n + s_count)
    print(f"forward parity n={n} S={s_count} P={p_count} slab={slab_hw}", flush=True)
    got, _, w = _run_hip(slab, probes, origins, shifts, n)
    ref = _oracle_patterns(slab, probes, origins, shifts, prop, w)
    err = rel_l2(got, ref)
    assert err < TOL, f"rel-L2 {err:.3e} at n={n}"


@pytest.mark.parametrize("n", [64, 256])
def test_forward_single_slice_and_modes(n):
    slab, probes, origins, shifts, prop = _random_case(n, 1, 4, (n, n + 16), seed=11, modes=2)
    print(f"modes/single-slice n={n}", flush=True)
    got, amp, w = _run_hip(slab, probes, origins, shifts, n, weights=[0.7, 0.3], amplitude=True)
    ref = _oracle_patterns(slab, probes, origins, shifts, prop, w)
    assert rel_l2(got, ref) < TOL
    _, ref_amp = o.window_stem(slab.astype(np.complex128), probes[1], origins, shifts, prop, return_amplitude=True)
    assert rel_l2(np.abs(amp[:, 1]), np.abs(ref_amp)) < TOL
    assert rel_l2(amp[:, 1].real, ref_amp.real) < 10 * TOL


def test_forward_vacuum_is_aperture_indicator():
    n = 128
    slab = np.ones((4, n, n), np.complex64)
    probe = o.make_probe(200.0, 20.0, 25.0, 0.0, 0.0, (n, n), 15.0)[None]
    got, _, _ = _run_hip(slab, probe, np.zeros((3, 2), np.int32), np.array([[0, 0], [3.3, -7.25], [60.0, 1.5]]), n,
                         kv=200.0, calib=0.15)
    q = o.probe_fourier_grid(n, n, 15.0)
    ind = np.fft.fftshift((q <= 0.02 / o.wavelength_ang(200.0)).astype(float))
    assert np.max(np.abs(got - ind[None])) < 1e-5


def test_forward_is_deterministic():
    n = 256
    slab, probes, origins, shifts, _ = _random_case(n, 3, 5, (300, 256), seed=3)
    a, _, _ = _run_hip(slab, probes, origins, shifts, n)
    b, _, _ = _run_hip(slab, probes, origins, shifts, n)
    assert np.array_equal(a, b)


def test_streaming_batches_agree_with_one_batch():
    """A batch cap that splits the positions into several launches changes nothing."""
    from ptyrodactyl_amd import engine, physics
    n = 256
    slab, probes, origins, shifts, _ = _random_case(n, 2, 11, (256, 256), seed=5)
    dev = torch.device("cuda", 0)
    args = (torch.tensor(slab, device=dev),
            torch.fft.fft2(torch.tensor(probes, device=dev)).to(torch.complex64).contiguous(),
            torch.ones(1, dtype=torch.float32, device=dev), torch.tensor(origins, device=dev),
            torch.tensor(shifts, device=dev), physics.propagator_factors(n, 1.5, 300.0, 0.1, device=dev))
    full, _ = engine.multislice_forward(*args)
    split, _ = engine.multislice_forward(*args, max_batch_waves=3)
    assert torch.equal(full, split)


def test_golden_fixture_through_hip_drop_in():
    """The reference's golden capture reproduced through the drop-in API on the GPU."""
    import ptyrodactyl_amd as pa
    ref = np.load(GOLDEN)
    pot = torch.tensor(o.jax_uniform_f64(42, (32, 32, 2)) * 10.0)
    slices = pa.create_potential_slices(pot, 2.0, 0.2)
    mic = pa.create_microscope_config(voltage_kv=100.0, aperture_mrad=20.0, probe_shape=(32, 32))
    det = pa.create_detector_config(real_space_calib_ang=0.2, probe_calibration_pm=20.0)
    probe = pa.make_probe(mic, det)
    pm = pa.create_probe_modes(probe[..., None], torch.tensor([1.0]), 0.2)
    cbed = pa.cbed_image(slices, pm, mic)
    assert rel_l2(cbed.data_array.cpu().numpy(), ref["cbed"]) < TOL
    assert cbed.calib_y == pytest.approx(1 / 6.4)
    positions = torch.tensor([[12.0, 12.0], [12.0, 20.0], [20.0, 12.0], [20.0, 20.0]], dtype=torch.float64)
    det2 = pa.create_detector_config(real_space_calib_ang=0.2, probe_calibration_pm=20.0, scan_positions_px=positions)
    s4d = pa.stem_4d(slices, pm, mic, det2)
    assert rel_l2(s4d.data.cpu().numpy(), ref["stem4d"]) < TOL
    np.testing.assert_allclose(s4d.scan_positions.cpu().numpy(), positions.numpy() * 0.2)


def test_cbed_amplitude_matches_oracle():
    import ptyrodactyl_amd as pa
    rng = np.random.default_rng(12)
    n = 64
    pot = rng.uniform(0, 20, (n, n, 3))
    mic = pa.create_microscope_config(voltage_kv=200.0, aperture_mrad=22.0, defocus_ang=20.0, probe_shape=(n, n))
    det = pa.create_detector_config(real_space_calib_ang=0.2)
    probe = pa.make_probe(mic, det)
    pm = pa.create_probe_modes(probe[..., None], [1.0], 0.2)
    amp = pa.cbed_amplitude(pa.create_potential_slices(torch.tensor(pot), 1.0, 0.2), pm, mic)
    ref = o.cbed_amplitude(pot, 1.0, 0.2, probe.numpy()[..., None], 0.2, 200.0)
    assert amp.shape == (n, n, 1)
    assert rel_l2(amp.cpu().numpy().real, ref.real) < 10 * TOL
    assert rel_l2(np.abs(amp.cpu().numpy()), np.abs(ref)) < TOL


def test_config2_subset_full_size():
    """Headline workload (256^2 windows, 64 slices, 1024^2 slab): the 64-position subset (every 64th)
    against the complex128 oracle."""
    from ptyrodactyl_amd import synth
    w = synth.config2(scan=64)
    sel = np.arange(0, 4096, 64)
    probe = w.probe[None]
    prop = o.propagation_func(w.n, w.n, w.slice_thickness, w.voltage_kv, w.calib)
    print("config2 subset", flush=True)
    got, _, _ = _run_hip(w.slab_t, probe, w.origins[sel], w.shifts_px[sel], w.n, w.slice_thickness, w.voltage_kv,
                         w.calib)
    ref = o.window_stem_batched(w.slab_t.astype(np.complex128), w.probe, w.origins[sel], w.shifts_px[sel], prop,
                                workers=16)
    err = rel_l2(got, ref)
    print(f"config2 subset (64 positions) rel-L2 {err:.3e}", flush=True)
    assert err < TOL


@pytest.mark.skipif(os.environ.get("PTY_FULL_CONFIG2") != "1",
                    reason="about 2.5 min of 16-thread CPU oracle work; set PTY_FULL_CONFIG2=1 to run it")
def test_config2_all_positions():
    """Headline workload at full size: all 4096 positions of the 256^2 x 64-slice scan over the 1024^2 slab
    against the complex128 oracle, in chunks of 256 positions; rel-L2 over the whole stack and the worst
    pattern."""
    from ptyrodactyl_amd import synth
    w = synth.config2()
    prop = o.propagation_func(w.n, w.n, w.slice_thickness, w.voltage_kv, w.calib)
    print("config2 all positions", flush=True)
    got, _, _ = _run_hip(w.slab_t, w.probe[None], w.origins, w.shifts_px, w.n, w.slice_thickness, w.voltage_kv,
                         w.calib)
    slab128 = w.slab_t.astype(np.complex128)
    num = den = 0.0
    worst = 0.0
    for c0 in range(0, len(w.origins), 256):
        sel = slice(c0, c0 + 256)
        ref = o.window_stem_batched(slab128, w.probe, w.origins[sel], w.shifts_px[sel], prop, workers=16)
        d = got[sel].astype(np.float64) - ref
        num += float(np.sum(d * d))
        den += float(np.sum(ref * ref))
        per = np.sqrt(np.sum(d * d, axis=(1, 2)) / np.sum(ref * ref, axis=(1, 2)))
        worst = max(worst, float(per.max()))
        print(f"positions {c0}-{c0 + 255}: running rel-L2 {np.sqrt(num / den):.3e}, worst pattern {worst:.3e}",
              flush=True)
    err = float(np.sqrt(num / den))
    print(f"config2 all 4096 positions rel-L2 {err:.3e}, worst pattern {worst:.3e}", flush=True)
    assert err < TOL and worst < TOL


def test_config1_full_workload():
    """Config 1 at its full size: the 5 x 5 x 10-cell perovskite crystal on a 128^2 grid at 0.15 A, 20 slices
    of 1.95 A, 200 kV / 20 mrad, all 256 positions of the 16 x 16 scan, against the complex128 oracle.  The
    same patterns through the drop-in stem_4d (potential slices, make_probe, positions passed in Angstrom as
    the reference's shift_beam_fourier takes them)."""
    import ptyrodactyl_amd as pa
    from ptyrodactyl_amd import synth
    w = synth.config1()
    prop = o.propagation_func(w.n, w.n, w.slice_thickness, w.voltage_kv, w.calib)
    print("config1 full workload (256 positions)", flush=True)
    got, _, _ = _run_hip(w.slab_t, w.probe[None], w.origins, w.shifts_px, w.n, w.slice_thickness, w.voltage_kv,
                         w.calib)
    ref = o.window_stem_batched(w.slab_t.astype(np.complex128), w.probe, w.origins, w.shifts_px, prop, workers=16)
    err = rel_l2(got, ref)
    print(f"config1 (256 positions) rel-L2 {err:.3e}", flush=True)
    assert got.shape == (256, 128, 128) and err < TOL
    # the drop-in API on the potential: positions (pixels x calibration) are taken as Angstrom by the shift
    pot = synth.projected_potential(synth.crystal_atoms(5, 5, 10, seed=0), w.n, w.calib, w.num_slices,
                                    w.slice_thickness)
    mic = pa.create_microscope_config(voltage_kv=w.voltage_kv, aperture_mrad=20.0, probe_shape=(w.n, w.n))
    det = pa.create_detector_config(real_space_calib_ang=w.calib, probe_calibration_pm=100.0 * w.calib,
                                    scan_positions_px=torch.tensor(w.shifts_px * w.calib))
    beam = pa.create_probe_modes(pa.make_probe(mic, det)[..., None], torch.tensor([1.0]), w.calib)
    pots = pa.create_potential_slices(torch.tensor(np.moveaxis(pot, 0, -1)), w.slice_thickness, w.calib)
    out = pa.stem_4d(pots, beam, mic, det).data.cpu().numpy()
    err_api = rel_l2(out, ref)
    print(f"config1 through the drop-in stem_4d: rel-L2 {err_api:.3e}", flush=True)
    assert out.shape == (256, 128, 128) and err_api < TOL


def test_transmission_kernel_matches_oracle():
    """pty_transmission (exp(i sigma V) in FP64, stored complex64) against the oracle's transmission_func:
    each component within float32 rounding of the FP64 value, over potentials from 0 to 1e5 V A (phases up
    to ~73 rad at 200 kV) and a negative range."""
    from ptyrodactyl_amd import _native, physics
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(17)
    pot = np.concatenate([rng.uniform(0.0, 500.0, 40000), rng.uniform(0.0, 1e5, 20000),
                          rng.uniform(-300.0, 0.0, 5000), [0.0, 1e-300, 2.5e4]])
    for kv in (100.0, 200.0, 300.0):
        sigma = physics.interaction_sigma(kv)
        assert sigma == pytest.approx(o.interaction_sigma(kv), rel=1e-14)
        pot_t = torch.tensor(pot, dtype=torch.float64, device=dev)
        out = torch.empty(pot.size, dtype=torch.complex64, device=dev)
        lib = _native.load()
        _native.check(lib.pty_transmission(_native.ptr(pot_t), sigma, _native.ptr(out), pot.size,
                                           _native.stream_handle(dev)), "pty_transmission")
        torch.cuda.synchronize()
        got = out.cpu().numpy().astype(np.complex128)
        want = o.transmission_func(pot, kv)
        err = np.max(np.abs(got - want))
        print(f"pty_transmission {kv} kV: max abs error {err:.2e}", flush=True)
        assert err < 1.2e-7
        assert np.array_equal(got, want.astype(np.complex64).astype(np.complex128)) or err < 1.2e-7


def test_golden_stem4d_sharded_through_hip_drop_in():
    """stem4d_sharded (atom slice provider) reproduces the reference's golden capture on the GPU."""
    import ptyrodactyl_amd as pa
    ref = np.load(GOLDEN)
    pot = o.jax_uniform_f64(42, (32, 32, 2)) * 10.0
    mic = pa.create_microscope_config(voltage_kv=100.0, aperture_mrad=20.0, probe_shape=(32, 32))
    det = pa.create_detector_config(real_space_calib_ang=0.2, probe_calibration_pm=20.0)
    pm = pa.create_probe_modes(pa.make_probe(mic, det)[..., None], torch.tensor([1.0]), 0.2)
    sample = pa.create_atomic_slice_data(
        np.array([[2.0, 2.5, 0.25], [4.5, 5.0, 2.25], [3.0, 4.0, 1.25]]), np.zeros(3, int),
        np.array([[0.0, 2.0], [2.0, 4.0]]), pot[..., 0][None])
    positions = torch.tensor([[12.0, 12.0], [12.0, 20.0], [20.0, 12.0], [20.0, 20.0]], dtype=torch.float64) * 0.2
    det2 = pa.create_detector_config(real_space_calib_ang=0.2, probe_calibration_pm=20.0,
                                     scan_positions_ang=positions)
    s4d = pa.stem4d_sharded(pm, sample, mic, det2)
    assert rel_l2(s4d.data.cpu().numpy(), ref["stem4d_sharded"]) < TOL
    chk = pa.checked_stem4d_sharded(pm, sample, mic, det2)
    assert torch.equal(chk.data, s4d.data)


def test_checked_wrappers_are_transparent():
    import ptyrodactyl_amd as pa
    rng = np.random.default_rng(7)
    n = 64
    pot = pa.create_potential_slices(torch.tensor(rng.uniform(0, 10, (n, n, 3))), 1.0, 0.2)
    mic = pa.create_microscope_config(voltage_kv=200.0, aperture_mrad=20.0, probe_shape=(n, n))
    det = pa.create_detector_config(real_space_calib_ang=0.2, scan_positions_px=[[3.0, 4.0], [10.5, 60.2]])
    beam = pa.create_probe_modes(pa.checked_make_probe(mic, det)[..., None], [1.0], 0.2)
    assert torch.equal(pa.checked_cbed_image(pot, beam, mic).data_array, pa.cbed_image(pot, beam, mic).data_array)
    assert torch.equal(pa.checked_stem_4d(pot, beam, mic, det).data, pa.stem_4d(pot, beam, mic, det).data)


@pytest.mark.parametrize("n,s_count,p_count", [(128, 4, 9), (256, 3, 6)])
def test_captured_forward_replays_match_eager(n, s_count, p_count):
    """Graph replays, interleaved with eager work and synchronisations, each equal the eager forward
    bitwise, also after the object is updated in place."""
    from ptyrodactyl_amd import engine, physics
    slab, probes, origins, shifts, _ = _random_case(n, s_count, p_count, (n + 40, n + 30), seed=21)
    dev = torch.device("cuda", 0)
    slab_t = torch.tensor(slab, device=dev)
    args = (slab_t, torch.fft.fft2(torch.tensor(probes, device=dev)).to(torch.complex64).contiguous(),
            torch.ones(1, dtype=torch.float32, device=dev), torch.tensor(origins, device=dev),
            torch.tensor(shifts, device=dev), physics.propagator_factors(n, 1.5, 300.0, 0.1, device=dev))
    cap = engine.CapturedForward(*args)
    for it in range(4):
        if it == 2:
            slab_t.mul_(torch.exp(1j * torch.full_like(slab_t.real, 0.05)).to(slab_t.dtype))  # in-place object update
        eager, _ = engine.multislice_forward(*args)
        torch.cuda.synchronize()
        got = cap.replay().clone()
        torch.cuda.synchronize()
        _ = torch.randn(1000, device=dev).sum().item()  # unrelated eager work between replays
        assert torch.equal(got, eager), f"replay {it} differs"


def test_cbed_from_atoms_matches_oracle_and_sharded_at_origin():
    """cbed_amplitude_from_atoms / cbed_image_from_atoms (parallelized.py:172-345) on the golden capture's atom
    sample: amplitudes vs the oracle's atom provider + multislice, and the image equals stem4d_sharded at a
    zero scan position."""
    import ptyrodactyl_amd as pa
    pot = o.jax_uniform_f64(42, (32, 32, 2)) * 10.0
    coords = np.array([[2.0, 2.5, 0.25], [4.5, 5.0, 2.25], [3.0, 4.0, 1.25]])
    types = np.zeros(3, int)
    bounds = np.array([[0.0, 2.0], [2.0, 4.0]])
    kern = pot[..., 0][None]
    mic = pa.create_microscope_config(voltage_kv=100.0, aperture_mrad=20.0, probe_shape=(32, 32))
    det = pa.create_detector_config(real_space_calib_ang=0.2, probe_calibration_pm=20.0)
    probe = pa.make_probe(mic, det)
    modes = torch.stack([probe, torch.roll(probe, 3, dims=1)], dim=2).to("cuda")
    amp = pa.cbed_amplitude_from_atoms(modes, coords, types, bounds, kern, 100.0, 0.2)
    trans = np.stack([o.transmission_func(o.slice_potential_from_atoms(coords, types, b[0], b[1], kern, (32, 32), 0.2),
                                          100.0) for b in bounds])
    prop = o.propagation_func(32, 32, 2.0, 100.0, 0.2)
    want = np.moveaxis(o.multislice_amplitude(trans, np.moveaxis(modes.cpu().numpy(), 2, 0), prop), 0, 2)
    assert rel_l2(amp.cpu().numpy(), want) < TOL
    dist = pa.create_distribution(torch.tensor([[0.0], [1.0]]), torch.tensor([0.25, 0.75]), "incoherent",
                                  "probe_modes")
    img = pa.cbed_image_from_atoms(modes, dist, coords, types, bounds, kern, 100.0, 0.2)
    want_img = 0.25 * np.abs(want[..., 0]) ** 2 + 0.75 * np.abs(want[..., 1]) ** 2
    assert rel_l2(img.cpu().numpy(), want_img) < TOL
    pm = pa.create_probe_modes(modes, torch.tensor([0.25, 0.75]), 0.2)
    det0 = pa.create_detector_config(real_space_calib_ang=0.2, probe_calibration_pm=20.0,
                                     scan_positions_ang=torch.zeros((1, 2), dtype=torch.float64))
    s4d = pa.stem4d_sharded(pm, pa.create_atomic_slice_data(coords, types, bounds, kern), mic, det0)
    assert rel_l2(s4d.data[0].double().cpu().numpy(), img.cpu().numpy()) < 1e-6


def test_stem4d_sharded_mode_weight_and_kernel_gradients():
    """test_parallelized.py:172-187: the sharded 4D-STEM is differentiable w.r.t. the mode weights (finite,
    nonzero); the weight gradient equals sum g |A_m|^2 of the oracle's amplitudes and the atom-kernel gradient
    matches central differences of the oracle."""
    import ptyrodactyl_amd as pa
    pot = o.jax_uniform_f64(42, (32, 32, 2)) * 10.0
    coords = np.array([[2.0, 2.5, 0.25], [4.5, 5.0, 2.25], [3.0, 4.0, 1.25]])
    types = np.zeros(3, int)
    bounds = np.array([[0.0, 2.0], [2.0, 4.0]])
    mic = pa.create_microscope_config(voltage_kv=100.0, aperture_mrad=20.0, probe_shape=(32, 32))
    det = pa.create_detector_config(real_space_calib_ang=0.2, probe_calibration_pm=20.0)
    probe = pa.make_probe(mic, det).cuda()
    modes = torch.stack([probe, torch.roll(probe, 3, dims=1)], dim=2)
    positions = torch.tensor([[2.4, 2.4], [2.4, 4.0], [4.0, 2.4]], dtype=torch.float64)
    det2 = pa.create_detector_config(real_space_calib_ang=0.2, probe_calibration_pm=20.0,
                                     scan_positions_ang=positions)
    w = torch.tensor([0.4, 0.6], dtype=torch.float64, device="cuda", requires_grad=True)
    kern = torch.tensor(pot[..., 0][None], device="cuda", requires_grad=True)
    g = torch.tensor(np.random.default_rng(3).standard_normal((3, 32, 32)), device="cuda")
    out = pa.stem4d_sharded(pa.ProbeModes(modes, w, 0.2), pa.create_atomic_slice_data(coords, types, bounds, kern),
                            mic, det2)
    gw, gk = torch.autograd.grad(torch.sum(out.data.double() * g), (w, kern))
    assert torch.all(torch.isfinite(gw)) and torch.any(gw != 0)
    gn = g.cpu().numpy()
    want = np.array([np.sum(gn * o.stem4d_from_atoms(modes.cpu().numpy(), sel, coords, types, bounds,
                                                     pot[..., 0][None], 100.0, 0.2, positions.numpy()))
                     for sel in ([1.0, 0.0], [0.0, 1.0])])
    np.testing.assert_allclose(gw.cpu().numpy(), want, rtol=1e-4)

    def oracle_loss(k):
        pats = o.stem4d_from_atoms(modes.cpu().numpy(), [0.4, 0.6], coords, types, bounds, k, 100.0, 0.2,
                                   positions.numpy())
        return float(np.sum(gn * pats))

    h = 1e-3
    for idx in ((0, 3, 5), (0, 17, 20)):
        kp, km = pot[..., 0][None].copy(), pot[..., 0][None].copy()
        kp[idx] += h
        km[idx] -= h
        fd = (oracle_loss(kp) - oracle_loss(km)) / (2 * h)
        assert abs(float(gk[idx]) - fd) <= 1e-2 * abs(fd) + 1e-3 * float(gk.abs().max())


def test_empty_scan_gives_an_empty_stack():
    """vmap over zero positions returns an empty stack in the reference; no kernel is launched here."""
    import ptyrodactyl_amd as pa
    mic = pa.create_microscope_config(voltage_kv=100.0, aperture_mrad=20.0, probe_shape=(32, 32))
    det = pa.create_detector_config(real_space_calib_ang=0.2, probe_calibration_pm=20.0,
                                    scan_positions_px=torch.zeros((0, 2), dtype=torch.float64))
    pm = pa.create_probe_modes(pa.make_probe(mic, det)[..., None], torch.tensor([1.0]), 0.2)
    slices = pa.create_potential_slices(torch.zeros((32, 32, 2), dtype=torch.float64), 1.0, 0.2)
    s4d = pa.stem_4d(slices, pm, mic, det)
    assert tuple(s4d.data.shape) == (0, 32, 32) and tuple(s4d.scan_positions.shape) == (0, 2)
    assert s4d.fourier_space_calib == pytest.approx(1 / 6.4)
