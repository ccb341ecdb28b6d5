"""GPU checks of the virtual annular detector (``pty_annular_integrate`` behind ``annular_detector``)."""

import numpy as np
import pytest

from oracle import ms_oracle as o

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

# reference golden vector: tests/test_ptyrodactyl/test_multislice/test_simulations.py:84-122
ANNULAR_EXPECTED = np.array([[120.0, 376.0, 632.0], [888.0, 1144.0, 1400.0]])


def test_golden_vector_through_drop_in():
    import ptyrodactyl_amd as pa
    data = torch.arange(6 * 4 * 4, dtype=torch.float32, device="cuda").reshape(6, 4, 4)
    positions = torch.tensor([[0.0, 0.0], [0.0, 1.0], [0.0, 2.0], [1.0, 0.0], [1.0, 1.0], [1.0, 2.0]])
    stem = pa.create_stem4d(data, 0.5, 0.25, positions, 80.0)
    det = pa.create_detector_config(real_space_calib_ang=0.5, collection_inner_mrad=0.0,
                                    collection_outer_mrad=1000.0, scan_shape=(2, 3))
    img = pa.annular_detector(stem, det)
    assert img.real_space and img.calib_y == 0.5 and img.calib_x == 0.5
    assert np.array_equal(img.data_array.cpu().numpy(), ANNULAR_EXPECTED)


@pytest.mark.parametrize("n,p,inner,outer", [(64, 37, 5.0, 20.0), (128, 20, 0.0, 12.0), (256, 9, 30.0, 90.0)])
def test_matches_oracle_on_random_patterns(n, p, inner, outer):
    import ptyrodactyl_amd as pa
    rng = np.random.default_rng(n + p)
    pats = rng.uniform(0, 1, (p, n, n)).astype(np.float32)
    stem = pa.create_stem4d(torch.tensor(pats, device="cuda"), 0.1, 1.0 / (n * 0.1), torch.zeros(p, 2), 300.0)
    det = pa.create_detector_config(real_space_calib_ang=0.1, collection_inner_mrad=inner,
                                    collection_outer_mrad=outer, scan_shape=(1, p))
    got = pa.annular_detector(stem, det).data_array.cpu().numpy()
    ref = o.annular_detector(pats.astype(np.float64), 1.0 / (n * 0.1), 300.0, inner, outer, (1, p))
    np.testing.assert_allclose(got, ref, rtol=1e-12)
    again = pa.annular_detector(stem, det).data_array.cpu().numpy()
    assert np.array_equal(got, again)


def test_vacuum_bright_field_counts_the_aperture():
    """V = 0: every pattern is the aperture indicator, so a detector that covers the aperture collects
    exactly the number of aperture pixels, and a dark-field annulus outside it collects nothing."""
    import ptyrodactyl_amd as pa
    n, calib, kv, ap = 64, 0.2, 200.0, 20.0
    pot = pa.create_potential_slices(torch.zeros(n, n, 2, dtype=torch.float64), 1.0, calib)
    mic = pa.create_microscope_config(voltage_kv=kv, aperture_mrad=ap, probe_shape=(n, n))
    det0 = pa.create_detector_config(real_space_calib_ang=calib)
    probe = pa.make_probe(mic, det0)
    beam = pa.create_probe_modes(probe[..., None], torch.tensor([1.0]), calib)
    pos = torch.tensor([[1.0, 2.0], [3.0, 4.0], [5.0, 1.0], [2.0, 2.0]], dtype=torch.float64)
    det = pa.create_detector_config(real_space_calib_ang=calib, scan_positions_px=pos)
    stem = pa.stem_4d(pot, beam, mic, det)
    pats = stem.data.cpu().numpy()
    aperture_pixels = float(np.count_nonzero(pats[0] > 0.5))
    bf = pa.create_detector_config(real_space_calib_ang=calib, collection_inner_mrad=0.0,
                                   collection_outer_mrad=2 * ap, scan_shape=(2, 2))
    df = pa.create_detector_config(real_space_calib_ang=calib, collection_inner_mrad=1.5 * ap,
                                   collection_outer_mrad=4 * ap, scan_shape=(2, 2))
    np.testing.assert_allclose(pa.annular_detector(stem, bf).data_array.cpu().numpy(), aperture_pixels, rtol=1e-5)
    assert float(pa.annular_detector(stem, df).data_array.abs().max()) < 1e-4


@pytest.mark.parametrize("n,s_count,p_count,modes", [(64, 3, 11, 1), (128, 4, 6, 2), (256, 3, 10, 1), (256, 2, 5, 2)])
def test_fused_detectors_match_integrating_the_patterns(n, s_count, p_count, modes):
    """Detector signals from the kernel epilogue equal annular integration of the stored patterns, with and
    without writing the patterns, and are bitwise reproducible."""
    from ptyrodactyl_amd import engine, physics, simulations
    print(f"fused detectors n={n} S={s_count} P={p_count} M={modes}", flush=True)
    rng = np.random.default_rng(n + p_count)
    hw = (n + 40, n + 24)
    slab = np.exp(1j * rng.uniform(0, 0.4, (s_count,) + hw)).astype(np.complex64)
    probes = np.stack([o.make_probe(300.0, 20.0 + 3 * m, 30.0 * m, 0.0, 0.0, (n, n), 10.0) for m in range(modes)])
    dev = torch.device("cuda", 0)
    args = (torch.tensor(slab, device=dev),
            torch.fft.fft2(torch.tensor(probes, device=dev)).to(torch.complex64).contiguous(),
            torch.full((modes,), 1.0 / modes, dtype=torch.float32, device=dev),
            torch.tensor(np.stack([rng.integers(0, hw[0], p_count), rng.integers(0, hw[1], p_count)], 1),
                         dtype=torch.int32, device=dev),
            torch.tensor(rng.uniform(-2, 2, (p_count, 2)), device=dev),
            physics.propagator_factors(n, 1.5, 300.0, 0.1, device=dev))
    angles = [(0.0, 10.0), (10.0, 25.0), (30.0, 80.0)]
    masks = torch.stack([simulations.annular_mask((n, n), 1.0 / (n * 0.1), 300.0, lo, hi, device=dev)
                         for lo, hi in angles]).contiguous()
    pats, sig = engine.multislice_forward_detectors(*args, masks, want_patterns=True)
    _, sig_only = engine.multislice_forward_detectors(*args, masks)
    _, sig_again = engine.multislice_forward_detectors(*args, masks)
    plain, _ = engine.multislice_forward(*args)
    torch.cuda.synchronize()
    assert torch.equal(pats, plain)
    assert torch.equal(sig_only, sig_again)
    ref = (pats.double()[:, None] * masks.double()[None]).sum(dim=(2, 3))
    np.testing.assert_allclose(sig.cpu().numpy(), ref.cpu().numpy(), rtol=1e-6)
    np.testing.assert_allclose(sig_only.cpu().numpy(), ref.cpu().numpy(), rtol=1e-6)
