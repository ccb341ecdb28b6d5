"""CPU tests of the drop-in API: signatures, validation tiers and message text
(tests/test_ptyrodactyl/test_multislice/test_checked.py:156-400,
test_simulations.py:59-81 in the reference), and the float64 set-up helpers
against the oracle."""

import inspect

import numpy as np
import pytest
import torch

import ptyrodactyl_amd as pa
from oracle import ms_oracle as o


def test_public_signatures_keep_reference_names_and_order():
    expect = {
        "stem_4d": ["pot_slice", "beam", "microscope", "detector"],
        "cbed_image": ["pot_slices", "beam", "microscope"],
        "cbed_amplitude": ["pot_slices", "beam", "microscope"],
        "make_probe": ["microscope", "detector"],
        "shift_beam_fourier": ["beam", "pos", "calib_ang"],
        "transmission_func": ["pot_slice", "voltage_kv"],
        "propagation_func": ["imsize_y", "imsize_x", "thickness_ang", "voltage_kv", "calib_ang"],
        "aberration": ["fourier_coord", "lambda_angstrom", "defocus", "c3", "c5"],
        "stem4d_sharded": ["probe_modes", "sample", "microscope", "detector", "mesh"],
        "checked_stem_4d": ["pot_slice", "beam", "microscope", "detector"],
        "checked_cbed_image": ["pot_slices", "beam", "microscope"],
        "annular_detector": ["stem4d_data", "detector"],
        "bind_cbed_axes": ["pot_slices", "probe_modes", "microscope", "detector", "axes", "column_maps"],
        "apply_distribution": ["distribution", "bound_amplitude_fn"],
        "apply_distributions": ["distributions", "bound_amplitude_fn"],
        "probe_modes_to_distribution": ["probe"],
        "coherence_to_distribution": ["energy_spread_ev", "angular_divergence_mrad", "n_quad"],
        "position_jitter_to_distribution": ["sigma_ang", "n_quad"],
        "decompose_beam_to_modes": ["beam", "num_modes", "key", "first_mode_weight"],
        "create_distribution": ["samples", "weights", "reduction", "axis_id"],
        "create_ensemble_axes": ["probe_modes", "position_jitter", "coherence"],
        "cbed_amplitude_from_atoms": ["beam", "atom_coords", "atom_types", "slice_z_bounds", "atom_potentials",
                                      "voltage_kv", "calib_ang", "atom_mask"],
        "cbed_image_from_atoms": ["beam", "mode_distribution", "atom_coords", "atom_types", "slice_z_bounds",
                                  "atom_potentials", "voltage_kv", "calib_ang", "atom_mask"],
        "bessel_kv": ["v", "x"],
        "single_atom_potential": ["atom_no", "pixel_size", "grid_shape", "center_coords", "supersampling",
                                  "parameterization"],
        "crystal_potential_slices": ["crystal_data", "pixel_size", "slice_thickness", "repeats", "padding",
                                     "supersampling", "grid_shape", "parameterization"],
        "atomic_form_factor": ["atom_no", "q", "parameterization"],
        "projected_atom_potential": ["atom_no", "r", "parameterization"],
        "lobato_form_factor": ["params", "q"],
        "kirkland_form_factor": ["params", "q"],
        "lobato_projected_potential": ["params", "r"],
        "kirkland_projected_potential": ["params", "r"],
        "lobato_bandlimited_peak": ["atom_no", "g_max"],
        "create_crystal_data": ["positions", "atomic_numbers", "lattice", "stress", "energy", "properties",
                                "comment"],
    }
    many = ("single_atom_potential", "crystal_potential_slices", "create_crystal_data")  # as the reference
    for name, params in expect.items():
        sig = inspect.signature(getattr(pa, name))
        assert list(sig.parameters) == params, name
        if not name.endswith("_from_atoms") and name not in many:  # parallelized.py:172-345, atom_potentials.py
            assert len(sig.parameters) <= 6


def test_make_probe_matches_oracle():
    mic = pa.create_microscope_config(voltage_kv=300.0, aperture_mrad=25.0, defocus_ang=50.0, c3_ang=1e-3,
                                      probe_shape=(64, 64))
    det = pa.create_detector_config(real_space_calib_ang=0.1)
    got = pa.make_probe(mic, det).numpy()
    ref = o.make_probe(300.0, 25.0, 50.0, 1e-3, 0.0, (64, 64), 10.0)
    assert np.allclose(got, ref, atol=1e-14)


def test_setup_helpers_match_oracle():
    rng = np.random.default_rng(0)
    v = rng.uniform(0, 20, (16, 16))
    assert np.allclose(pa.transmission_func(torch.tensor(v), 200.0).numpy(), o.transmission_func(v, 200.0))
    assert np.allclose(pa.propagation_func(16, 24, 1.5, 300.0, 0.1).numpy(), o.propagation_func(16, 24, 1.5, 300.0, 0.1))
    q = rng.uniform(0, 2, (8, 8))
    assert np.allclose(pa.aberration(torch.tensor(q), 0.025, 10.0, 1e-3, 2e-3).numpy(),
                       o.aberration(q, 0.025, 10.0, 1e-3, 2e-3))
    beam = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    pos = np.array([[0.3, -1.1], [2.0, 0.7]])
    assert np.allclose(pa.shift_beam_fourier(torch.tensor(beam), torch.tensor(pos), 0.2).numpy(),
                       o.shift_beam_fourier(beam, pos, 0.2))
    assert pa.relativistic_wavelength_ang(200.0) == pytest.approx(o.wavelength_ang(200.0), rel=1e-14)
    assert pa.interaction_sigma(300.0) == pytest.approx(o.interaction_sigma(300.0), rel=1e-14)


def test_factory_tiers():
    with pytest.raises(ValueError, match="modes must be 3D"):
        pa.create_probe_modes(torch.zeros(4, 4, dtype=torch.complex128), [1.0], 0.1)
    with pytest.raises(ValueError, match="weights must have shape"):
        pa.create_probe_modes(torch.zeros(4, 4, 2, dtype=torch.complex128), [1.0], 0.1)
    with pytest.raises(pa.RuntimeCheckError, match="weights must be non-negative"):
        pa.create_probe_modes(torch.zeros(4, 4, 2, dtype=torch.complex128), [1.0, -0.5], 0.1)
    pm = pa.create_probe_modes(torch.zeros(4, 4, 2, dtype=torch.complex128), [1.0, 3.0], 0.1)
    assert torch.allclose(pm.weights, torch.tensor([0.25, 0.75], dtype=torch.float64))
    with pytest.raises(pa.RuntimeCheckError, match="voltage_kv must be positive"):
        pa.create_microscope_config(voltage_kv=-1.0, aperture_mrad=20.0)
    with pytest.raises(ValueError, match="slices must be 3D"):
        pa.create_potential_slices(torch.zeros(4, 4), 1.0, 0.1)
    with pytest.raises(pa.RuntimeCheckError, match="slices contain non-finite values"):
        pa.create_potential_slices(torch.full((4, 4, 1), float("nan")), 1.0, 0.1)


def test_runtime_checks_can_be_disabled(monkeypatch):
    monkeypatch.setenv("PTYRODACTYL_DISABLE_RUNTIME_CHECKS", "1")
    m = pa.create_microscope_config(voltage_kv=-1.0, aperture_mrad=20.0)
    assert m.voltage_kv == -1.0


def _carriers(n=32):
    pot = pa.create_potential_slices(torch.zeros(n, n, 2, dtype=torch.float64), 2.0, 0.2)
    mic = pa.create_microscope_config(voltage_kv=100.0, aperture_mrad=20.0, probe_shape=(n, n))
    det = pa.create_detector_config(real_space_calib_ang=0.2)
    beam = pa.create_probe_modes(pa.make_probe(mic, det)[..., None], [1.0], 0.2)
    return pot, mic, beam


def test_checked_rejections_before_any_device_work():
    """test_checked.py:195-206, 259-268, 313-324: rejection conditions and messages."""
    pot, mic, beam = _carriers()
    bad = pa.PotentialSlices(torch.full((32, 32, 2), float("nan"), dtype=torch.float64), 2.0, 0.2)
    with pytest.raises(pa.RuntimeCheckError, match="pot_slices.slices contain non-finite values"):
        pa.checked_cbed_image(bad, beam, mic)
    neg = pa.MicroscopeConfig(-100.0, 20.0, 0.0, 0.0, 0.0, pa.EnsembleAxes(), (32, 32))
    with pytest.raises(pa.RuntimeCheckError, match="voltage_kv must be positive"):
        pa.checked_cbed_image(pot, beam, neg)
    det = pa.create_detector_config(real_space_calib_ang=0.2, scan_positions_px=[[5.0, 40.0]])
    with pytest.raises(pa.RuntimeCheckError, match="positions must be within pot_slice grid bounds"):
        pa.checked_stem_4d(pot, beam, mic, det)
    with pytest.raises(ValueError, match="scan_positions_px is required"):
        pa.checked_stem_4d(pot, beam, mic, pa.create_detector_config(real_space_calib_ang=0.2))
    small = pa.create_probe_modes(torch.zeros(16, 16, 1, dtype=torch.complex128), [1.0], 0.2)
    with pytest.raises(ValueError, match="beam.modes spatial shape must match pot_slices.slices"):
        pa.checked_cbed_image(pot, small, mic)


def test_atomic_slice_data_validation():
    with pytest.raises(ValueError, match="atom_coords must have shape"):
        pa.create_atomic_slice_data(np.zeros((3, 2)), np.zeros(3, int), [[0, 1]], np.zeros((1, 4, 4)))
    with pytest.raises(pa.RuntimeCheckError, match="atom_types must index atom_potentials"):
        pa.create_atomic_slice_data(np.zeros((3, 3)), [0, 1, 2], [[0, 1]], np.zeros((2, 4, 4)))
    with pytest.raises(pa.RuntimeCheckError, match="thicknesses must be positive"):
        pa.create_atomic_slice_data(np.zeros((1, 3)), [0], [[1.0, 1.0]], np.zeros((1, 4, 4)))


# The reference's golden vector (tests/test_ptyrodactyl/test_multislice/test_simulations.py:84-122): 4x4 patterns
# arange(96), fourier calib 0.25 1/A, 80 kV, collection 0 - 1000 mrad, raster (2, 3).
ANNULAR_EXPECTED = np.array([[120.0, 376.0, 632.0], [888.0, 1144.0, 1400.0]])


def test_annular_oracle_golden_and_ring():
    data = np.arange(6 * 4 * 4, dtype=np.float64).reshape(6, 4, 4)
    assert np.array_equal(o.annular_detector(data, 0.25, 80.0, 0.0, 1000.0, (2, 3)), ANNULAR_EXPECTED)
    # calib = 1 / (1000 lambda) makes the radius in pixels equal the angle in mrad: the 0.5 - 2.5 mrad ring
    # holds the lattice points with r^2 in {1, 2, 4, 5}: 4 + 4 + 4 + 8 = 20 pixels
    lam = o.wavelength_ang(200.0)
    assert o.annular_detector(np.ones((1, 8, 8)), 1 / (1000 * lam), 200.0, 0.5, 2.5, (1, 1))[0, 0] == 20.0
    # the centre pixel is (H//2, W//2): a 0 - 0.5 mrad disk selects exactly that pixel
    d = np.zeros((1, 7, 6))
    d[0, 3, 3] = 5.0
    d[0, 3, 2] = 7.0
    assert o.annular_detector(d, 1 / (1000 * lam), 200.0, 0.0, 0.5, (1, 1))[0, 0] == 5.0


def test_annular_mask_matches_oracle_mask():
    """The device-independent mask builder reproduces the reference's float64 mask bit for bit."""
    rng = np.random.default_rng(4)
    for _ in range(20):
        h, w = int(rng.integers(4, 40)), int(rng.integers(4, 40))
        calib, kv = float(rng.uniform(0.01, 0.5)), float(rng.choice([80.0, 200.0, 300.0]))
        inner = float(rng.uniform(0, 30))
        outer = inner + float(rng.uniform(0, 60))
        got = pa.simulations.annular_mask((h, w), calib, kv, inner, outer).numpy().astype(bool)
        eye = np.zeros((h * w, h, w))
        eye[np.arange(h * w), np.arange(h * w) // w, np.arange(h * w) % w] = 1.0
        ref = o.annular_detector(eye, calib, kv, inner, outer, (h, w)).astype(bool)
        assert np.array_equal(got, ref)


def test_annular_detector_host_errors():
    stem = pa.create_stem4d(torch.zeros(6, 4, 4, dtype=torch.float32), 0.5, 0.25, torch.zeros(6, 2), 80.0)
    det = pa.create_detector_config(real_space_calib_ang=0.5, collection_inner_mrad=0.0,
                                    collection_outer_mrad=1000.0)
    with pytest.raises(ValueError, match="scan_shape is required"):
        pa.annular_detector(stem, det)
    det = pa.create_detector_config(real_space_calib_ang=0.5, collection_inner_mrad=0.0,
                                    collection_outer_mrad=1000.0, scan_shape=(2, 3))
    with pytest.raises(ValueError, match="GPU tensor"):
        pa.annular_detector(stem, det)


def test_grid_edges_beyond_the_kernels_raise_before_any_device_work():
    """Any grid shape with edges up to 4096 is accepted; larger ones raise ValueError from the host checks
    (before any device work: this runs without a GPU)."""
    import ptyrodactyl_amd as pa
    from ptyrodactyl_amd import simulations
    assert simulations._grid_shape(15, 32) == (15, 32)
    assert simulations._grid_shape(1, 4096) == (1, 4096)
    with pytest.raises(ValueError, match=r"each edge must be in \[1, 4096\]"):
        simulations._grid_shape(4097, 8)
    with pytest.raises(ValueError, match="the GPU kernels take edges of 1 to 4096"):
        pa.crystal2stem4d(pa.create_crystal_data([[0.0, 0.0, 0.0], [3.0, 2.0, 1.0]], [6, 14]), [[1.0, 1.0]],
                          300.0, 20.0, real_space_pixel_size_ang=0.002)
