"""Crystal -> 4D-STEM workflow host logic (``workflows/stem_4d.py`` in the reference): the antialiased linear
resize against torch's independent antialiased bilinear interpolation and its closed forms, the memory estimate,
the FFT-grid choice and argument validation.  No GPU needed."""

import numpy as np
import pytest
import torch

import ptyrodactyl_amd as pa
from ptyrodactyl_amd import workflows as wf
from oracle import workflow_oracle as wo


@pytest.mark.parametrize("n_in, n_out", [(66, 32), (66, 100), (64, 64), (17, 5), (5, 17), (130, 256), (96, 48)])
def test_resize_weights_match_the_oracle(n_in, n_out):
    np.testing.assert_allclose(wf.resize_weights(n_in, n_out).numpy(), wo.resize_weights(n_in, n_out),
                               rtol=0, atol=1e-15)


@pytest.mark.parametrize("shape_in, shape_out", [((66, 66), (32, 32)), ((66, 66), (256, 256)), ((40, 52), (17, 80)),
                                                 ((33, 33), (33, 12))])
def test_resize_equals_torch_antialiased_bilinear(shape_in, shape_out):
    """torch's antialiased bilinear (PIL-style support) is an independent implementation of the same resampler."""
    rng = np.random.default_rng(0)
    img = torch.as_tensor(rng.uniform(0, 1, (3, *shape_in)))
    got = wf.resize_linear(img, shape_out)
    want = torch.nn.functional.interpolate(img[:, None], size=shape_out, mode="bilinear", align_corners=False,
                                           antialias=True)[:, 0]
    np.testing.assert_allclose(got.numpy(), want.numpy(), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(got.numpy(), wo.resize_linear(img.numpy(), shape_out), rtol=1e-13, atol=1e-14)


def test_resize_closed_forms():
    w = wo.resize_weights(16, 8)
    np.testing.assert_allclose(w[3:7, 2], [0.125, 0.375, 0.375, 0.125], atol=1e-15)   # 2x reduction, taps 3-6
    np.testing.assert_allclose(w[0:3, 0], np.asarray([0.75, 0.75, 0.25]) / 1.75, atol=1e-15)
    img = np.random.default_rng(1).uniform(size=(9, 9))
    assert np.array_equal(wo.resize_linear(img, (9, 9)), img)
    ramp = np.tile(np.arange(10.0), (3, 1))
    up = wo.resize_linear(ramp, (3, 20))
    interior = (np.arange(20) + 0.5) / 2.0 - 0.5
    np.testing.assert_allclose(up[0, 1:-1], interior[1:-1], atol=1e-14)
    np.testing.assert_allclose(wo.resize_linear(np.ones((30, 30)), (7, 11)), 1.0, atol=1e-15)


def test_memory_estimate_and_fourier_grid():
    """``stem_4d.py:57-138`` and ``:735-760``."""
    assert wf.estimate_memory_gb(4096, 256, 256, 1, 64) == pytest.approx(wo.estimate_memory_gb(4096, 256, 256, 1, 64))
    assert wf.estimate_memory_gb(4096, 256, 256, 1, 64) == pytest.approx((4096 * 65536 * 24 + 65536 * 64 * 8) * 2.5
                                                                         / 2 ** 30)
    assert wf.fourier_pixels_for(300.0, 50.0, (256, 256), 0.1, 64) == 512
    assert wf.fourier_pixels_for(300.0, 50.0, (256, 256), 0.02, 4096) == 4096
    assert wf.fourier_pixels_for(200.0, 40.0, (64, 64), 0.1, 32) == wo.fft_pixels_for(200.0, 40.0, (64, 64), 0.1, 32)


def test_workflow_argument_validation():
    crystal = pa.create_crystal_data([[0.0, 0.0, 0.0], [3.0, 2.0, 1.0]], [6, 14])
    with pytest.raises(ValueError, match="Tile size"):
        pa.crystal2stem4d_tiled(crystal, [[1.0, 1.0]], 300.0, 20.0, tile_size_ang=10.0, grid_pixels=64,
                                pixel_size_ang=0.1)
    with pytest.raises(ValueError, match=r"fourier_pixels \(64\) must be >= grid_pixels \(128\)"):
        pa.crystal2stem4d_tiled(crystal, [[1.0, 1.0]], 300.0, 20.0, tile_size_ang=5.0, grid_pixels=128,
                                pixel_size_ang=0.1, fourier_pixels=64)
    with pytest.raises(ValueError, match="the GPU kernels take edges of 1 to 4096"):
        pa.crystal2stem4d_tiled(crystal, [[1.0, 1.0]], 300.0, 20.0, grid_pixels=5000)   # FFT grid 5000
    with pytest.raises(ValueError, match="the GPU kernels take edges of 1 to 4096"):
        pa.crystal2stem4d(crystal, [[1.0, 1.0]], 300.0, 20.0,
                          real_space_pixel_size_ang=0.002)             # 5500 x 5000 pixels
