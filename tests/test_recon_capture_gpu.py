"""Reconstructor iterations recorded in a HIP graph (``PTYRODACTYL_CAPTURE_ITERATIONS=1``) against the eager loop.

Every iteration after the first is a replay of the same graph, with the eager Optax update between replays.
Snapshots are taken after every iteration, so each replay's result is compared with the eager run; the losses
are compared too.  The graph replays the same kernels on the same inputs as the eager iteration, so the
results are expected to agree bitwise (asserted to 1e-12 relative, with the exact-equality count printed)."""

import numpy as np
import pytest

from test_recon_gpu import _experimental, _problem

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _run(monkeypatch, capture, fn):
    import warnings
    if capture:
        monkeypatch.setenv("PTYRODACTYL_CAPTURE_ITERATIONS", "1")
    else:
        monkeypatch.delenv("PTYRODACTYL_CAPTURE_ITERATIONS", raising=False)
    with warnings.catch_warnings(record=True) as seen:
        warnings.simplefilter("always")
        out = fn()
        torch.cuda.synchronize()
    # recording must not leave autograd syncing across streams (an accumulator made on another stream)
    assert not [w for w in seen if "AccumulateGrad" in str(w.message)], [str(w.message)[:120] for w in seen]
    return [o.detach().cpu().clone() if isinstance(o, torch.Tensor) else o for o in out]


def _compare(eager, captured, what):
    exact = 0
    for a, b in zip(eager, captured):
        a, b = np.asarray(a), np.asarray(b)
        assert a.shape == b.shape, what
        exact += int(np.array_equal(a, b))
        scale = max(np.abs(a).max(), 1e-300)
        assert np.abs(a - b).max() <= 1e-12 * scale, what
    print(f"{what}: {exact}/{len(eager)} outputs bitwise equal", flush=True)


@pytest.mark.parametrize("loss_type", ["mse", "mae"])
def test_single_slice_ptychography_capture_matches_eager(monkeypatch, loss_type):
    import ptyrodactyl_amd as pa
    pb = _problem(n=64, p_count=9, seed=7)
    exp = _experimental(pb)
    pot = pa.create_calibrated_array(torch.tensor(pb["pot0"]), 0.5, 0.5, True)
    beam = pa.create_calibrated_array(torch.tensor(pb["beam0"][..., 0]), 0.5, 0.5, True)

    def run():
        p, b, sp, sb = pa.single_slice_ptychography(exp, pot, beam, pb["dz"], save_every=1, num_iterations=6,
                                                    learning_rate=0.01, loss_type=loss_type)
        return p.data_array, b.data_array, sp, sb

    _compare(_run(monkeypatch, False, run), _run(monkeypatch, True, run), f"single_slice_ptychography {loss_type}")


def test_poscorrected_and_multi_slice_capture_match_eager(monkeypatch):
    import ptyrodactyl_amd as pa
    pb = _problem(n=64, p_count=9, seed=8)
    exp = _experimental(pb)
    pot = pa.create_calibrated_array(torch.tensor(pb["pot0"]), 0.5, 0.5, True)
    beam = pa.create_calibrated_array(torch.tensor(pb["beam0"][..., 0]), 0.5, 0.5, True)

    def run_pos():
        out = pa.single_slice_poscorrected(exp, pot, beam, pb["dz"], save_every=1, num_iterations=5,
                                           learning_rate=0.01)
        return [out[0].data_array, out[1].data_array] + list(out[2:])

    _compare(_run(monkeypatch, False, run_pos), _run(monkeypatch, True, run_pos), "single_slice_poscorrected")

    def run_ms():
        return pa.multi_slice_multi_modal(exp, torch.tensor(pb["pot0"]), torch.tensor(pb["beam0"][..., 0]),
                                          pb["dz"], save_every=1, num_iterations=5, learning_rate=0.005,
                                          pos_learning_rate=0.05)

    _compare(_run(monkeypatch, False, run_ms), _run(monkeypatch, True, run_ms), "multi_slice_multi_modal")


def test_capture_on_runtime_size_grid_matches_eager(monkeypatch):
    """Capture also works on a grid that is not a square power of two (30 x 44, the runtime-size kernels).  Their
    adjoint accumulates with float32 atomics, so captured and eager runs agree to rounding, not bitwise: the
    parameters after 4 iterations within 1e-5 of their scale."""
    import ptyrodactyl_amd as pa
    pb = _problem(n=30, w=44, p_count=6, seed=9)
    exp = _experimental(pb)
    pot = pa.create_calibrated_array(torch.tensor(pb["pot0"]), 0.5, 0.5, True)
    beam = pa.create_calibrated_array(torch.tensor(pb["beam0"][..., 0]), 0.5, 0.5, True)

    def run():
        p, b, _, _ = pa.single_slice_ptychography(exp, pot, beam, pb["dz"], save_every=2, num_iterations=4,
                                                  learning_rate=0.01, loss_type="mse")
        return p.data_array, b.data_array

    eager = _run(monkeypatch, False, run)
    captured = _run(monkeypatch, True, run)
    for a, b in zip(eager, captured):
        a, b = np.asarray(a), np.asarray(b)
        assert a.shape == b.shape == (30, 44)
        err = np.abs(a - b).max() / np.abs(a).max()
        print(f"capture vs eager on 30x44: max difference {err:.2e} of the scale", flush=True)
        assert err < 1e-5
