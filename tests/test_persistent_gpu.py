"""GPU checks of the single-launch persistent forward (n = 256, one probe mode).

It applies each slice's row operator before its column operator, while the
streaming forward alternates the order slice by slice (the two operators
commute), so the two paths agree to float32 rounding (AGREE below) rather
than bitwise; the oracle comparison uses the forward path's tolerance
(relative L2 <= 1e-5).
"""

import numpy as np
import pytest

from oracle import ms_oracle as o

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL = 1e-5
AGREE = 2e-6


def _agree(got, ref, what):
    err = float(torch.linalg.vector_norm(got - ref) / torch.linalg.vector_norm(ref))
    print(f"{what}: persistent vs streaming rel-L2 {err:.3e}", flush=True)
    assert err < AGREE


def _case(s_count, p_count, slab_hw, seed):
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0, 0.4, (s_count,) + slab_hw)
    amp = rng.uniform(0.95, 1.0, (s_count,) + slab_hw)
    slab = (amp * np.exp(1j * phase)).astype(np.complex64)
    probe = o.make_probe(300.0, 20.0, 30.0, 0.0, 0.0, (256, 256), 10.0)
    origins = np.stack([rng.integers(0, slab_hw[0], p_count), rng.integers(0, slab_hw[1], p_count)], 1)
    shifts = rng.uniform(-3.0, 3.0, (p_count, 2))
    return slab, probe, origins.astype(np.int32), shifts


def _device_args(slab, probe, origins, shifts):
    from ptyrodactyl_amd import physics
    dev = torch.device("cuda", 0)
    return (torch.tensor(slab, device=dev),
            torch.fft.fft2(torch.tensor(probe[None], device=dev)).to(torch.complex64).contiguous(),
            torch.ones(1, dtype=torch.float32, device=dev), torch.tensor(origins, device=dev),
            torch.tensor(shifts, device=dev), physics.propagator_factors(256, 1.5, 300.0, 0.1, device=dev))


@pytest.mark.parametrize("s_count,p_count,group_waves", [(1, 3, 2), (3, 37, 2), (2, 20, 1), (4, 50, 3)])
def test_persistent_agrees_with_streaming(s_count, p_count, group_waves):
    from ptyrodactyl_amd import engine
    print(f"persistent S={s_count} P={p_count} g={group_waves}", flush=True)
    args = _device_args(*_case(s_count, p_count, (300, 280), seed=s_count * 100 + p_count))
    ref, _ = engine.multislice_forward(*args)
    got = engine.multislice_forward_persistent(*args, group_waves=group_waves)
    torch.cuda.synchronize()
    _agree(got, ref, f"S={s_count}")


def test_persistent_matches_oracle():
    from ptyrodactyl_amd import engine
    slab, probe, origins, shifts = _case(3, 4, (256, 300), seed=7)
    got = engine.multislice_forward_persistent(*_device_args(slab, probe, origins, shifts))
    prop = o.propagation_func(256, 256, 1.5, 300.0, 0.1)
    ref = o.window_stem(slab.astype(np.complex128), probe, origins, shifts, prop)
    err = float(np.linalg.norm(got.cpu().numpy() - ref) / np.linalg.norm(ref))
    print(f"persistent vs oracle rel-L2 {err:.3e}", flush=True)
    assert err < TOL


def test_persistent_reuses_state_and_workspace():
    """Two calls in a row on shared state and workspace, the second smaller: the second call is bitwise equal to
    the same call on fresh state, and each agrees with the streaming path."""
    from ptyrodactyl_amd import engine
    big = _device_args(*_case(2, 40, (256, 256), seed=11))
    small = _device_args(*_case(2, 9, (256, 256), seed=12))
    state = torch.empty(engine._native.load().pty_persistent_state_bytes() // 4, dtype=torch.int32, device="cuda")
    ws = torch.empty(engine._native.load().pty_persistent_workspace_bytes(256, 2), dtype=torch.uint8, device="cuda")
    a = engine.multislice_forward_persistent(*big, state=state, workspace=ws)
    b = engine.multislice_forward_persistent(*small, state=state, workspace=ws)
    _agree(a, engine.multislice_forward(*big)[0], "first call")
    _agree(b, engine.multislice_forward(*small)[0], "second call")
    assert torch.equal(b, engine.multislice_forward_persistent(*small))


def test_persistent_config2_slice_count():
    """64 slices on the 1024^2 slab of the headline workload, 256 positions: agrees with the streaming path."""
    from ptyrodactyl_amd import engine, synth
    w = synth.config2(scan=16)
    args = _device_args(w.slab_t, w.probe, w.origins, w.shifts_px)
    from ptyrodactyl_amd import physics
    args = args[:5] + (physics.propagator_factors(256, w.slice_thickness, w.voltage_kv, w.calib, device="cuda"),)
    print("persistent config2 subset", flush=True)
    ref, _ = engine.multislice_forward(*args)
    got = engine.multislice_forward_persistent(*args)
    _agree(got, ref, "config 2, 64 slices")


def test_persistent_rejects_unsupported_shapes():
    from ptyrodactyl_amd import engine
    slab, probe, origins, shifts = _case(1, 2, (256, 256), seed=1)
    args = list(_device_args(slab, probe, origins, shifts))
    args[1] = torch.cat([args[1], args[1]])
    args[2] = torch.ones(2, dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        engine.multislice_forward_persistent(*args)
