"""GPU parity of the ensemble-axis path of cbed_image / stem_4d.

All (position, sample) waves run in one batched HIP launch (pty_forward_batched
with per-wave probes, energy-grouped transmissions and propagators, and an
amplitude output for coherent axes).  Compared with the complex128 oracle
(``oracle/ensemble_oracle.py``) at relative L2 1e-5, with the per-sample
binder (``bind_cbed_axes`` + ``apply_distributions``), and with the plain
path for zero-width axes.
"""

import numpy as np
import pytest

from oracle import ensemble_oracle as eo
from oracle import ms_oracle as o

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL = 1e-5
DEV = "cuda"


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _case(n=32, calib=0.3, kv=120.0, seed=0, modes=2):
    rng = np.random.default_rng(seed)
    pot = rng.uniform(0, 25, (n, n, 3))
    probes = [o.make_probe(kv, 22.0, 20.0 + 15 * m, 0.0, 0.0, (n, n), calib * 100.0) for m in range(modes)]
    return pot, np.stack(probes, axis=2), np.array([0.6, 0.4][:modes]), calib, kv


def _objects(pot, modes, weights, calib, kv, jitter=None, coherence=None, explicit_modes=None):
    import ptyrodactyl_amd as pa
    ens = pa.create_ensemble_axes(probe_modes=explicit_modes, position_jitter=jitter, coherence=coherence)
    mic = pa.create_microscope_config(kv, 22.0, ensemble=ens)
    pots = pa.create_potential_slices(torch.tensor(pot, device=DEV), 1.2, calib)
    beam = pa.create_probe_modes(torch.tensor(modes, device=DEV), torch.tensor(weights), calib)
    return pots, beam, mic


def test_cbed_image_jitter_and_coherence_match_oracle():
    from ptyrodactyl_amd import producers
    import ptyrodactyl_amd as pa
    pot, modes, w, calib, kv = _case()
    jit, coh = producers.position_jitter_to_distribution(0.25, 2), producers.coherence_to_distribution(0.8, 0.3, 2)
    pots, beam, mic = _objects(pot, modes, w, calib, kv, jit, coh)
    got = pa.cbed_image(pots, beam, mic).data_array.cpu().numpy()
    want = eo.cbed_image(pot, 1.2, calib, modes, w, calib, kv, eo.position_jitter(0.25, 2), eo.coherence(0.8, 0.3, 2))
    err = rel_l2(got, want)
    print(f"ensemble cbed_image (2 modes x 4 jitter x 8 coherence): rel-L2 {err:.2e}", flush=True)
    assert err < TOL


def test_stem4d_with_jitter_matches_oracle():
    import ptyrodactyl_amd as pa
    from ptyrodactyl_amd import producers
    pot, modes, w, calib, kv = _case(seed=1)
    jit = producers.position_jitter_to_distribution(0.2, 3)
    pots, beam, mic = _objects(pot, modes, w, calib, kv, jitter=jit)
    positions = np.array([[0.5, 1.0], [2.0, -1.5], [-0.7, 0.3], [3.1, 2.2], [0.0, 0.0]])
    det = pa.create_detector_config(real_space_calib_ang=calib, scan_positions_px=torch.tensor(positions))
    out = pa.stem_4d(pots, beam, mic, det)
    want = eo.stem_4d(pot, 1.2, calib, modes, w, calib, kv, positions, calib, jitter=eo.position_jitter(0.2, 3))
    err = rel_l2(out.data.cpu().numpy(), want)
    print(f"ensemble stem_4d (5 positions x 2 modes x 9 jitter): rel-L2 {err:.2e}", flush=True)
    assert err < TOL


def test_coherent_axis_uses_kernel_amplitudes():
    import ptyrodactyl_amd as pa
    from ptyrodactyl_amd import distributions as dist
    pot, modes, w, calib, kv = _case(seed=2, modes=1)
    samples = np.array([[0.0, 0.0], [0.35, -0.2], [-0.1, 0.4]])
    cw = np.array([0.5, 0.3, 0.2])
    jit = dist.create_distribution(samples, cw, "coherent", "position_jitter")
    pots, beam, mic = _objects(pot, modes, w, calib, kv, jitter=jit)
    got = pa.cbed_image(pots, beam, mic).data_array.cpu().numpy()
    want = eo.cbed_image(pot, 1.2, calib, modes, w, calib, kv, jitter=eo.distribution(samples, cw, "coherent",
                                                                                        "position_jitter"))
    err = rel_l2(got, want)
    print(f"coherent jitter axis: rel-L2 {err:.2e}", flush=True)
    assert err < TOL


def test_batched_path_equals_per_sample_binder():
    import ptyrodactyl_amd as pa
    from ptyrodactyl_amd import producers, reduce, simulations
    pot, modes, w, calib, kv = _case(seed=3)
    jit, coh = producers.position_jitter_to_distribution(0.3, 2), producers.coherence_to_distribution(0.5, 0.2, 2)
    pots, beam, mic = _objects(pot, modes, w, calib, kv, jit, coh)
    fast = pa.cbed_image(pots, beam, mic).data_array
    axes = simulations._ensemble_axes_for_cbed(mic.ensemble, beam)
    det = pa.create_detector_config(real_space_calib_ang=calib)
    slow = reduce.apply_distributions(axes, simulations.bind_cbed_axes(pots, beam, mic, det, axes))
    err = rel_l2(fast.cpu().numpy(), slow.cpu().numpy())
    print(f"batched vs per-sample binder: rel-L2 {err:.2e}", flush=True)
    assert err < 1e-6


def test_zero_width_axes_equal_plain_path():
    import ptyrodactyl_amd as pa
    from ptyrodactyl_amd import producers
    pot, modes, w, calib, kv = _case(seed=4)
    pots, beam, mic = _objects(pot, modes, w, calib, kv, producers.position_jitter_to_distribution(0.0, 1),
                               producers.coherence_to_distribution(0.0, 0.0, 1))
    ens = pa.cbed_image(pots, beam, mic).data_array.cpu().numpy()
    _, plain_beam, plain_mic = _objects(pot, modes, w, calib, kv)
    plain = pa.cbed_image(pots, plain_beam, plain_mic).data_array.cpu().numpy()
    assert rel_l2(ens, plain) < 1e-6


def test_ensemble_cbed_image_gradient_wrt_potential():
    """With grad enabled the ensemble runs its waves in one batched launch and one batched adjoint."""
    import ptyrodactyl_amd as pa
    from ptyrodactyl_amd import producers
    pot, modes, w, calib, kv = _case(seed=5, modes=1)
    rng = np.random.default_rng(6)
    g = rng.standard_normal((32, 32))
    v = torch.tensor(pot, device=DEV, requires_grad=True)
    ens = pa.create_ensemble_axes(position_jitter=producers.position_jitter_to_distribution(0.2, 2))
    mic = pa.create_microscope_config(kv, 22.0, ensemble=ens)
    beam = pa.create_probe_modes(torch.tensor(modes, device=DEV), torch.tensor(w), calib)
    img = pa.cbed_image(pa.create_potential_slices(v, 1.2, calib), beam, mic).data_array
    loss = torch.sum(img.double() * torch.tensor(g, device=DEV))
    (gv,) = torch.autograd.grad(loss, v)
    gv = gv.cpu().numpy()
    jit = eo.position_jitter(0.2, 2)

    def oracle_loss(p):
        return float(np.sum(g * eo.cbed_image(p, 1.2, calib, modes, w, calib, kv, jitter=jit)))

    h = 1e-3
    worst = 0.0
    for idx in [(3, 4, 0), (16, 17, 1), (30, 2, 2), (10, 20, 1)]:
        pp, pm = pot.copy(), pot.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (oracle_loss(pp) - oracle_loss(pm)) / (2 * h)
        worst = max(worst, abs(gv[idx] - fd) / (abs(fd) + 1e-3 * np.abs(gv).max()))
    print(f"ensemble dL/dV vs oracle FD: worst {worst:.2e}", flush=True)
    assert worst < 1e-2


def test_random_normal_matches_jax_restatement():
    from ptyrodactyl_amd import random as prandom
    from oracle import recon_oracle as ro
    for seed, shape in ((0, (16, 3)), (11, (4, 4)), (2**33 + 7, (1000,))):
        for key_p, key_o in zip(prandom.split(prandom.PRNGKey(seed), 2), ro.jax_split(ro.jax_key(seed), 2)):
            got = prandom.normal(key_p, shape, DEV).cpu().numpy()
            want = ro.jax_normal_f64(key_o, shape)
            np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)


def test_decompose_beam_to_modes_pinned_value():
    """test_simulations.py:140-189: PRNGKey(0), 3 modes of arange(16) + 1j, first weight 0.6."""
    import ptyrodactyl_amd as pa
    from ptyrodactyl_amd import random as prandom
    beam = pa.create_calibrated_array(torch.arange(16, dtype=torch.float64, device=DEV).reshape(4, 4) + 1j,
                                      0.5, 0.5, True)
    first = pa.decompose_beam_to_modes(beam, 3, prandom.PRNGKey(0), first_mode_weight=0.6)
    second = pa.decompose_beam_to_modes(beam, 3, prandom.PRNGKey(0), first_mode_weight=0.6)
    assert torch.equal(first.modes, second.modes)
    value = complex(first.modes[0, 0, 0])
    print(f"decompose_beam_to_modes modes[0,0,0] = {value!r}", flush=True)
    assert abs(value - (-0.34562459928563927 + 0.25753846176370626j)) < 1e-13
    np.testing.assert_array_equal(first.weights.cpu().numpy(), [0.6, 0.2, 0.2])
    assert first.calib == 0.5
    with pytest.raises(TypeError):
        pa.decompose_beam_to_modes(beam, 3)


def test_cbed_image_gradient_wrt_jitter_width():
    """d/dsigma of the summed ensemble pattern through the Gauss-Hermite samples (the reference differentiates
    the same, test_producers.py:160-228); compared with central differences of the oracle."""
    import ptyrodactyl_amd as pa
    from ptyrodactyl_amd import producers
    pot, modes, w, calib, kv = _case(seed=7, modes=1)
    rng = np.random.default_rng(8)
    g = rng.standard_normal((32, 32))
    beam = pa.create_probe_modes(torch.tensor(modes, device=DEV), torch.tensor(w), calib)
    pots = pa.create_potential_slices(torch.tensor(pot, device=DEV), 1.2, calib)
    sigma = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    ens = pa.create_ensemble_axes(position_jitter=producers.position_jitter_to_distribution(sigma, 2))
    img = pa.cbed_image(pots, beam, pa.create_microscope_config(kv, 22.0, ensemble=ens)).data_array
    (gs,) = torch.autograd.grad(torch.sum(img.double() * torch.tensor(g, device=DEV)), sigma)

    def oracle_loss(s):
        return float(np.sum(g * eo.cbed_image(pot, 1.2, calib, modes, w, calib, kv, jitter=eo.position_jitter(s, 2))))

    h = 1e-4
    fd = (oracle_loss(0.3 + h) - oracle_loss(0.3 - h)) / (2 * h)
    print(f"d/dsigma: GPU {float(gs):.6e} oracle FD {fd:.6e}", flush=True)
    assert abs(float(gs) - fd) <= 1e-3 * abs(fd) + 1e-9


def test_stem4d_ensemble_gradient_wrt_positions_and_potential():
    """stem_4d with a jitter axis and grad enabled (batched forward with checkpoints, batched adjoint): the loss
    and its gradients agree with the no-grad forward and with oracle central differences."""
    import ptyrodactyl_amd as pa
    from ptyrodactyl_amd import producers
    pot, modes, w, calib, kv = _case(seed=9, modes=1)
    rng = np.random.default_rng(10)
    pos0 = np.array([[0.4, -0.3], [1.1, 0.8]])
    g = rng.standard_normal((2, 32, 32))
    beam = pa.create_probe_modes(torch.tensor(modes, device=DEV), torch.tensor(w), calib)
    ens = pa.create_ensemble_axes(position_jitter=producers.position_jitter_to_distribution(0.2, 2))
    mic = pa.create_microscope_config(kv, 22.0, ensemble=ens)
    v = torch.tensor(pot, device=DEV, requires_grad=True)
    pos = torch.tensor(pos0, device=DEV, requires_grad=True)
    det = pa.create_detector_config(real_space_calib_ang=calib, scan_positions_px=pos)
    out = pa.stem_4d(pa.create_potential_slices(v, 1.2, calib), beam, mic, det)
    loss = torch.sum(out.data.double() * torch.tensor(g, device=DEV))
    g_v, g_pos = torch.autograd.grad(loss, (v, pos))
    with torch.no_grad():
        fast = pa.stem_4d(pa.create_potential_slices(v.detach(), 1.2, calib), beam, mic,
                          pa.create_detector_config(real_space_calib_ang=calib, scan_positions_px=pos.detach())).data
    assert rel_l2(out.data.detach().cpu().numpy(), fast.cpu().numpy()) < 1e-6
    jit = eo.position_jitter(0.2, 2)

    def oracle_loss(p_arr, positions):
        return float(np.sum(g * eo.stem_4d(p_arr, 1.2, calib, modes, w, calib, kv, positions, calib, jitter=jit)))

    h = 1e-4
    for i, ax in ((0, 0), (1, 1)):
        pp, pm = pos0.copy(), pos0.copy()
        pp[i, ax] += h
        pm[i, ax] -= h
        fd = (oracle_loss(pot, pp) - oracle_loss(pot, pm)) / (2 * h)
        assert abs(float(g_pos[i, ax]) - fd) <= 1e-3 * abs(fd) + 1e-8
    idx = (12, 20, 1)
    pp, pm = pot.copy(), pot.copy()
    pp[idx] += 1e-3
    pm[idx] -= 1e-3
    fd = (oracle_loss(pp, pos0) - oracle_loss(pm, pos0)) / 2e-3
    assert abs(float(g_v[idx]) - fd) <= 1e-2 * abs(fd) + 1e-3 * float(g_v.abs().max())


def test_batched_ensemble_gradients_equal_the_per_sample_binder(monkeypatch):
    """The batched differentiable path (BatchedWavesFunction) against the per-sample binder (one launch
    sequence per (position, sample)) for the same stem_4d loss: gradients w.r.t. the potential, the modes, the
    mode weights, the scan positions and the jitter width, with two modes, a 2 x 2 jitter and 2^3 samples of
    partial coherence (energy and tilts).  rel-L2 1e-5 (the batched adjoint adds waves with float32 atomics);
    5e-5 for the position and width gradients, and the weight gradient on the loss's scale (see below)."""
    import ptyrodactyl_amd as pa
    from ptyrodactyl_amd import producers, simulations
    pot, modes, w, calib, kv = _case(seed=12, modes=2)
    rng = np.random.default_rng(13)
    pos0 = np.array([[0.4, -0.3], [1.1, 0.8], [-0.6, 0.2]])
    g = torch.tensor(rng.standard_normal((3, 32, 32)), device=DEV)

    def run():
        v = torch.tensor(pot, device=DEV, requires_grad=True)
        m = torch.tensor(modes, device=DEV, requires_grad=True)
        wt = torch.tensor(w, device=DEV, requires_grad=True)
        pos = torch.tensor(pos0, device=DEV, requires_grad=True)
        sigma = torch.tensor(0.25, dtype=torch.float64, requires_grad=True)
        ens = pa.create_ensemble_axes(position_jitter=producers.position_jitter_to_distribution(sigma, 2),
                                      coherence=producers.coherence_to_distribution(0.8, 0.3, 2))
        mic = pa.create_microscope_config(kv, 22.0, ensemble=ens)
        det = pa.create_detector_config(real_space_calib_ang=calib, scan_positions_px=pos)
        out = pa.stem_4d(pa.create_potential_slices(v, 1.2, calib), pa.ProbeModes(m, wt, calib), mic, det)
        loss = torch.sum(out.data.double() * g)
        grads = torch.autograd.grad(loss, (v, m, wt, pos, sigma))
        return float(loss), [t.detach().cpu().numpy() for t in (out.data, *grads)]

    loss_b, batched = run()
    monkeypatch.setattr(simulations, "_batched_gradient_ok", lambda *a: False)
    loss_p, per_sample = run()
    assert loss_b == pytest.approx(loss_p, rel=1e-6)
    for name, a, b in zip(("patterns", "potential", "modes", "weights", "positions", "sigma"), batched, per_sample):
        err = rel_l2(a, b)
        print(f"batched vs per-sample {name}: rel-L2 {err:.2e}", flush=True)
        if name == "weights":
            # dL/dw_m = (L_m - L) / sum(w) subtracts two nearly equal float32 sums: compare on the loss's scale
            assert np.abs(a - b).max() < 1e-5 * abs(loss_p) / min(w), name
        elif name in ("positions", "sigma"):
            # sums of phase-ramp derivatives whose terms cancel; the oracle FD tests above hold them to 1e-3
            assert err < 5e-5, name
        else:
            assert err < 1e-5, name
