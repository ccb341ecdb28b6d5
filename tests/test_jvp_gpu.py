"""GPU parity of the forward-mode multislice (``pty_jvp``) and of forward-mode autodiff through it.

* d_patterns against the complex128 oracle's tangent recurrence (``multislice_jvp``, itself pinned by
  central differences and the dot-product identity on the CPU) at relative L2 <= 1e-5, for slab and
  probe tangents, one or two modes, n = 32 ... 512;
* the primal patterns it returns equal ``pty_forward``'s to 1e-6 (the same transforms and tables, in kernels
  the compiler schedules differently);
* the dot-product identity <J v, u> = <v, J^T u> against the HIP adjoint for slab, probe and shift
  tangents, for intensities and (``pty_adjoint_amplitude``) amplitudes, within 1e-5 relative;
* torch.func.jvp through the drop-in ``stem_4d`` (potential and probe tangents) against central
  differences of the oracle's reference-level stem_4d.
"""

import numpy as np
import pytest

from oracle import ms_oracle as o

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

DEV = "cuda"
TOL = 1e-5


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a.astype(np.result_type(a, np.float64)) - b) / np.linalg.norm(b))


def _case(n, s_count, p_count, slab_hw, seed, modes=1):
    rng = np.random.default_rng(seed)
    slab = (rng.uniform(0.95, 1.0, (s_count,) + slab_hw) * np.exp(1j * rng.uniform(0, 0.4, (s_count,) + slab_hw)))
    slab = slab.astype(np.complex64)
    probes = np.stack([o.make_probe(300.0, 20.0 + 4 * m, 25.0 * (m + 1), 0.0, 0.0, (n, n), 10.0)
                       for m in range(modes)])
    origins = np.stack([rng.integers(0, slab_hw[0], p_count), rng.integers(0, slab_hw[1], p_count)], 1)
    shifts = rng.uniform(-2.5, 2.5, (p_count, 2))
    d_slab = ((rng.standard_normal((s_count,) + slab_hw) + 1j * rng.standard_normal((s_count,) + slab_hw)) * 0.1
              ).astype(np.complex64)
    d_probes = (rng.standard_normal(probes.shape) + 1j * rng.standard_normal(probes.shape)) * np.abs(probes).max()
    return slab, probes, origins.astype(np.int32), shifts, d_slab, d_probes


def _operands(slab, probes, origins, shifts, n, dz=1.5, kv=300.0, calib=0.1, weights=None):
    from ptyrodactyl_amd import physics
    m_count = probes.shape[0]
    w = np.full(m_count, 1.0 / m_count) if weights is None else np.asarray(weights)
    return (torch.tensor(slab, device=DEV),
            torch.fft.fft2(torch.tensor(probes, device=DEV)).to(torch.complex64).contiguous(),
            torch.tensor(w, dtype=torch.float32, device=DEV), torch.tensor(origins, device=DEV),
            torch.tensor(shifts, device=DEV), physics.propagator_factors(n, dz, kv, calib, device=DEV)), w


@pytest.mark.parametrize("n,s_count,p_count,slab_hw,modes", [
    (32, 3, 5, (40, 36), 1), (64, 4, 4, (70, 64), 2), (128, 6, 3, (140, 128), 1), (128, 16, 2, (160, 150), 1),
    (256, 4, 3, (280, 256), 1), (512, 2, 2, (512, 530), 1),
])
def test_jvp_matches_oracle_tangent_recurrence(n, s_count, p_count, slab_hw, modes):
    from ptyrodactyl_amd import engine
    print(f"jvp n={n} S={s_count} P={p_count} M={modes}", flush=True)
    slab, probes, origins, shifts, d_slab, d_probes = _case(n, s_count, p_count, slab_hw, seed=n + s_count, modes=modes)
    ops, w = _operands(slab, probes, origins, shifts, n)
    d_pk = torch.fft.fft2(torch.tensor(d_probes, device=DEV)).to(torch.complex64).contiguous()
    pat, _, d_pat, _ = engine.multislice_jvp(*ops, torch.tensor(d_slab, device=DEV), d_pk, want_patterns=True)
    ref_pat, _ = engine.multislice_forward(*ops)
    torch.cuda.synchronize()
    e_pat = rel_l2(pat.cpu().numpy(), ref_pat.cpu().numpy())
    print(f"  primal patterns vs pty_forward: rel-L2 {e_pat:.2e}", flush=True)
    assert e_pat < 1e-6
    prop = o.propagation_func(n, n, 1.5, 300.0, 0.1)
    want = 0
    for m in range(modes):
        want = want + w[m] * o.multislice_jvp(slab.astype(np.complex128), probes[m], origins, shifts, prop,
                                              d_slab.astype(np.complex128), d_probes[m])
    err = rel_l2(d_pat.cpu().numpy(), want)
    print(f"  d_patterns rel-L2 {err:.2e}", flush=True)
    assert err < TOL


@pytest.mark.parametrize("n", [64, 128, 256])
def test_jvp_adjoint_dot_product_identity(n):
    """<J v, u> = <v, J^T u> with J^T from the HIP adjoint, for intensity and amplitude outputs; the tangent
    v covers the slab, probe_k and the shifts."""
    from ptyrodactyl_amd import engine
    slab, probes, origins, shifts, d_slab, d_probes = _case(n, 5, 4, (n + 20, n + 10), seed=3 * n, modes=2)
    ops, _ = _operands(slab, probes, origins, shifts, n)
    rng = np.random.default_rng(n)
    d_slab_t = torch.tensor(d_slab, device=DEV)
    d_pk = torch.fft.fft2(torch.tensor(d_probes, device=DEV)).to(torch.complex64).contiguous()
    d_sh = torch.tensor(rng.standard_normal((4, 2)), device=DEV)
    u_i = torch.tensor(rng.standard_normal((4, n, n)), dtype=torch.float32, device=DEV)
    u_a = torch.tensor(rng.standard_normal((4, 2, n, n)) + 1j * rng.standard_normal((4, 2, n, n)),
                       dtype=torch.complex64, device=DEV)
    _, _, d_pat, d_amp = engine.multislice_jvp(*ops, d_slab_t, d_pk, d_sh, want_d_amplitude=True)
    k = 2
    _, ckpt = engine.multislice_forward_checkpointed(*ops, k)

    def inner(a, b):
        return float(torch.sum(torch.conj(a.to(torch.complex128)) * b.to(torch.complex128)).real)

    for seed_name, seed in (("intensity", u_i), ("amplitude", u_a)):
        g_sh = torch.zeros((4, 2), dtype=torch.float64, device=DEV)
        if seed_name == "intensity":
            g_slab, g_pk = engine.multislice_adjoint(*ops, seed, ckpt, k, grad_shifts=g_sh)
            lhs = float(torch.sum(d_pat.double() * seed.double()))
        else:
            g_slab, g_pk = engine.multislice_adjoint(*ops, None, ckpt, k, grad_amplitude=seed, grad_shifts=g_sh)
            lhs = inner(seed, d_amp)
        rhs = inner(g_slab, d_slab_t) + inner(g_pk, d_pk) + float(torch.sum(g_sh * d_sh))
        print(f"jvp/adjoint identity n={n} {seed_name}: <Jv,u> {lhs:.8e}  <v,J^T u> {rhs:.8e}", flush=True)
        assert lhs == pytest.approx(rhs, rel=TOL)


def test_func_jvp_through_drop_in_stem_4d_matches_oracle_differences():
    """torch.func.jvp of the drop-in stem_4d w.r.t. the potential slices and the probe modes (the GN / LM
    solvers' forward-mode path) against central differences (h = 1e-4) of the oracle's stem_4d."""
    import ptyrodactyl_amd as pa
    rng = np.random.default_rng(71)
    n, s_count, calib, kv, dz = 64, 3, 0.2, 200.0, 1.5
    pot = rng.uniform(0, 20, (n, n, s_count))
    d_pot = rng.standard_normal((n, n, s_count))
    mic = pa.create_microscope_config(voltage_kv=kv, aperture_mrad=20.0, probe_shape=(n, n))
    pos = torch.tensor(rng.uniform(0, 5, (4, 2)))
    det = pa.create_detector_config(real_space_calib_ang=calib, probe_calibration_pm=100.0 * calib,
                                    scan_positions_px=pos)
    probe = pa.make_probe(mic, det).to(torch.complex128)
    d_probe = torch.tensor(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * probe.abs().max()

    def f(v, pr):
        pots = pa.create_potential_slices(v, dz, calib)
        beam = pa.create_probe_modes(pr[..., None], torch.tensor([1.0]), calib)
        return pa.stem_4d(pots, beam, mic, det).data

    v0 = torch.tensor(pot, device=DEV)
    _, d_out = torch.func.jvp(f, (v0, probe.to(DEV)), (torch.tensor(d_pot, device=DEV), d_probe.to(DEV)))
    d_out = d_out.double().cpu().numpy()

    def oracle(pv, pr):
        return o.stem_4d(pv, dz, calib, pr[..., None], np.array([1.0]), calib, kv, pos.numpy(), calib)["data"]

    h = 1e-4
    probe_np = probe.numpy()
    fd = (oracle(pot + h * d_pot, probe_np + h * d_probe.numpy())
          - oracle(pot - h * d_pot, probe_np - h * d_probe.numpy())) / (2 * h)
    err = rel_l2(d_out, fd)
    print(f"torch.func.jvp through stem_4d vs oracle central differences: rel-L2 {err:.2e}", flush=True)
    assert err < 2e-5
