"""GPU tests of the FP64 normal-equation kernels and the mixed-precision
aberration Levenberg-Marquardt (complex64 JVP columns, FP64 normal
equations and damped solve on the GPU) against the FP64 oracle run."""

import numpy as np
import pytest

from oracle import ms_oracle as o

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_normal_equations_match_fp64():
    from ptyrodactyl_amd import _native
    lib = _native.load()
    rng = np.random.default_rng(0)
    p_count, q, n = 3, 12, 64
    amp = (rng.standard_normal((p_count, q + 1, n, n)) + 1j * rng.standard_normal((p_count, q + 1, n, n)))
    amp = amp.astype(np.complex64)
    data = rng.uniform(0, 2, (p_count, n, n)).astype(np.float32)
    dev = torch.device("cuda", 0)
    ta, td = torch.tensor(amp, device=dev), torch.tensor(data, device=dev)
    ws = int(lib.pty_normal_equations_workspace_bytes(p_count, n, q))
    wbuf = torch.empty(ws // 8 + 1, dtype=torch.float64, device=dev)
    jtj = torch.empty((q, q), dtype=torch.float64, device=dev)
    jtr = torch.empty(q, dtype=torch.float64, device=dev)
    rss = torch.empty(1, dtype=torch.float64, device=dev)
    _native.check(lib.pty_normal_equations(_native.ptr(ta), _native.ptr(td), p_count, n, q, _native.ptr(wbuf), ws,
                                           _native.ptr(jtj), _native.ptr(jtr), _native.ptr(rss),
                                           _native.stream_handle(dev)), "pty_normal_equations")
    a = amp.astype(np.complex128)
    j = 2.0 * np.real(np.conj(a[:, :1]) * a[:, 1:]).transpose(1, 0, 2, 3).reshape(q, -1)
    r = (np.abs(a[:, 0]) ** 2 - data.astype(np.float64)).ravel()
    assert np.allclose(jtj.cpu().numpy(), j @ j.T, rtol=1e-11)
    assert np.allclose(jtr.cpu().numpy(), j @ r, rtol=1e-11)
    assert float(rss.cpu()[0]) == pytest.approx(float(r @ r), rel=1e-11)


def test_damped_solve_matches_numpy():
    from ptyrodactyl_amd import _native
    lib = _native.load()
    rng = np.random.default_rng(1)
    q = 12
    m = rng.standard_normal((q, 40))
    a = m @ m.T
    b = rng.standard_normal(q)
    lam = 0.37
    dev = torch.device("cuda", 0)
    ta, tb = torch.tensor(a, device=dev), torch.tensor(b, device=dev)
    x = torch.empty(q, dtype=torch.float64, device=dev)
    mr = torch.empty(1, dtype=torch.float64, device=dev)
    st = torch.empty(1, dtype=torch.int32, device=dev)
    _native.check(lib.pty_damped_solve(_native.ptr(ta), _native.ptr(tb), lam, q, _native.ptr(x), _native.ptr(mr),
                                       _native.ptr(st), _native.stream_handle(dev)), "pty_damped_solve")
    ref = np.linalg.solve(a + lam * np.eye(q), b)
    assert int(st.cpu()[0]) == 1
    assert np.allclose(x.cpu().numpy(), ref, rtol=1e-11)
    assert float(mr.cpu()[0]) == pytest.approx(float(ref @ b - 0.5 * ref @ a @ ref), rel=1e-11)


def _fit(w, data_shifted, iterations):
    from ptyrodactyl_amd import jacobian
    dev = torch.device("cuda", 0)
    return jacobian.fit_probe_aberrations(
        torch.tensor(w.slab_t, device=dev), torch.tensor(w.origins, device=dev), torch.tensor(w.shifts_px, device=dev),
        torch.tensor(data_shifted, dtype=torch.float32, device=dev), w.coeffs_init, calib_ang=w.calib,
        voltage_kv=w.voltage_kv, aperture_mrad=w.aperture_mrad, slice_thickness=w.slice_thickness,
        iterations=iterations)


def _oracle_data(w):
    prop = o.propagation_func(w.n, w.n, w.slice_thickness, w.voltage_kv, w.calib)
    pk = o.probe_k_from_coeffs(w.n, w.calib, w.voltage_kv, w.aperture_mrad, w.coeffs_true)
    amp = o.window_amplitudes(w.slab_t.astype(np.complex128), pk[None], w.origins, w.shifts_px, prop)[:, 0]
    return np.fft.fftshift(np.abs(amp) ** 2, axes=(-2, -1))


def test_aberration_lm_matches_fp64_oracle_small():
    from ptyrodactyl_amd import synth
    w = synth.config4(scan=6, n=64, object_px=128, num_slices=4)
    data = _oracle_data(w)
    fit = _fit(w, data, iterations=12)
    ref, hist = o.lm_aberration_fit(w.slab_t.astype(np.complex128), w.origins, w.shifts_px, data, w.coeffs_init,
                                    w.calib, w.voltage_kv, w.aperture_mrad, w.slice_thickness, iterations=12)
    got = fit.coefficients.cpu().numpy()
    err_oracle = np.abs(got - ref) / np.abs(w.coeffs_true)
    err_true = np.abs(got - w.coeffs_true) / np.abs(w.coeffs_true)
    print("GPU vs FP64 oracle (relative, per coefficient):", np.array2string(err_oracle, precision=2), flush=True)
    print("GPU vs truth:", np.array2string(err_true, precision=2), flush=True)
    print("residual history GPU", [f"{h:.3e}" for h in fit.history[:6]], "oracle", [f"{h:.3e}" for h in hist[:6]])
    assert np.max(err_oracle) < 1e-3
    assert np.max(err_true) < 1e-3


def test_config4_full_against_fp64_oracle_fixture():
    """Aberration fit workload at full size (24x24 positions, 128^2 windows, 8 slices, 256^2 object,
    30 iterations) against the FP64 oracle run stored in tests/golden/config4_fp64_lm.npz."""
    from pathlib import Path
    from ptyrodactyl_amd import synth
    ref = np.load(Path(__file__).resolve().parent / "golden" / "config4_fp64_lm.npz")
    w = synth.config4()
    assert np.allclose(ref["coeffs_init"], w.coeffs_init) and np.allclose(ref["coeffs_true"], w.coeffs_true)
    data = _oracle_data(w)
    fit = _fit(w, data, iterations=30)
    got = fit.coefficients.cpu().numpy()
    err = np.abs(got - ref["coefficients"]) / np.abs(w.coeffs_true)
    print("config4 GPU vs FP64 oracle, relative per coefficient:", np.array2string(err, precision=2), flush=True)
    print(f"config4 max relative parameter error vs FP64 oracle {err.max():.2e}; residual {fit.residual_norm:.3e}"
          f" after {fit.iterations} iterations", flush=True)
    assert err.max() < 1e-3


def _joint_data(w):
    from oracle import joint_oracle as jo
    prop = o.propagation_func(w.n, w.n, w.slice_thickness, w.voltage_kv, w.calib)
    pk = o.probe_k_from_coeffs(w.n, w.calib, w.voltage_kv, w.aperture_mrad, w.coeffs_true)
    return jo.patterns(w.object_true.astype(np.complex128), w.num_slices, pk, w.origins, w.shifts_px, prop)


def _joint_fit(w, data, iterations, cg_iterations):
    from ptyrodactyl_amd import jacobian
    dev = torch.device("cuda", 0)
    return jacobian.fit_aberrations_and_object(
        torch.tensor(w.object_init, device=dev), w.num_slices, torch.tensor(w.origins, device=dev),
        torch.tensor(w.shifts_px, device=dev), torch.tensor(data, dtype=torch.float32, device=dev), w.coeffs_init,
        calib_ang=w.calib, voltage_kv=w.voltage_kv, aperture_mrad=w.aperture_mrad,
        slice_thickness=w.slice_thickness, iterations=iterations, cg_iterations=cg_iterations)


def _joint_errors(fit, coeffs_ref, obj_ref, w):
    got_c = fit.coefficients.cpu().numpy()
    err_c = np.abs(got_c - coeffs_ref) / np.abs(w.coeffs_true)
    obj = fit.object.cpu().numpy()
    start = w.object_init.astype(np.complex128)
    err_o = np.linalg.norm(obj - obj_ref) / np.linalg.norm(obj_ref - start)
    return err_c, err_o


def test_joint_fit_matches_fp64_oracle_small():
    """Joint aberration + object LM on a reduced workload (6 x 6 positions, 64^2 windows, 4 slices, 128^2
    object, 6 iterations) against the FP64 oracle run of the same algorithm."""
    from oracle import joint_oracle as jo
    from ptyrodactyl_amd import synth
    w = synth.config4(scan=6, n=64, object_px=128, num_slices=4)
    data = _joint_data(w)
    fit = _joint_fit(w, data, 6, 3)
    ref_c, ref_o, hist = jo.joint_fit(w.object_init.astype(np.complex128), w.num_slices, w.origins, w.shifts_px,
                                      data, w.coeffs_init, w.calib, w.voltage_kv, w.aperture_mrad,
                                      w.slice_thickness, iterations=6, cg_iterations=3)
    err_c, err_o = _joint_errors(fit, ref_c, ref_o, w)
    print("joint GPU vs FP64 oracle, coefficients:", np.array2string(err_c, precision=2), flush=True)
    print(f"joint object change rel-L2 vs oracle {err_o:.2e}; residual GPU "
          f"{[f'{h[1]:.4e}' for h in fit.history]} oracle {[f'{h[1]:.4e}' for h in hist]}", flush=True)
    assert np.allclose(np.array(fit.history), np.array(hist), rtol=1e-4)
    assert err_c.max() < 1e-4 and err_o < 1e-4


def test_config4_joint_full_against_fp64_oracle_fixture():
    """Config 4 with both unknowns: 12 aberrations (seeded +-20 % offsets) and the 256^2 object slice (seeded
    +-5 % amplitude noise), 24 x 24 positions, 128^2 windows, 8 slices, 30 block-coordinate LM iterations with
    3 CG steps per object step, against the FP64 oracle run stored in tests/golden/config4_joint_fp64.npz
    (generator tests/golden/make_config4_joint_reference.py)."""
    from pathlib import Path
    from ptyrodactyl_amd import synth
    ref = np.load(Path(__file__).resolve().parent / "golden" / "config4_joint_fp64.npz")
    w = synth.config4()
    assert np.allclose(ref["coeffs_init"], w.coeffs_init) and np.allclose(ref["coeffs_true"], w.coeffs_true)
    data = _joint_data(w)
    fit = _joint_fit(w, data, int(ref["iterations"]), int(ref["cg_iterations"]))
    err_c, err_o = _joint_errors(fit, ref["coefficients"], ref["object"].astype(np.complex128), w)
    hist = np.array(fit.history)
    print("config4 joint: GPU vs FP64 oracle, relative per coefficient:", np.array2string(err_c, precision=2),
          flush=True)
    print(f"config4 joint: max coefficient error {err_c.max():.2e}; object change rel-L2 {err_o:.2e}; residual GPU "
          f"{hist[-1, 1]:.6e} oracle {ref['history'][-1, 1]:.6e}", flush=True)
    assert np.allclose(hist, ref["history"], rtol=1e-5)
    assert err_c.max() < 1e-5 and err_o < 1e-5
