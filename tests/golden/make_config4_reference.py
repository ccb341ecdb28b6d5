"""Writes tests/golden/config4_fp64_lm.npz: the FP64 oracle Levenberg-Marquardt run of
the aberration-fit workload (24x24 positions, 128^2 windows, 8 slices, 256^2 object,
30 iterations) that the GPU mixed-precision fit is compared with.

Run from the repository root:  python tests/golden/make_config4_reference.py
"""

import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import ms_oracle as o  # noqa: E402
from ptyrodactyl_amd import synth  # noqa: E402


def main():
    t0 = time.time()
    w = synth.config4()
    slab = w.slab_t.astype(np.complex128)
    prop = o.propagation_func(w.n, w.n, w.slice_thickness, w.voltage_kv, w.calib)
    pk = o.probe_k_from_coeffs(w.n, w.calib, w.voltage_kv, w.aperture_mrad, w.coeffs_true)
    amp = o.window_amplitudes(slab, pk[None], w.origins, w.shifts_px, prop)[:, 0]
    data = np.fft.fftshift(np.abs(amp) ** 2, axes=(-2, -1))
    coeffs, hist = o.lm_aberration_fit(slab, w.origins, w.shifts_px, data, w.coeffs_init, w.calib, w.voltage_kv,
                                       w.aperture_mrad, w.slice_thickness, iterations=30)
    out = Path(__file__).resolve().parent / "config4_fp64_lm.npz"
    np.savez(out, coefficients=coeffs, history=np.asarray(hist), coeffs_true=w.coeffs_true,
             coeffs_init=w.coeffs_init)
    print(f"wrote {out} in {time.time() - t0:.0f}s; final residual {hist[-1]:.3e}")


if __name__ == "__main__":
    main()
