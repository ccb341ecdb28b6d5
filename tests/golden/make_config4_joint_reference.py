"""Writes tests/golden/config4_joint_fp64.npz: the FP64 oracle run of the joint aberration + object
reconstruction of the config-4 workload (24x24 positions, 128^2 windows, 8 slices repeating one 256^2
object slice, 12 aberration coefficients from seeded +-20 % offsets, the object from seeded +-5 %
amplitude noise; 30 block-coordinate LM iterations, 3 CG steps per object step) that the GPU
mixed-precision fit is compared with.  Noise-free data from the true coefficients and object.

Run from the repository root:  python tests/golden/make_config4_joint_reference.py
"""

import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import joint_oracle as jo  # noqa: E402
from oracle import ms_oracle as o  # noqa: E402
from ptyrodactyl_amd import synth  # noqa: E402

ITERATIONS = 30
CG_ITERATIONS = 3


def main():
    t0 = time.time()
    w = synth.config4()
    prop = o.propagation_func(w.n, w.n, w.slice_thickness, w.voltage_kv, w.calib)
    pk = o.probe_k_from_coeffs(w.n, w.calib, w.voltage_kv, w.aperture_mrad, w.coeffs_true)
    data = jo.patterns(w.object_true.astype(np.complex128), w.num_slices, pk, w.origins, w.shifts_px, prop)
    coeffs, obj, hist = jo.joint_fit(w.object_init.astype(np.complex128), w.num_slices, w.origins, w.shifts_px,
                                     data, w.coeffs_init, w.calib, w.voltage_kv, w.aperture_mrad, w.slice_thickness,
                                     iterations=ITERATIONS, cg_iterations=CG_ITERATIONS,
                                     log=lambda m: print(f"{time.time() - t0:7.1f}s {m}", flush=True))
    out = Path(__file__).resolve().parent / "config4_joint_fp64.npz"
    np.savez(out, coefficients=coeffs, object=obj.astype(np.complex64), history=np.asarray(hist),
             coeffs_init=w.coeffs_init, coeffs_true=w.coeffs_true, iterations=ITERATIONS,
             cg_iterations=CG_ITERATIONS)
    print(f"wrote {out} in {time.time() - t0:.0f}s; final residual {hist[-1][1]:.6e}")


if __name__ == "__main__":
    main()
