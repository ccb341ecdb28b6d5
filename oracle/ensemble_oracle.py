"""CPU oracle for the ensemble axes of cbed_image / stem_4d.

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` may import this module, and only
as the checker; the product path never calls into it.

A complex128 numpy restatement (paths relative to the reference root) of:

* ``types/distributions.py:105-215``: distributions with normalised weights;
* ``multislice/producers.py:46-301``: Gauss-Hermite position-jitter and
  coherence distributions (numpy's ``hermgauss``, as the reference calls it);
* ``multislice/reduce.py:29-201``: single- and multi-axis reduction
  (coherent axes summed in amplitude, then |.|^2, then incoherent axes);
* ``multislice/simulations.py:682-1233``: probe-mode axis, the per-sample
  binder (mode selection, jitter shift, tilt ramp, energy-shifted voltage)
  and the ensemble branch of ``cbed_image``;
* ``multislice/simulations.py:1330-1443``: ``stem_4d`` shifting the modes by
  each scan position before the per-position ensemble ``cbed_image``.

Pins (``tests/test_ensemble_oracle.py``): the reference's hand-computed
reducer values (``test_reduce.py``) and its cursor-walk field
(``test_producers.py:95-136``); zero-width axes equal the plain
``cbed_image`` (``test_producers.py:77-93``); Gauss-Hermite moments; an
integer-pixel tilt of a vacuum probe rolls the aperture disc; an energy
offset equals the plain image at the shifted voltage; jitter equals the
explicit weighted sum over shifted probes.
"""

from __future__ import annotations

import numpy as np

from . import ms_oracle as o

COHERENT = "coherent"
INCOHERENT = "incoherent"
DIMS = {"probe_modes": 1, "position_jitter": 2, "coherence": 3}


def distribution(samples, weights, reduction=INCOHERENT, axis_id=None):
    s = np.asarray(samples, np.float64)
    w = np.clip(np.asarray(weights, np.float64), 0.0, None)
    return {"samples": s, "weights": w / w.sum(), "reduction": reduction, "axis_id": axis_id}


def gauss_hermite_normal(n_quad: int):
    x, w = np.polynomial.hermite.hermgauss(int(n_quad))
    return np.sqrt(2.0) * x, w / np.sqrt(np.pi)


def position_jitter(sigma_ang: float, n_quad: int):
    nodes, w = gauss_hermite_normal(n_quad)
    d = sigma_ang * nodes if sigma_ang > 0 else np.zeros_like(nodes)
    dy, dx = np.meshgrid(d, d, indexing="ij")
    wy, wx = np.meshgrid(w, w, indexing="ij")
    return distribution(np.stack((dy.ravel(), dx.ravel()), 1), (wy * wx).ravel(), INCOHERENT, "position_jitter")


def coherence(energy_spread_ev: float, divergence_mrad: float, n_quad: int):
    nodes, w = gauss_hermite_normal(n_quad)
    e = energy_spread_ev * nodes if energy_spread_ev > 0 else np.zeros_like(nodes)
    t = divergence_mrad * nodes if divergence_mrad > 0 else np.zeros_like(nodes)
    eg, xg, yg = np.meshgrid(e, t, t, indexing="ij")
    we, wx, wy = np.meshgrid(w, w, w, indexing="ij")
    return distribution(np.stack((eg.ravel(), xg.ravel(), yg.ravel()), 1), (we * wx * wy).ravel(), INCOHERENT,
                        "coherence")


def probe_modes_axis(weights):
    m = len(weights)
    return distribution(np.arange(m, dtype=np.float64)[:, None], weights, INCOHERENT, "probe_modes")


def apply_distributions(axes, bound):
    counts = [a["samples"].shape[0] for a in axes]
    grids = np.meshgrid(*[np.arange(c) for c in counts], indexing="ij")
    rows = np.concatenate([a["samples"][g.ravel()] for a, g in zip(axes, grids)], axis=1)
    amps = np.stack([bound(r) for r in rows])
    field = amps.reshape(*counts, *amps.shape[-2:])
    labels = list(range(len(axes)))
    for ax in sorted([i for i, a in enumerate(axes) if a["reduction"] == COHERENT], reverse=True):
        cur = labels.index(ax)
        field = np.tensordot(axes[ax]["weights"].astype(np.complex128), np.moveaxis(field, cur, 0), axes=(0, 0))
        labels.pop(cur)
    inten = np.abs(field) ** 2
    for ax in sorted([i for i, a in enumerate(axes) if a["reduction"] == INCOHERENT], reverse=True):
        cur = labels.index(ax)
        inten = np.tensordot(axes[ax]["weights"], np.moveaxis(inten, cur, 0), axes=(0, 0))
        labels.pop(cur)
    return inten


def axis_update(sample, axes):
    """(energy eV, position (2,) Angstrom, tilt (2,) mrad, mode index) of one concatenated sample row."""
    cursor, energy, mode = 0, 0.0, 0
    pos, tilt = np.zeros(2), np.zeros(2)
    for a in axes:
        d = a["samples"].shape[1]
        cols = sample[cursor:cursor + d]
        if a["axis_id"] == "probe_modes":
            mode = int(cols[0])
        elif a["axis_id"] == "position_jitter":
            pos = pos + cols
        elif a["axis_id"] == "coherence":
            energy += cols[0]
            tilt = tilt + cols[1:3]
        cursor += d
    return energy, pos, tilt, mode


def tilt_ramp(n_y, n_x, tilt_mrad, voltage_kv, calib):
    y = (np.arange(n_y) - (n_y - 1) / 2.0) * calib
    x = (np.arange(n_x) - (n_x - 1) / 2.0) * calib
    yy, xx = np.meshgrid(y, x, indexing="ij")
    phase = (2.0 * np.pi / o.wavelength_ang(voltage_kv)) * (tilt_mrad[0] / 1000.0 * xx + tilt_mrad[1] / 1000.0 * yy)
    return np.exp(1j * phase)


def bind(pot, thickness, pot_calib, modes, beam_calib, voltage_kv, axes, det_calib):
    """The per-sample amplitude of ``bind_cbed_axes`` (``simulations.py:769-887``)."""
    modes = np.asarray(modes, np.complex128)

    def bound(sample):
        energy, dpos, tilt, m = axis_update(sample, axes)
        kv = voltage_kv + energy / 1000.0
        shifted = o.shift_beam_fourier(modes[..., m:m + 1], dpos[None, :], det_calib)[0]
        tilted = shifted * tilt_ramp(modes.shape[0], modes.shape[1], tilt, kv, det_calib)[..., None]
        return o.cbed_amplitude(pot, thickness, pot_calib, tilted, beam_calib, kv)[..., 0]

    return bound


def ensemble_axes(mode_weights, jitter=None, coh=None, explicit_modes=None):
    axes = []
    if explicit_modes is not None:
        axes.append(explicit_modes)
    elif len(mode_weights) > 1:
        axes.append(probe_modes_axis(mode_weights))
    if jitter is not None:
        axes.append(jitter)
    if coh is not None:
        axes.append(coh)
    if not axes:
        axes.append(probe_modes_axis(mode_weights))
    return axes


def cbed_image(pot, thickness, pot_calib, modes, weights, beam_calib, voltage_kv, jitter=None, coh=None,
               explicit_modes=None):
    """Ensemble branch of ``cbed_image`` (``simulations.py:1130-1233``)."""
    calib = min(float(pot_calib), float(beam_calib))
    axes = ensemble_axes(np.asarray(weights, np.float64) / np.sum(weights), jitter, coh, explicit_modes)
    return apply_distributions(axes, bind(pot, thickness, pot_calib, modes, beam_calib, voltage_kv, axes, calib))


def stem_4d(pot, thickness, pot_calib, modes, weights, beam_calib, voltage_kv, positions, det_calib, jitter=None,
            coh=None):
    """``stem_4d`` with ensemble axes: per position, the shifted modes through ``cbed_image``."""
    shifted = o.shift_beam_fourier(modes, positions, det_calib)
    return np.stack([cbed_image(pot, thickness, pot_calib, shifted[p], weights, beam_calib, voltage_kv, jitter, coh)
                     for p in range(shifted.shape[0])])


def decompose_beam_to_modes(beam, num_modes: int, seed: int, first_mode_weight: float = 0.6):
    """``simulations.py:1446-1536`` with key = PRNGKey(seed): (modes (H, W, M), weights (M,))."""
    from . import recon_oracle as ro
    k1, k2 = ro.jax_split(ro.jax_key(seed), 2)
    hh, ww = beam.shape
    m = int(num_modes)
    mat = ro.jax_normal_f64(k1, (hh * ww, m)) + 1j * ro.jax_normal_f64(k2, (hh * ww, m))
    q, _ = np.linalg.qr(mat, mode="reduced")
    weights = np.zeros(m)
    weights[0] = first_mode_weight
    if m > 1:
        weights[1:] = (1.0 - first_mode_weight) / (m - 1)
    return (q * np.abs(np.asarray(beam).reshape(-1, 1))).reshape(hh, ww, m), weights
