"""Ptychographic reconstructors: Optax loops over the HIP 4D-STEM gradient.

Same names, parameters, defaults and return values as
``src/ptyrodactyl/multislice/multislice_recon.py``:

* :func:`single_slice_ptychography` (``:242-667``) — potential and beam;
* :func:`single_slice_poscorrected` (``:670-1220``) — plus scan positions;
* :func:`single_slice_multi_modal` (``:1223-1736``) — ``ProbeModes`` beam;
* :func:`multi_slice_multi_modal` (``:1739-2227``) — separate position
  learning rate.

Each iteration evaluates the loss of :func:`stem_4d` (looked up on this
module, so it can be replaced as the reference's tests do) and its gradient
with ``torch.autograd``: the potential, probe modes and positions reach the
HIP forward through torch and the HIP adjoint returns the gradients w.r.t.
the transmission, probe and shifts.  torch's gradient dL/dRe + i dL/dIm is
the ``conj(jax.grad)`` the reference hands to Optax
(``_apply_optimizer_step`` ``:166-210``); the update rules run in place in
the fused kernel of :mod:`.optim`.  Snapshots follow the reference: slot 0
after the first update, then slot ``ii // save_every`` after iteration
``ii`` whenever ``ii % save_every == 0`` (slots past the end are dropped).

Precision: parameters and optimizer moments keep the promoted parameter
dtype (complex128 for float64 inputs); the multislice itself runs in
complex64 arithmetic on the GPU.

With ``PTYRODACTYL_CAPTURE_ITERATIONS=1`` the iterations after the first
replay a HIP graph of the forward and the gradients (MSE and MAE losses),
with the same kernels on the same data as the eager loop.
"""

from __future__ import annotations

import os
from typing import Callable, Dict, List, Tuple, Union

import torch

from .optim import OPTIMIZERS, OptaxRule, OptaxState, get_optimizer
from .simulations import stem_4d
from .types import (STEM4D, CalibratedArray, LossType, ProbeModes, create_calibrated_array, create_detector_config,
                    create_microscope_config, create_potential_slices, create_probe_modes, error_if,
                    suspended_checks)

__all__ = ["OPTIMIZERS", "multi_slice_multi_modal", "single_slice_multi_modal", "single_slice_poscorrected",
           "single_slice_ptychography"]


def _reduce_loss(model_output: torch.Tensor, experimental_data: torch.Tensor, loss_mode: LossType) -> torch.Tensor:
    """Mean absolute, mean squared or root-mean-squared error (``multislice_recon.py:127-163``).

    The absolute value is written so that its derivative at a zero residual is +1, as JAX's is
    (torch's ``abs`` has derivative 0 there)."""
    r = model_output - experimental_data
    if loss_mode is LossType.MAE:
        return torch.mean(torch.where(r >= 0, r, -r))
    if loss_mode is LossType.MSE:
        return torch.mean(r * r)
    return torch.sqrt(torch.mean(r * r))


def _promote_to_complex(values) -> torch.Tensor:
    """float32 / complex64 -> complex64, anything wider -> complex128 (``multislice_recon.py:213-239``)."""
    t = values if isinstance(values, torch.Tensor) else torch.as_tensor(values)
    if t.dtype in (torch.float32, torch.complex64, torch.float16, torch.bfloat16):
        return t.to(torch.complex64)
    return t.to(torch.complex128)


def _apply_optimizer_step(params: torch.Tensor, grads: torch.Tensor, optimizer: OptaxRule,
                          optimizer_state: OptaxState) -> None:
    """One in-place update with torch's gradient (= the conjugated JAX gradient Optax receives)."""
    optimizer.update_(params, grads, optimizer_state)


def _owned(t: torch.Tensor, dev) -> torch.Tensor:
    """A contiguous copy on ``dev``: the update rules write in place and the caller's arrays stay unchanged,
    as the reference's functional loop leaves them."""
    return t.detach().to(device=dev, copy=True).contiguous()


def _learning_rates(learning_rate) -> Tuple[float, float]:
    """A scalar applies to parameters and positions; a length-2 array is (parameters, positions)."""
    lr = torch.as_tensor(learning_rate, dtype=torch.float64)
    if lr.ndim == 0:
        return float(lr), float(lr)
    return float(lr[0]), float(lr[1])


def _device(experimental_data: STEM4D) -> torch.device:
    d = experimental_data.data
    if isinstance(d, torch.Tensor) and d.is_cuda:
        return d.device
    if not torch.cuda.is_available():
        raise RuntimeError("the reconstructors run on the GPU (no CPU fallback) and no GPU is visible")
    return torch.device("cuda", torch.cuda.current_device())


def _capture_enabled() -> bool:
    return os.environ.get("PTYRODACTYL_CAPTURE_ITERATIONS", "0").lower() in ("1", "true", "yes")


def _loss_seed(model_output: torch.Tensor, experimental_data: torch.Tensor, loss_mode: LossType):
    """dL/d(model output) of the MSE or MAE loss, elementwise in autograd's operation order, or None (RMSE needs
    the reduced loss in its derivative)."""
    r = model_output - experimental_data
    inv_n = 1.0 / r.numel()
    if loss_mode is LossType.MSE:
        t = r * inv_n
        return (t + t).to(model_output.dtype)
    if loss_mode is LossType.MAE:
        return torch.where(r >= 0, torch.full_like(r, inv_n), torch.full_like(r, -inv_n)).to(model_output.dtype)
    return None


class _CapturedGradient:
    """The forward and the parameter gradients of one iteration recorded once in a HIP graph and replayed.

    The parameters are read through views of their storage, which the eager Optax update changes in place
    between replays. The graph is seeded with the elementwise loss derivative, so it holds no reductions. The
    reported loss is reduced eagerly from the graph's static output. Value checks are off while recording, as
    a graph may not synchronise."""

    def __init__(self, forward, params, names, data, loss_mode):
        self.names, self.data, self.loss_mode = names, data, loss_mode
        with suspended_checks():
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                for _ in range(2):  # warm-up on a side stream before recording; the parameters are not updated
                    warm = {k: params[k].detach().requires_grad_(True) for k in names}
                    out = forward(**warm)
                    torch.autograd.grad(out, [warm[k] for k in names], grad_outputs=_loss_seed(out, data, loss_mode),
                                        allow_unused=True)
                    del warm, out
            torch.cuda.current_stream().wait_stream(side)
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                # leaves made on the capture stream: a leaf's gradient accumulator belongs to the stream that
                # first used it, and warm-up leaves reused here made autograd warn of (and sync across) the
                # stream mismatch between the side stream and the capture stream
                self.leaves = {k: params[k].detach().requires_grad_(True) for k in names}
                self.out = forward(**self.leaves)
                self.grads = torch.autograd.grad(self.out, [self.leaves[k] for k in names],
                                                 grad_outputs=_loss_seed(self.out, data, loss_mode), allow_unused=True)

    def run(self):
        self.graph.replay()
        return self.grads, _reduce_loss(self.out, self.data, self.loss_mode).detach()


def _optimise(forward: Callable[..., torch.Tensor], params: Dict[str, torch.Tensor],
              optimizers: Dict[str, OptaxRule], data: torch.Tensor, loss_mode: LossType, save_every: int,
              num_iterations: int) -> Tuple[Dict[str, torch.Tensor], Dict[str, torch.Tensor], List[float]]:
    """The loop shared by the four reconstructors; ``params`` are updated in place."""
    if save_every < 1:
        raise ValueError("save_every must be >= 1")
    states = {k: optimizers[k].init(p) for k, p in params.items()}
    count = num_iterations // save_every
    snaps = {k: torch.zeros(tuple(p.shape) + (count,), dtype=p.dtype, device=p.device) for k, p in params.items()}
    losses: List[float] = []
    names = list(params)

    def step() -> None:
        leaves = {k: params[k].detach().requires_grad_(True) for k in names}
        loss = _reduce_loss(forward(**leaves), data, loss_mode)
        grads = torch.autograd.grad(loss, [leaves[k] for k in names], allow_unused=True)
        for k, g in zip(names, grads):
            _apply_optimizer_step(params[k], torch.zeros_like(params[k]) if g is None else g, optimizers[k],
                                  states[k])
        losses.append(loss.detach())

    def save(slot: int) -> None:
        if 0 <= slot < count:
            for k in names:
                snaps[k][..., slot] = params[k]

    captured = None
    if num_iterations > 0:
        step()
        save(0)
        if num_iterations > 1 and _capture_enabled() and loss_mode is not LossType.RMSE and \
                next(iter(params.values())).is_cuda:
            captured = _CapturedGradient(forward, params, names, data, loss_mode)
        for ii in range(1, num_iterations):
            if captured is None:
                step()
            else:
                grads, loss = captured.run()
                for k, g in zip(names, grads):
                    _apply_optimizer_step(params[k], torch.zeros_like(params[k]) if g is None else g,
                                          optimizers[k], states[k])
                losses.append(loss)
            if ii % save_every == 0:
                save(ii // save_every)
    if captured is not None:  # the recorded iterations ran without value checks
        for k in names:
            error_if(~torch.isfinite(torch.view_as_real(params[k]) if params[k].is_complex() else params[k]).all(),
                     f"{k} contains non-finite values")
    return params, snaps, [float(v) for v in losses]


def _setup(experimental_data: STEM4D, loss_mode: LossType):
    dev = _device(experimental_data)
    data = experimental_data.data.to(dev)
    voltage_kv = float(experimental_data.voltage_kv)
    calib_ang = float(experimental_data.real_space_calib)
    microscope = create_microscope_config(voltage_kv=voltage_kv, aperture_mrad=1.0)
    positions = experimental_data.scan_positions.to(device=dev, dtype=torch.float64).contiguous()
    return loss_mode, dev, data, calib_ang, microscope, positions


def _single_mode_forward(slice_thickness, calib_ang, microscope, data_dtype, fixed_positions=None):
    def forward(pot_slice, beam, pos_list=None):
        pos = fixed_positions if pos_list is None else pos_list
        potential_slices = create_potential_slices(slices=torch.real(pot_slice)[..., None],
                                                   slice_thickness=slice_thickness, calib=calib_ang)
        probe_modes = create_probe_modes(modes=beam[..., None],
                                         weights=torch.ones((1,), dtype=data_dtype, device=beam.device),
                                         calib=calib_ang)
        detector = create_detector_config(real_space_calib_ang=calib_ang, scan_positions_px=pos)
        return stem_4d(potential_slices, probe_modes, microscope, detector).data
    return forward


def single_slice_ptychography(experimental_data: STEM4D, initial_potential: CalibratedArray,
                              initial_beam: CalibratedArray, slice_thickness, save_every: int = 10,
                              num_iterations: int = 1000, learning_rate: float = 0.001, loss_type: str = "mse",
                              optimizer_name: str = "adam"):
    """Reconstruct the potential slice and the beam (``multislice_recon.py:242-667``).

    A reciprocal-space ``initial_beam`` (``real_space=False``) is inverse-transformed first.  Returns
    (final_potential, final_beam) as real-space ``CalibratedArray`` with the initial calibrations, and
    the potential and beam snapshots (H, W, num_iterations // save_every).
    """
    loss_mode = LossType(loss_type)
    optimizer = get_optimizer(optimizer_name, learning_rate)
    loss_mode, dev, data, calib_ang, microscope, positions = _setup(experimental_data, loss_mode)
    pot = _owned(_promote_to_complex(initial_potential.data_array), dev)
    beam = _owned(_promote_to_complex(initial_beam.data_array), dev)
    if not initial_beam.real_space:
        beam = torch.fft.ifft2(beam).contiguous()
    forward = _single_mode_forward(slice_thickness, calib_ang, microscope, data.dtype, fixed_positions=positions)
    params, snaps, _ = _optimise(lambda pot_slice, beam: forward(pot_slice, beam), {"pot_slice": pot, "beam": beam},
                                 {"pot_slice": optimizer, "beam": optimizer}, data, loss_mode, save_every,
                                 num_iterations)
    final_potential = create_calibrated_array(params["pot_slice"], initial_potential.calib_y,
                                              initial_potential.calib_x, True)
    final_beam = create_calibrated_array(params["beam"], initial_beam.calib_y, initial_beam.calib_x, True)
    return final_potential, final_beam, snaps["pot_slice"], snaps["beam"]


def single_slice_poscorrected(experimental_data: STEM4D, initial_potential: CalibratedArray,
                              initial_beam: CalibratedArray, slice_thickness, save_every: int = 10,
                              num_iterations: int = 1000, learning_rate: Union[float, torch.Tensor] = 0.01,
                              loss_type: str = "mse", optimizer_name: str = "adam"):
    """Reconstruct potential, beam and scan positions (``multislice_recon.py:670-1220``).

    ``learning_rate`` is a scalar or (parameter_lr, position_lr).  The beam is used as given (no
    reciprocal-space branch, as in the reference).  Returns (final_potential, final_beam, positions
    (P, 2), potential snapshots, beam snapshots, position snapshots (P, 2, S)).
    """
    loss_mode = LossType(loss_type)
    param_lr, pos_lr = _learning_rates(learning_rate)
    param_opt, pos_opt = get_optimizer(optimizer_name, param_lr), get_optimizer(optimizer_name, pos_lr)
    loss_mode, dev, data, calib_ang, microscope, positions = _setup(experimental_data, loss_mode)
    pot = _owned(_promote_to_complex(initial_potential.data_array), dev)
    beam = _owned(_promote_to_complex(initial_beam.data_array), dev)
    forward = _single_mode_forward(slice_thickness, calib_ang, microscope, data.dtype)
    params, snaps, _ = _optimise(forward, {"pot_slice": pot, "beam": beam, "pos_list": positions.clone()},
                                 {"pot_slice": param_opt, "beam": param_opt, "pos_list": pos_opt}, data, loss_mode,
                                 save_every, num_iterations)
    final_potential = create_calibrated_array(params["pot_slice"], initial_potential.calib_y,
                                              initial_potential.calib_x, True)
    final_beam = create_calibrated_array(params["beam"], initial_beam.calib_y, initial_beam.calib_x, True)
    return (final_potential, final_beam, params["pos_list"], snaps["pot_slice"], snaps["beam"],
            snaps["pos_list"])


def single_slice_multi_modal(experimental_data: STEM4D, initial_pot_slice: torch.Tensor, initial_beam: ProbeModes,
                             slice_thickness, save_every: int = 10, num_iterations: int = 1000,
                             learning_rate: Union[float, torch.Tensor] = 0.01, loss_type: str = "mse",
                             optimizer_name: str = "adam"):
    """Reconstruct potential, probe modes and positions (``multislice_recon.py:1223-1736``).

    Only the modes of ``initial_beam`` are optimised; its weights and calibration are kept.  Returns
    (pot_slice (H, W), ProbeModes, positions (P, 2), potential snapshots (H, W, S), mode snapshots
    (H, W, M, S)).
    """
    loss_mode = LossType(loss_type)
    param_lr, pos_lr = _learning_rates(learning_rate)
    param_opt, pos_opt = get_optimizer(optimizer_name, param_lr), get_optimizer(optimizer_name, pos_lr)
    loss_mode, dev, data, calib_ang, microscope, positions = _setup(experimental_data, loss_mode)
    pot = _owned(_promote_to_complex(initial_pot_slice), dev)
    modes = _owned(_promote_to_complex(initial_beam.modes), dev)
    weights, beam_calib = initial_beam.weights, initial_beam.calib

    def forward(pot_slice, modes, pos_list):
        potential_slices = create_potential_slices(slices=torch.real(pot_slice)[..., None],
                                                   slice_thickness=slice_thickness, calib=calib_ang)
        detector = create_detector_config(real_space_calib_ang=calib_ang, scan_positions_px=pos_list)
        return stem_4d(potential_slices, ProbeModes(modes, weights, beam_calib), microscope, detector).data

    params, snaps, _ = _optimise(forward, {"pot_slice": pot, "modes": modes, "pos_list": positions.clone()},
                                 {"pot_slice": param_opt, "modes": param_opt, "pos_list": pos_opt}, data, loss_mode,
                                 save_every, num_iterations)
    beam = ProbeModes(params["modes"], weights, beam_calib)
    return params["pot_slice"], beam, params["pos_list"], snaps["pot_slice"], snaps["modes"]


def multi_slice_multi_modal(experimental_data: STEM4D, initial_pot_slice: torch.Tensor,
                            initial_beam: torch.Tensor, slice_thickness, save_every: int = 10,
                            num_iterations: int = 1000, learning_rate: float = 0.001,
                            pos_learning_rate: float = 0.01, loss_type: str = "mse", optimizer_name: str = "adam"):
    """Reconstruct potential, beam and positions with separate learning rates
    (``multislice_recon.py:1739-2227``).  Returns (pot_slice (H, W), beam (H, W), positions (P, 2),
    potential snapshots, beam snapshots)."""
    loss_mode = LossType(loss_type)
    param_opt = get_optimizer(optimizer_name, learning_rate)
    pos_opt = get_optimizer(optimizer_name, pos_learning_rate)
    loss_mode, dev, data, calib_ang, microscope, positions = _setup(experimental_data, loss_mode)
    pot = _owned(_promote_to_complex(initial_pot_slice), dev)
    beam = _owned(_promote_to_complex(initial_beam), dev)
    forward = _single_mode_forward(slice_thickness, calib_ang, microscope, data.dtype)
    params, snaps, _ = _optimise(forward, {"pot_slice": pot, "beam": beam, "pos_list": positions.clone()},
                                 {"pot_slice": param_opt, "beam": param_opt, "pos_list": pos_opt}, data, loss_mode,
                                 save_every, num_iterations)
    return params["pot_slice"], params["beam"], params["pos_list"], snaps["pot_slice"], snaps["beam"]
