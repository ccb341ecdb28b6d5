"""Continuous-batching request loop for mixed 4D-STEM simulation and gradient requests.

A request is a forward 4D-STEM simulation or a gradient evaluation
(loss sum (sqrt(I + eps) - sqrt(I_meas))^2 and its gradient w.r.t. the
transmission slab and probe) with its own window n, slice count S, scan
size P and specimen.  :class:`RequestLoop` serves a stream of such requests:

* admission by HBM budget: a request is admitted when it has arrived and
  its device footprint (inputs, outputs, checkpoints) fits the budget; its
  inputs are then built on the device and stay resident until it completes;
* continuous batching: every step takes chunks of positions from all
  admitted requests of each size class (n, S), in arrival order, and runs
  them in ONE batched launch sequence (``pty_forward_batched`` /
  ``pty_adjoint_batched``): per-position descriptors point each wave at its
  own request's slab, probe, propagator and outputs, so requests that
  arrive while others run join the next step;
* size classes run on their own HIP streams so their work overlaps;
* per-request latency = completion - arrival (queueing included), reported
  as p50 / p99 beside patterns/s;
* failure isolation: a request that cannot be built, or whose work raises
  when launched, is marked ``failed`` with the error and its memory is
  released while the other requests continue (a failing batched step is
  retried request by request, so only the offending request fails); a
  request larger than the whole budget is ``rejected`` at admission;
  ``cancel(rid)`` drops a pending or running request at the next step;
* per-step metrics (time, active and pending requests, HBM in use) are kept
  in ``metrics`` and can be streamed to a callback as JSON-ready dicts.

:class:`PlainExecutor` is the comparison: the same requests one at a time,
in arrival order, each through the ordinary per-request engine calls.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _native, autograd, engine, physics, synth

EPS_SQRT = 1e-6                                        # loss regulariser: sqrt(I + eps)
CALIB = {32: 0.25, 64: 0.2, 128: 0.15, 256: 0.1, 512: 0.1}
SLICE_DZ = {32: 1.95, 64: 1.95, 128: 1.95, 256: 1.0, 512: 1.0}
VOLTAGE_KV, APERTURE_MRAD, DEFOCUS, C3 = 300.0, 25.0, 50.0, 1e-3
CKPT_EVERY = 4
DESC_DTYPE = np.dtype(_native.POSITION_DESC_FIELDS)


@dataclass
class Request:
    rid: int
    kind: str          # "forward" or "gradient"
    n: int
    slices: int
    positions: int
    seed: int
    arrival: float = 0.0
    # results
    patterns: Optional[torch.Tensor] = None
    loss: Optional[float] = None
    grad_slab: Optional[torch.Tensor] = None
    grad_probe_k: Optional[torch.Tensor] = None
    t_admit: Optional[float] = None
    t_done: Optional[float] = None
    inputs: Optional["RequestInputs"] = None  # kept after completion only when the loop is asked to
    status: str = "pending"   # pending, running, done, failed, rejected, cancelled
    error: Optional[str] = None

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


@dataclass
class RequestInputs:
    slab: torch.Tensor       # (S, L, L) complex64
    probe_k: torch.Tensor    # (n, n) complex64
    prop: torch.Tensor       # (2n,) complex64: prop_y then prop_x
    origins: torch.Tensor    # (P, 2) int32
    shifts: torch.Tensor     # (P, 2) float64
    measured: Optional[torch.Tensor] = None  # (P, n, n) float32 for gradient requests
    origins_host: Optional[np.ndarray] = None
    shifts_host: Optional[np.ndarray] = None


def slab_side(n: int) -> int:
    return 2 * n


def request_bytes(req: Request) -> int:
    """Device footprint of a request while resident (inputs, outputs, checkpoints, gradients)."""
    n, s, p = req.n, req.slices, req.positions
    side = slab_side(n)
    b = 8 * s * side * side + 8 * n * n + 16 * n + 16 * p + 4 * p * n * n
    if req.kind == "gradient":
        nc = (s - 1) // CKPT_EVERY
        b += 4 * p * n * n * 2 + 8 * s * side * side + 8 * n * n + 8 * p * nc * n * n
    return b


def build_inputs(req: Request, device) -> RequestInputs:
    """Per-request specimen and scan from the request's seed, built on the device.

    A perovskite-like crystal (the synth generator) fills a periodic slab of
    side 2n; the scan is a sqrt(P) x sqrt(P) raster over the slab with seeded
    sub-pixel remainders.  Gradient requests measure the vacuum-like aperture
    disk with Poisson noise at 1e5 electrons per pattern.
    """
    n, s_count, p_count = req.n, req.slices, req.positions
    calib, dz = CALIB[n], SLICE_DZ[n]
    side = slab_side(n)
    rng = np.random.default_rng(1000 + req.seed)
    cells = int(math.ceil(side * calib / synth.LATTICE_A))
    cells_z = int(math.ceil(s_count * dz / synth.LATTICE_A)) + 1
    atoms = synth.crystal_atoms(cells, cells, cells_z, seed=1000 + req.seed)
    atoms[:, :2] += rng.uniform(0, synth.LATTICE_A, size=2)
    pot = synth.projected_potential_torch(atoms, side, calib, s_count, dz, device)
    slab = torch.empty(pot.shape, dtype=torch.complex64, device=device)
    lib = _native.load()
    _native.check(lib.pty_transmission(_native.ptr(pot), physics.interaction_sigma(VOLTAGE_KV), _native.ptr(slab),
                                       pot.numel(), _native.stream_handle(device)), "pty_transmission")
    probe = synth.probe_for(n, calib, VOLTAGE_KV, APERTURE_MRAD, DEFOCUS, C3)
    probe_k = torch.fft.fft2(torch.from_numpy(probe).to(device)).to(torch.complex64).contiguous()
    pf = physics.propagator_factors(n, dz, VOLTAGE_KV, calib, device=device)
    prop = torch.cat([pf, pf]).contiguous()
    side_pos = int(math.ceil(math.sqrt(p_count)))
    idx = np.arange(p_count)
    step = side / side_pos
    pos = np.stack([(idx // side_pos) * step, (idx % side_pos) * step], axis=1) + rng.uniform(0, 1, (p_count, 2))
    origins = np.floor(pos).astype(np.int32) - n // 2
    shifts = pos - np.floor(pos)
    measured = None
    if req.kind == "gradient":
        lam = physics.relativistic_wavelength_ang(VOLTAGE_KV)
        q = torch.fft.fftshift(physics.fftfreq(n, calib, device=device))
        disk = (torch.sqrt(q[:, None] ** 2 + q[None, :] ** 2) <= (APERTURE_MRAD / 1000.0) / lam).to(torch.float64)
        dose = 1e5
        gen = torch.Generator(device=device)
        gen.manual_seed(2000 + req.seed)
        counts = torch.poisson((dose * disk / disk.sum()).expand(p_count, n, n).contiguous(), generator=gen)
        measured = (counts * (disk.sum() / dose)).to(torch.float32).contiguous()
    return RequestInputs(slab.contiguous(), probe_k, prop, torch.from_numpy(origins).to(device),
                         torch.from_numpy(shifts).to(device), measured, origins, shifts)


def _sqrt_loss_kernel(patterns: torch.Tensor, measured: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """One ``pty_sqrt_intensity_loss`` launch: the FP64 loss (0-dim tensor) and dL/dI (float32)."""
    if not (patterns.is_cuda and measured.is_cuda) or patterns.shape != measured.shape:
        raise ValueError("patterns and measured must be device tensors of the same shape")
    pat = patterns.detach().to(torch.float32).contiguous()
    meas = measured.detach().to(torch.float32).contiguous()
    dev = pat.device
    lib = _native.load()
    grad = torch.empty_like(pat)
    loss = torch.empty((), dtype=torch.float64, device=dev)
    work = torch.empty(int(lib.pty_sqrt_intensity_loss_workspace_bytes()) // 8, dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        status = lib.pty_sqrt_intensity_loss(pat.numel(), _native.ptr(pat), _native.ptr(meas), EPS_SQRT,
                                             _native.ptr(grad), _native.ptr(loss), _native.ptr(work),
                                             work.numel() * 8, _native.stream_handle(dev))
    _native.check(status, "pty_sqrt_intensity_loss")
    return loss, grad


class _SqrtIntensityLoss(torch.autograd.Function):
    """The kernel's loss as a differentiable scalar of the patterns; backward multiplies the kernel's dL/dI by the
    incoming gradient."""

    @staticmethod
    def forward(ctx, patterns, loss, grad):
        ctx.save_for_backward(grad)
        return loss.clone()

    @staticmethod
    def backward(ctx, grad_loss):
        (grad,) = ctx.saved_tensors
        return (grad * grad_loss.to(grad.dtype)).to(grad.dtype), None, None


def loss_and_grad_intensity(patterns: torch.Tensor, measured: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """L = sum (sqrt(I + eps) - sqrt(I_meas))^2 (FP64, fixed summation order) and dL/dI (float32), from one fused
    ``pty_sqrt_intensity_loss`` launch.  The loss is differentiable w.r.t. the patterns when they require grad."""
    loss, grad = _sqrt_loss_kernel(patterns, measured)
    if patterns.requires_grad and torch.is_grad_enabled():
        loss = _SqrtIntensityLoss.apply(patterns, loss, grad)
    return loss, grad


def make_request_stream(count: int, qps: float, seed: int = 0, sizes=(64, 128, 256), slices=(8, 16, 32, 64),
                        scans=(16, 256, 1024, 4096), gradient_fraction: float = 0.3) -> List[Request]:
    """Poisson arrivals at ``qps``; 70 % forward / 30 % gradient; sizes uniform over the given sets."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(count):
        t += rng.exponential(1.0 / qps)
        kind = "gradient" if rng.uniform() < gradient_fraction else "forward"
        reqs.append(Request(i, kind, int(rng.choice(sizes)), int(rng.choice(slices)), int(rng.choice(scans)),
                            seed=int(rng.integers(0, 2**31)), arrival=t))
    return reqs


# ---------------------------------------------------------------------------
# Plain per-request path
# ---------------------------------------------------------------------------


class PlainExecutor:
    """One request at a time, in arrival order, through the ordinary engine calls."""

    def __init__(self, device=None):
        self.device = device or torch.device("cuda", 0)

    def run_one(self, req: Request, inp: Optional[RequestInputs] = None) -> Request:
        dev = self.device
        inp = inp or build_inputs(req, dev)
        one = torch.ones(1, dtype=torch.float32, device=dev)
        prop = inp.prop[: req.n].contiguous()
        pk = inp.probe_k[None]
        if req.kind == "forward":
            req.patterns, _ = engine.multislice_forward(inp.slab, pk, one, inp.origins, inp.shifts, prop)
        else:
            slab = inp.slab.clone().requires_grad_(True)
            pkg = pk.clone().requires_grad_(True)
            pat = autograd.multislice_patterns(slab, pkg, one, inp.origins, inp.shifts, prop, CKPT_EVERY)
            loss, _ = loss_and_grad_intensity(pat, inp.measured)
            loss.backward()
            req.patterns, req.loss = pat.detach(), float(loss.detach())
            req.grad_slab, req.grad_probe_k = slab.grad, pkg.grad[0]
        return req

    def run(self, requests: List[Request], realtime: bool = True) -> Dict:
        reqs = sorted(requests, key=lambda r: r.arrival)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in reqs:
            if realtime:
                wait = r.arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
            r.t_admit = time.perf_counter() - t0
            self.run_one(r)
            torch.cuda.synchronize()
            r.t_done = time.perf_counter() - t0
        return summarize(reqs, time.perf_counter() - t0)


def summarize(reqs: List[Request], wall: float, others: Optional[List[Request]] = None) -> Dict:
    """Throughput and latency of the completed requests; counts of the others by status."""
    lat = np.array([r.latency for r in reqs]) * 1000.0
    pats = sum(r.positions for r in reqs)
    out = {"requests": len(reqs), "patterns": pats, "wall_s": round(wall, 4),
           "patterns_per_s": round(pats / wall, 2) if wall > 0 else 0.0,
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3) if len(lat) else None,
                          "p99": round(float(np.percentile(lat, 99)), 3) if len(lat) else None,
                          "mean": round(float(lat.mean()), 3) if len(lat) else None}}
    for r in others or []:
        out[r.status] = out.get(r.status, 0) + 1
    return out


# ---------------------------------------------------------------------------
# Continuous batching
# ---------------------------------------------------------------------------


@dataclass
class _Active:
    req: Request
    inp: RequestInputs
    next_pos: int = 0
    done_pos: int = 0
    loss: Optional[torch.Tensor] = None  # FP64 device accumulator, read once at completion
    ready: Optional[torch.cuda.Event] = None  # inputs built (on the builder stream)
    pending_loss: Optional[torch.Tensor] = None  # this step's loss, added once the step succeeded


@dataclass
class RequestLoop:
    """Continuous-batching executor (see module docstring)."""

    hbm_budget_bytes: int = 64 << 30
    keep_inputs: bool = False
    admit_per_step: int = 4   # requests built per step while the GPU runs (bounds the host work between launches)
    max_steps_ahead: int = 2  # launched steps the host may queue ahead of the GPU
    chunk_positions: Dict[int, int] = field(default_factory=lambda: {32: 8192, 64: 8192, 128: 4096, 256: 512,
                                                                     512: 128})
    device: Optional[torch.device] = None
    metrics_sink: Optional[Callable[[Dict], None]] = None

    def __post_init__(self):
        self.device = self.device or torch.device("cuda", 0)
        self.base = engine.arena_base(self.device)
        self.streams: Dict[Tuple[int, int], torch.cuda.Stream] = {}
        self.metrics: List[Dict] = []
        self._cancelled: set = set()

    def cancel(self, rid: int) -> None:
        """Drop request ``rid``: not admitted if pending, released at the next step if running."""
        self._cancelled.add(rid)

    def _stream(self, key) -> torch.cuda.Stream:
        if key not in self.streams:
            self.streams[key] = torch.cuda.Stream(device=self.device)
        return self.streams[key]

    def _descs(self, items: List[Tuple[_Active, int, int]], grad: bool, out_tensors) -> torch.Tensor:
        """Descriptor records for (active, start, count) chunks; out_tensors[i] receives chunk i's patterns."""
        total = sum(c for _, _, c in items)
        rec = np.zeros(total, dtype=DESC_DTYPE)
        k = 0
        for (a, start, count), out in zip(items, out_tensors):
            inp = a.inp
            sl = slice(k, k + count)
            rec["slab_offset"][sl] = engine.element_offset(inp.slab, self.base)
            rec["slab_h"][sl] = inp.slab.shape[1]
            rec["slab_w"][sl] = inp.slab.shape[2]
            n2 = a.req.n * a.req.n
            rec["pattern_offset"][sl] = engine.element_offset(out, self.base) + np.arange(count) * n2
            rec["probe_offset"][sl] = engine.element_offset(inp.probe_k, self.base)
            rec["prop_offset"][sl] = engine.element_offset(inp.prop, self.base)
            if grad:
                rec["grad_offset"][sl] = engine.element_offset(a.req.grad_slab, self.base)
                rec["gprobe_offset"][sl] = engine.element_offset(a.req.grad_probe_k, self.base)
            o = inp.origins_host[start:start + count]
            s = inp.shifts_host[start:start + count]
            rec["origin_y"][sl], rec["origin_x"][sl] = o[:, 0], o[:, 1]
            rec["shift_y"][sl], rec["shift_x"][sl] = s[:, 0], s[:, 1]
            k += count
        # pinned staging + asynchronous copy: the upload is ordered on the class stream without
        # making the host wait for that stream's earlier launches
        return torch.from_numpy(rec.view(np.uint8)).pin_memory().to(self.device, non_blocking=True)

    def _step_class(self, key, actives: List[_Active]) -> List[Tuple[_Active, str]]:
        """One batched step of size class key = (n, S) over chunks of the given requests.  Returns the
        requests whose work raised (with the message); if the batch raises, every chunk is retried alone so
        that only the offending requests fail."""
        n = key[0]
        budget = self.chunk_positions.get(n, 512)
        items = []
        for a in actives:
            if budget <= 0:
                break
            count = min(a.req.positions - a.next_pos, budget)
            if count <= 0:
                continue
            items.append((a, a.next_pos, count))
            budget -= count
        completed: list = []
        try:
            self._step_items(key, items, completed)
        except (ValueError, RuntimeError, KeyError):
            # chunks that finished before the error keep their results; the others are retried one by one
            self._advance(completed)
            done_ids = {id(it) for it in completed}
            failed = []
            for item in items:
                if id(item) in done_ids:
                    continue
                alone: list = []
                try:
                    self._step_items(key, [item], alone)
                except (ValueError, RuntimeError, KeyError) as e:
                    failed.append((item[0], f"{type(e).__name__}: {e}"))
                    continue
                self._advance(alone)
            return failed
        self._advance(completed)
        return []

    @staticmethod
    def _advance(items) -> None:
        for a, _, count in items:
            a.next_pos += count
            a.done_pos += count
            if a.pending_loss is not None:
                a.loss = a.pending_loss if a.loss is None else a.loss + a.pending_loss
                a.pending_loss = None

    def _step_items(self, key, items, completed: list) -> None:
        """Launch the forward (and for gradient requests the loss, gradient and adjoint) of the chunks;
        each chunk is appended to ``completed`` once all of its work is enqueued."""
        n, s_count = key
        fwd_items = [it for it in items if it[0].req.kind != "gradient"]
        grad_items = [it for it in items if it[0].req.kind == "gradient"]
        dev = self.device
        if fwd_items:
            outs = [a.req.patterns[start:start + count] for a, start, count in fwd_items]
            descs = self._descs(fwd_items, False, outs)
            engine.forward_batched(n, s_count, descs, self.base)
            completed.extend(fwd_items)
        if grad_items:
            outs = [a.req.patterns[start:start + count] for a, start, count in grad_items]
            total = sum(c for _, _, c in grad_items)
            nc = (s_count - 1) // CKPT_EVERY
            ckpt = torch.empty((total, nc, n, n), dtype=torch.complex64, device=dev) if nc else None
            descs = self._descs(grad_items, True, outs)
            engine.forward_batched(n, s_count, descs, self.base, ckpt, CKPT_EVERY)
            gis = []
            for a, start, count in grad_items:
                loss, gi = loss_and_grad_intensity(a.req.patterns[start:start + count],
                                                   a.inp.measured[start:start + count])
                a.pending_loss = loss
                gis.append(gi)
            # adjoint per request chunk (each a full-GPU launch sequence): one slab per call lets the
            # gradients be gathered in a fixed order instead of accumulated with atomics
            one = torch.ones(1, dtype=torch.float32, device=dev)
            k0 = 0
            for item, gi in zip(grad_items, gis):
                a, start, count = item
                inp = a.inp
                engine.multislice_adjoint(inp.slab, inp.probe_k[None], one, inp.origins[start:start + count],
                                          inp.shifts[start:start + count], inp.prop[:n].contiguous(), gi,
                                          ckpt[k0:k0 + count] if nc else None, CKPT_EVERY,
                                          grad_slab=a.req.grad_slab, grad_probe_k=a.req.grad_probe_k[None])
                completed.append(item)
                k0 += count

    def _admit(self, req: Request, now: float) -> _Active:
        """Build the request's inputs and outputs on the builder stream; the class streams wait for
        the recorded event before the request's first step."""
        dev = self.device
        builder = self._stream("build")
        with torch.cuda.stream(builder):
            inp = build_inputs(req, dev)
            req.patterns = torch.empty((req.positions, req.n, req.n), dtype=torch.float32, device=dev)
            if req.kind == "gradient":
                req.grad_slab = torch.zeros_like(inp.slab)
                req.grad_probe_k = torch.zeros_like(inp.probe_k)
            ready = torch.cuda.Event()
            ready.record(builder)
        req.t_admit = now
        return _Active(req, inp, ready=ready)

    def _admit_arrivals(self, pending: List[Request], active: List[_Active], in_use: int, t0: float,
                        realtime: bool, limit: int, closed: List[Request]) -> int:
        now = time.perf_counter() - t0
        admitted = 0
        while pending and admitted < limit and (not realtime or pending[0].arrival <= now):
            r = pending[0]
            if r.rid in self._cancelled:
                pending.pop(0)
                self._close(r, "cancelled", None, now, closed)
                continue
            need = request_bytes(r)
            if need > self.hbm_budget_bytes:
                pending.pop(0)
                self._close(r, "rejected", f"needs {need} bytes, above the HBM budget of {self.hbm_budget_bytes}",
                            now, closed)
                continue
            if in_use > 0 and in_use + need > self.hbm_budget_bytes:  # something resident (running or draining)
                break
            pending.pop(0)
            try:
                active.append(self._admit(r, now))
            except (ValueError, RuntimeError, KeyError) as e:
                self._close(r, "failed", f"{type(e).__name__}: {e}", now, closed)
                continue
            r.status = "running"
            in_use += need
            admitted += 1
        return in_use

    @staticmethod
    def _close(r: Request, status: str, error: Optional[str], now: float, closed: List[Request]) -> None:
        r.status, r.error, r.t_done = status, error, now
        r.patterns = r.grad_slab = r.grad_probe_k = None
        closed.append(r)

    def run(self, requests: List[Request], realtime: bool = True) -> Dict:
        """Serve ``requests`` to completion.  Steps are launched without waiting for the GPU: the host builds
        arriving requests and launches the next step while earlier steps run, at most ``max_steps_ahead``
        steps ahead.  A request whose positions are all enqueued is retired once the event recorded after its
        last step has completed; only then are its buffers (allocated on the builder stream, read on the class
        streams) released."""
        pending = sorted(requests, key=lambda r: r.arrival)
        active: List[_Active] = []
        draining: List[Tuple[_Active, torch.cuda.Event]] = []  # every position enqueued, GPU work outstanding
        in_flight: List[torch.cuda.Event] = []  # one per launched step, oldest first
        in_use = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        finished: List[Request] = []
        closed: List[Request] = []
        main = torch.cuda.current_stream(self.device)
        while pending or active or draining:
            if not active and not draining:
                in_use = self._admit_arrivals(pending, active, in_use, t0, realtime, 1, closed)
                if not active:
                    if pending:
                        now = time.perf_counter() - t0
                        time.sleep(max(0.0, min(0.005, pending[0].arrival - now)))
                    continue
            classes: Dict[Tuple[int, int], List[_Active]] = {}
            for a in active:
                if a.next_pos < a.req.positions:
                    classes.setdefault((a.req.n, a.req.slices), []).append(a)
            failures: Dict[int, str] = {}
            for key, members in classes.items():
                st = self._stream(key)
                for a in members:
                    if a.ready is not None:
                        st.wait_event(a.ready)
                with torch.cuda.stream(st):
                    for a, message in self._step_class(key, members):
                        failures[a.req.rid] = message
            if classes:
                for key in classes:
                    main.wait_stream(self._stream(key))
                ev = torch.cuda.Event()
                ev.record(main)
                in_flight.append(ev)
            # requests that arrived meanwhile are built while the GPU runs the queued steps; they join the next one
            in_use = self._admit_arrivals(pending, active, in_use, t0, realtime, self.admit_per_step, closed)
            if failures or any(a.req.rid in self._cancelled for a in active):
                torch.cuda.synchronize()  # a dropped request's earlier chunks may still read its buffers
            now = time.perf_counter() - t0
            still = []
            for a in active:
                if a.req.rid in failures or a.req.rid in self._cancelled:
                    in_use -= request_bytes(a.req)
                    if a.req.rid in failures:
                        self._close(a.req, "failed", failures[a.req.rid], now, closed)
                    else:
                        self._close(a.req, "cancelled", None, now, closed)
                elif a.done_pos >= a.req.positions:
                    if a.ready is not None:
                        main.wait_event(a.ready)  # covers the build of a request with no positions
                    done = torch.cuda.Event()
                    done.record(main)
                    draining.append((a, done))
                else:
                    still.append(a)
            active = still
            while len(in_flight) > self.max_steps_ahead:
                in_flight.pop(0).synchronize()
            if draining and not active:
                draining[0][1].synchronize()  # nothing else to launch: wait for the oldest request
            now = time.perf_counter() - t0
            waiting = []
            for a, done in draining:
                if not done.query():
                    waiting.append((a, done))
                    continue
                if a.req.kind == "gradient":
                    a.req.loss = float(a.loss)
                a.req.t_done = now
                if self.keep_inputs:
                    a.req.inputs = a.inp
                a.req.status = "done"
                in_use -= request_bytes(a.req)
                finished.append(a.req)
            draining = waiting
            m = {"t_s": round(now, 6), "active": len(active) + len(draining), "pending": len(pending),
                 "hbm_in_use": in_use, "done": len(finished), "closed": len(closed)}
            self.metrics.append(m)
            if self.metrics_sink is not None:
                self.metrics_sink(m)
        torch.cuda.synchronize()
        return summarize(finished, time.perf_counter() - t0, closed)


# ---------------------------------------------------------------------------
# Scheduling ceiling
# ---------------------------------------------------------------------------


def scheduling_ceiling(requests: List[Request], device=None, floor_positions: int = 4096) -> Dict:
    """Measured bound on what any scheduler of these requests can gain over :class:`PlainExecutor`.

    Each request is run alone on the plain path (warm, synchronised): its input build and its compute are
    timed separately.  For every class (kind, n, S) of the stream one request of ``floor_positions``
    positions is run alone as well; its time per position is the class's saturated cost, the rate at which
    the kernels run when a launch fills the GPU.  A scheduler that hides every input build behind other
    work and runs every position at its class's saturated cost needs ``floor_ms`` = sum of positions x
    class cost; ``max_speedup`` = plain wall time / floor_ms bounds the speed-up of batching, overlap and
    graph capture together, without faster kernels."""
    dev = device or torch.device("cuda", 0)
    plain = PlainExecutor(dev)

    def timed(req, inp=None):
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        inp = inp or build_inputs(req, dev)
        torch.cuda.synchronize(dev)
        t1 = time.perf_counter()
        plain.run_one(req, inp)
        torch.cuda.synchronize(dev)
        return inp, (t1 - t0) * 1e3, (time.perf_counter() - t1) * 1e3

    rows = []
    for r in sorted(requests, key=lambda q: q.arrival):
        probe = Request(r.rid, r.kind, r.n, r.slices, r.positions, r.seed)
        inp, build_ms, _ = timed(probe)
        _, _, compute_ms = timed(probe, inp)          # second run: warm caches and workspaces
        rows.append({"rid": r.rid, "kind": r.kind, "n": r.n, "slices": r.slices, "positions": r.positions,
                     "build_ms": round(build_ms, 3), "compute_ms": round(compute_ms, 3)})
        del inp
    classes = {}
    for row in rows:
        key = (row["kind"], row["n"], row["slices"])
        if key in classes:
            continue
        big = Request(-1, key[0], key[1], key[2], floor_positions, seed=7)
        inp, _, _ = timed(big)
        _, _, ms = timed(big, inp)
        classes[key] = ms / floor_positions
        del inp
    for row in rows:
        row["floor_ms"] = round(row["positions"] * classes[(row["kind"], row["n"], row["slices"])], 4)
    plain_ms = sum(r["build_ms"] + r["compute_ms"] for r in rows)
    floor_ms = sum(r["floor_ms"] for r in rows)
    top = max(rows, key=lambda r: r["compute_ms"])
    return {"plain_build_ms": round(sum(r["build_ms"] for r in rows), 2),
            "plain_compute_ms": round(sum(r["compute_ms"] for r in rows), 2),
            "floor_ms": round(floor_ms, 2), "max_speedup": round(plain_ms / floor_ms, 3),
            "largest_request": {k: top[k] for k in ("rid", "kind", "n", "slices", "positions", "compute_ms")},
            "class_us_per_position": {f"{k}_{n}_{s}": round(v * 1e3, 3) for (k, n, s), v in sorted(classes.items())},
            "requests": rows}
