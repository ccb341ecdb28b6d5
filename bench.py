"""Benchmark: 4D-STEM patterns/s of the multislice forward on one MI355X.

Workload (default): 256^2 complex64 windows cropped from a 1024^2 slab,
64 slices of 1 Angstrom, 300 kV / 25 mrad / C1 = 50 A / Cs = 1 um probe,
64 x 64 = 4096 scan positions, float32 patterns.  One step = one pass of the
forward over all 4096 positions with inputs resident in HBM.

Prints ONE JSON line.  ``roofline`` uses the algorithmic bytes per pattern
B = 8 S n^2 + 8 n^2 + 4 n^2 (transmission window per slice, incident probe,
float32 pattern) over the device time of the forward measured with HIP
events on the launch stream.  ``cpu_baseline`` times the oracle's numpy /
pocketfft complex128 restatement of the reference algorithm on a bounded
sample of the same workload (rank 0, N = 1 only).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
# VALU issue rate: one wave64 vector instruction per CU per cycle (4 SIMDs x 16 lanes), 256 CUs at 2.4 GHz
VALU_ISSUE_PER_S = 256 * 2.4e9

# HBM-side bytes per step of the config-2 forward, measured by two rocprofv3 --pmc passes (FETCH_SIZE, then
# WRITE_SIZE) and calibrated with bench.py --workload calib; the measurement records the sha256 of the kernel
# sources it was taken on, and is reported only while those sources are unchanged.
TRAFFIC_PROFILE = ROOT / "profiles" / "forward_256x64_traffic.json"


def measured_traffic(workload: str, path: str, batch: int):
    """(bytes per step, note) from the committed counter measurement, or (None, reason)."""
    import hashlib
    if not TRAFFIC_PROFILE.exists():
        return None, "no counter measurement committed"
    doc = json.loads(TRAFFIC_PROFILE.read_text())
    if (doc.get("workload"), doc.get("path"), doc.get("batch_waves")) != (workload, path, batch):
        return None, f"measured for {doc.get('workload')} / {doc.get('path')} / {doc.get('batch_waves')} waves"
    h = hashlib.sha256()
    for src in doc["source_files"]:
        h.update((ROOT / src).read_bytes())
    if h.hexdigest() != doc["source_sha256"]:
        return None, "kernel sources changed since the counter measurement"
    return int(doc["traffic_bytes"]), f"rocprofv3 FETCH_SIZE/WRITE_SIZE, {TRAFFIC_PROFILE.relative_to(ROOT)}"


def measured_valu(workload: str, path: str, batch: int, kernel_ms: float):
    """Wavefront VALU instructions per step (rocprofv3 SQ_INSTS_VALU, committed beside the traffic counts under
    the same source hash) and the fraction of the device's issue rate they take in ``kernel_ms``, or None."""
    traffic, _ = measured_traffic(workload, path, batch)
    if traffic is None:
        return None
    doc = json.loads(TRAFFIC_PROFILE.read_text())
    if "valu_instructions" not in doc:
        return None
    instr = float(doc["valu_instructions"])
    issue_s = instr / VALU_ISSUE_PER_S
    return {"instructions": instr, "issue_ms": round(issue_s * 1e3, 2), "peak_per_s": VALU_ISSUE_PER_S,
            "frac": round(issue_s / (kernel_ms * 1e-3), 4),
            "source": f"rocprofv3 SQ_INSTS_VALU, {TRAFFIC_PROFILE.relative_to(ROOT)}"}


def algorithmic_bytes(n: int, s: int) -> int:
    return 8 * s * n * n + 8 * n * n + 4 * n * n


def algorithmic_flops(n: int, s: int) -> float:
    return 2 * s * 5 * n * n * np.log2(n * n) + 6 * n * n * 2 * s + 3 * n * n


def cpu_baseline(w, sample: int, cores: int):
    """Oracle (numpy/pocketfft, complex128, `cores` workers) on `sample` positions."""
    from oracle import ms_oracle as o
    prop = o.propagation_func(w.n, w.n, w.slice_thickness, w.voltage_kv, w.calib)
    sel = np.linspace(0, len(w.origins) - 1, sample).astype(int)
    slab = w.slab_t  # complex64 input, arithmetic promoted to complex128 inside
    o.window_stem_batched(slab, w.probe, w.origins[sel[:2]], w.shifts_px[sel[:2]], prop, workers=cores)
    def timed(dtype):
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            o.window_stem_batched(slab, w.probe, w.origins[sel], w.shifts_px[sel], prop, workers=cores, dtype=dtype)
            times.append(time.perf_counter() - t0)
        return sample / float(np.median(times))

    c128 = timed(np.complex128)
    c64 = timed(np.complex64)
    return {"value": round(c128, 3), "unit": "patterns/s", "cores": cores, "kind": "port",
            "sample": f"{sample} of {len(w.origins)} positions (every {len(w.origins) // sample}th), complex128, "
                      f"median of 3 after 1 warm-up; reference-algorithm CPU restatement (numpy/pocketfft)",
            "complex64_value": round(c64, 3)}


def detectors_bench(w, slab, pk, weights, origins, shifts, prop, steps: int):
    """The same forward with three annular virtual detectors (BF 0-25, ABF 12.5-25, ADF 50-95 mrad) fused into
    the epilogue and no pattern stack written."""
    import torch
    from ptyrodactyl_amd import engine, simulations
    angles = [(0.0, 25.0), (12.5, 25.0), (50.0, 95.0)]
    masks = torch.stack([simulations.annular_mask((w.n, w.n), 1.0 / (w.n * w.calib), w.voltage_kv, lo, hi,
                                                  device=slab.device) for lo, hi in angles]).contiguous()
    engine.multislice_forward_detectors(slab, pk, weights, origins, shifts, prop, masks)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        engine.multislice_forward_detectors(slab, pk, weights, origins, shifts, prop, masks)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    return {"detectors_mrad": angles, "patterns_written": False, "ms_per_step": round(ms, 3),
            "patterns_per_s": round(len(w.origins) / (ms * 1e-3), 1)}


def captured_small_bench(w, slab, pk, weights, origins, shifts, prop, positions: int = 16, reps: int = 20):
    """Latency of a small request (the first `positions` positions of the workload): the eager launch sequence
    against its HIP-graph replay (engine.CapturedForward)."""
    import torch
    from ptyrodactyl_amd import engine
    o, s = origins[:positions].contiguous(), shifts[:positions].contiguous()
    cap = engine.CapturedForward(slab, pk, weights, o, s, prop)
    out = torch.empty((positions, w.n, w.n), dtype=torch.float32, device=slab.device)
    res = {}
    for name, fn in (("eager", lambda: engine.multislice_forward(slab, pk, weights, o, s, prop, patterns_out=out)),
                     ("graph", cap.replay)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        res[f"{name}_ms"] = round((time.perf_counter() - t0) * 1000.0 / reps, 3)
    res["positions"] = positions
    res["launches_per_call"] = w.num_slices + 2 if w.n >= 256 else 1
    return res


def gradient_algorithmic_bytes(n: int, s: int, k: int) -> int:
    """Per pattern: transmission read three times (forward, segment recompute, adjoint) and its gradient
    written once, checkpoints written and read, probe read and probe gradient written, pattern written
    and dL/dI read."""
    nc = (s - 1) // k
    return 8 * n * n * (4 * s + 2 * nc + 2) + 8 * n * n


def gradient_bench(steps: int, warmup: int):
    """Gradient workload (16 x 512^2 object, 128^2 probe, 32 x 32 jittered scan, sqrt-intensity loss,
    checkpoints every 4 slices): one step = forward with checkpoints + loss gradient + adjoint over all
    1024 positions, gradients w.r.t. every slice and the probe."""
    import torch
    from ptyrodactyl_amd import engine, physics, serving, synth
    w = synth.config3()
    dev = torch.device("cuda", 0)
    obj = torch.from_numpy(w.object_t).to(dev)
    pk = torch.fft.fft2(torch.from_numpy(w.probe).to(dev)).to(torch.complex64)[None].contiguous()
    weights = torch.ones(1, dtype=torch.float32, device=dev)
    origins = torch.from_numpy(w.origins).to(dev)
    shifts = torch.from_numpy(w.shifts_px).to(dev)
    prop = physics.propagator_factors(w.n, w.slice_thickness, w.voltage_kv, w.calib, device=dev)
    # measured data: Poisson counts of the unperturbed object's patterns (GPU forward, seeded)
    clean, _ = engine.multislice_forward(torch.from_numpy(w.object_true).to(dev), pk, weights, origins, shifts, prop)
    measured = torch.from_numpy(synth.poisson_measure(clean.double().cpu().numpy())).to(dev, torch.float32)
    grad_obj = torch.zeros_like(obj)
    grad_pk = torch.zeros_like(pk)

    def step():
        grad_obj.zero_()
        grad_pk.zero_()
        pat, ckpt = engine.multislice_forward_checkpointed(obj, pk, weights, origins, shifts, prop, w.ckpt_every)
        _, g = serving.loss_and_grad_intensity(pat, measured)
        engine.multislice_adjoint(obj, pk, weights, origins, shifts, prop, g, ckpt, w.ckpt_every,
                                  grad_slab=grad_obj, grad_probe_k=grad_pk)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * steps)]
    for i in range(steps):
        ev[2 * i].record()
        step()
        ev[2 * i + 1].record()
    torch.cuda.synchronize()
    ms = [ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(steps)]
    p_count = len(w.origins)
    kernel_ms = float(np.mean(ms))
    b = gradient_algorithmic_bytes(w.n, w.num_slices, w.ckpt_every)
    achieved = b * p_count / (kernel_ms * 1e-3) / 1e9
    return {"workload": "gradient_128x16_p1024_obj512_ckpt4", "patterns_per_s": round(p_count / (kernel_ms * 1e-3), 1),
            "ms_per_step": round(kernel_ms, 3), "steps": steps,
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_pattern": b,
                         "formula": "8 n^2 (4 S + 2 floor((S-1)/k) + 2) + 8 n^2"}}


def gradient256_bench(steps: int, warmup: int):
    """The request loop's largest request as a workload of its own: a gradient at 256^2 windows, 32 slices, 4096
    positions over a 512^2 synthetic crystal slab (the serving input builder, seed 11): checkpointed forward, the
    sqrt-intensity loss gradient and the adjoint w.r.t. the slab and the probe, through the drop-in autograd path.
    Timed with the drop-in's own checkpoint interval (every slice here: 62 GiB of store) and with every 4th slice
    (what the request loop uses to keep its HBM budget)."""
    import torch
    from ptyrodactyl_amd import autograd, serving
    dev = torch.device("cuda", 0)
    req = serving.Request(0, "gradient", 256, 32, 4096, seed=11)
    inp = serving.build_inputs(req, dev)
    one = torch.ones(1, dtype=torch.float32, device=dev)
    prop = inp.prop[:256].contiguous()

    def timed(k):
        def step():
            slab = inp.slab.clone().requires_grad_(True)
            pk = inp.probe_k[None].clone().requires_grad_(True)
            pat = autograd.multislice_patterns(slab, pk, one, inp.origins, inp.shifts, prop, k)
            loss, _ = serving.loss_and_grad_intensity(pat, inp.measured)
            loss.backward()

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * steps)]
        for i in range(steps):
            ev[2 * i].record()
            step()
            ev[2 * i + 1].record()
        torch.cuda.synchronize()
        return float(np.mean([ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(steps)]))

    k_default = autograd.checkpoint_interval(inp.slab, inp.probe_k[None], inp.origins)
    ms = timed(None)
    ms4 = timed(serving.CKPT_EVERY)
    b = gradient_algorithmic_bytes(256, 32, 4)  # one definition of the work for every checkpoint interval
    achieved = b * req.positions / (ms * 1e-3) / 1e9
    return {"workload": "gradient_256x32_p4096_slab512", "patterns_per_s": round(req.positions / (ms * 1e-3), 1),
            "ms_per_step": round(ms, 3), "steps": steps, "ckpt_every": k_default,
            "ckpt_every_4": {"ms_per_step": round(ms4, 3), "patterns_per_s": round(req.positions / (ms4 * 1e-3), 1)},
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_pattern": b,
                         "formula": "8 n^2 (4 S + 2 floor((S-1)/4) + 2) + 8 n^2"}}


def serving_bench(count: int, qps: float):
    """Mixed request stream (70 % forward / 30 % gradient, n in {64,128,256}, S in {8,16,32,64},
    P in {16,256,1024,4096}, Poisson arrivals): continuous batching vs one request at a time.
    Throughput is taken at saturation (all requests arrive within a few ms); latency is reported both
    there and at the paced rate `qps`."""
    from ptyrodactyl_amd import serving
    warm = serving.make_request_stream(6, qps=1e6, seed=99, scans=(16, 64))
    serving.RequestLoop().run(warm, realtime=False)
    serving.PlainExecutor().run(serving.make_request_stream(6, qps=1e6, seed=99, scans=(16, 64)), realtime=False)
    out = {"requests": count,
           "mix": "70% forward / 30% gradient; n {64,128,256}; S {8,16,32,64}; P {16,256,1024,4096}; Poisson arrivals"}
    for label, rate in (("saturated", 1e4), (f"paced_{qps:g}qps", qps)):
        plain = serving.PlainExecutor().run(serving.make_request_stream(count, rate, seed=5))
        loop = serving.RequestLoop().run(serving.make_request_stream(count, rate, seed=5))
        print(f"# serving {label} plain {plain}", file=sys.stderr, flush=True)
        print(f"# serving {label} batched {loop}", file=sys.stderr, flush=True)
        out[label] = {"qps": rate, "batched": loop, "plain": plain,
                      "speedup": round(loop["patterns_per_s"] / plain["patterns_per_s"], 3)}
    ceiling = serving.scheduling_ceiling(serving.make_request_stream(count, 1e4, seed=5))
    print(f"# serving ceiling {json.dumps({k: v for k, v in ceiling.items() if k != 'requests'})}", file=sys.stderr,
          flush=True)
    out["ceiling"] = {k: v for k, v in ceiling.items() if k != "requests"}
    return out


def potential_bench(steps: int = 3):
    """Projected-potential slices of the headline slab (1024^2 at 0.1 A, 64 slices of 1 A, perovskite-like
    crystal from ``synth.crystal_atoms``, Lobato form, supersampling 4) built on the GPU.  The coefficient rows
    for Sr, Ti and O are synthetic (positive amplitudes and scales): the kernels' cost does not depend on the
    values.  Times the whole builder and, with HIP events on its stream, the structure-factor kernel alone; the
    kernel's algorithmic work is 8 FP64 flops (4 FMA) per (atom, Fourier pixel) of its slice."""
    import math
    import torch
    from ptyrodactyl_amd import _native, atom_potentials as ap, form_factors as ff, synth
    from ptyrodactyl_amd.types import create_crystal_data
    grid, calib, s_count = 1024, 0.1, 64
    cells = int(math.ceil(grid * calib / synth.LATTICE_A))
    atoms = synth.crystal_atoms(cells, cells, int(math.ceil(s_count / synth.LATTICE_A)) + 1, seed=0)
    saved = dict(ff._TABLES)
    ff.set_form_factor_table("lobato", {z: [z / 50.0, 0.5, z / 60.0, 1.2, z / 80.0, 3.0, z / 100.0, 8.0, z / 120.0,
                                            20.0] for z in (8, 22, 38)})
    try:
        dev = torch.device("cuda", 0)
        crystal = create_crystal_data(atoms[:, [1, 0, 2]], atoms[:, 3].astype(np.int64))
        crystal = type(crystal)(crystal.positions.to(dev), crystal.atomic_numbers.to(dev), crystal.lattice)
        shape = (grid, grid, s_count)

        def build():
            return ap.crystal_potential_slices(crystal, calib, 1.0, padding=0.0, supersampling=4, grid_shape=shape)

        build()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            out = build()
        torch.cuda.synchronize()
        total_ms = (time.perf_counter() - t0) * 1e3 / steps
        plan = ap.plan_slices(crystal, calib, 1.0, (1, 1, 1), 0.0, shape, dev)
        rows = np.stack([ff.coefficient_row(int(z), "lobato") for z in plan.elements.tolist()])
        kern_k = torch.fft.fft2(ap._atom_grids(rows, "lobato", calib, grid, grid, 4, None, dev)).contiguous()
        fk = torch.empty((s_count, grid, grid), dtype=torch.complex128, device=dev)
        lib = _native.load()

        def kernel():
            st = lib.pty_potential_slices_fourier(s_count, rows.shape[0], grid, grid, plan.shifts.shape[0],
                                                  _native.ptr(plan.shifts), _native.ptr(plan.segments),
                                                  _native.ptr(kern_k), _native.ptr(fk), _native.stream_handle(dev))
            _native.check(st, "pty_potential_slices_fourier")

        kernel()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            kernel()
        e1.record()
        torch.cuda.synchronize()
        k_ms = e0.elapsed_time(e1) / steps
        n_atoms = int(plan.shifts.shape[0])
        flops = 8.0 * n_atoms * grid * grid
        tf = flops / (k_ms * 1e-3) / 1e12
        return {"workload": "crystal_potential_slices_1024x1024x64_lobato_ss4", "atoms": n_atoms,
                "elements": plan.elements.tolist(), "coefficients": "synthetic (timing only)",
                "builder_ms": round(total_ms, 2), "slices_per_s": round(s_count / (total_ms * 1e-3), 1),
                "structure_kernel": {"kernel": "k_slice_structure", "ms": round(k_ms, 3), "bound": "fp64",
                                     "achieved": round(tf, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                     "frac": round(tf / FP64_PEAK_TFLOPS, 4),
                                     "formula": "8 flops x atoms x H x W (each atom over its slice's grid)"},
                "output_checksum": float(out.slices.sum())}
    finally:
        ff._TABLES.update(saved)


def grid_bench(steps: int = 3, cases=((100, 100, 16, 1024), (250, 250, 16, 512), (550, 500, 16, 128))):
    """Forward and gradient on grids that are not square powers of two (the runtime-size kernels): per case
    (h, w, slices, positions) a random transmission slab 64 pixels larger than the window, one probe mode,
    random origins and shifts.  Algorithmic bytes per pattern as for the headline: S h w 8 + h w 8 + h w 4.
    The gradient step is the checkpointed forward (every 4 slices) and the adjoint with slab, probe and shift
    gradients."""
    import torch
    from ptyrodactyl_amd import engine, physics
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(0)
    out = []

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    for i, (h, w, s_count, p_count) in enumerate(cases):
        gen = torch.Generator(device=dev).manual_seed(i)
        slab = torch.polar(torch.ones((s_count, h + 64, w + 64), device=dev),
                           0.4 * torch.rand((s_count, h + 64, w + 64), device=dev, generator=gen)).contiguous()
        pk = torch.fft.fft2(torch.randn((1, h, w), dtype=torch.complex64, device=dev, generator=gen)).contiguous()
        wt = torch.ones(1, dtype=torch.float32, device=dev)
        org = torch.tensor(rng.integers(0, 64, (p_count, 2)), dtype=torch.int32, device=dev)
        sh = torch.tensor(rng.uniform(-2, 2, (p_count, 2)), device=dev)
        prop = physics.propagator_pair(h, w, 1.0, 300.0, 0.1, device=dev)
        ms = timed(lambda: engine.multislice_forward(slab, pk, wt, org, sh, prop))
        bytes_per = (s_count * h * w * 8 + h * w * 8 + h * w * 4)
        row = {"grid": [h, w], "slices": s_count, "positions": p_count, "forward_ms": round(ms, 3),
               "patterns_per_s": round(p_count / (ms * 1e-3), 1),
               "algorithmic_gbs": round(bytes_per * p_count / (ms * 1e-3) / 1e9, 1)}
        g = torch.randn((p_count, h, w), device=dev, generator=gen)

        def grad_step():
            _, ck = engine.multislice_forward_checkpointed(slab, pk, wt, org, sh, prop, 4)
            gs = torch.zeros((p_count, 2), dtype=torch.float64, device=dev)
            engine.multislice_adjoint(slab, pk, wt, org, sh, prop, g, ck, 4, grad_shifts=gs)

        gms = timed(grad_step)
        row["gradient_ms"] = round(gms, 3)
        row["gradient_patterns_per_s"] = round(p_count / (gms * 1e-3), 1)
        out.append(row)
        print(f"# grid {row}", file=sys.stderr, flush=True)
    return out


def calibration_kernels(sizes_mib=(64, 2048), reps: int = 5):
    """Kernels of known traffic for calibrating the TCC byte counters under rocprofv3 --pmc: per size, `reps`
    launches each of a read-only reduction (a.sum(): n * 4 bytes read), a write-only fill (b.fill_(): n * 4
    bytes written) and a copy (b.copy_(a): n * 4 read + n * 4 written).  64 MiB stays in the 256 MB
    Infinity Cache between launches, 2 GiB does not.  Prints the launch plan; timing is incidental."""
    import torch
    dev = torch.device("cuda", 0)
    plan = []
    for mib in sizes_mib:
        n = mib * (1 << 20) // 4
        a = torch.ones(n, device=dev)
        b = torch.empty_like(a)
        s = torch.empty((), device=dev)
        for _ in range(reps):
            s.copy_(a.sum())
        for _ in range(reps):
            b.fill_(2.0)
        for _ in range(reps):
            b.copy_(a)
        torch.cuda.synchronize()
        plan.append({"mib": mib, "bytes": n * 4, "reps": reps, "order": ["sum (read)", "fill (write)", "copy"]})
        del a, b
    return plan


def membw_bench(sizes_mib=(256, 512, 1024, 4096), reps: int = 20):
    """Achievable HBM bandwidth on this box: torch device copies (read + write) and reductions (read) of
    float32 buffers, timed with events; the practical ceiling the streaming passes are compared with."""
    import torch
    dev = torch.device("cuda", 0)
    out = []
    for mib in sizes_mib:
        n = mib * (1 << 20) // 4
        a = torch.ones(n, device=dev)
        b = torch.empty_like(a)
        s = torch.empty((), device=dev)
        row = {"mib": mib}
        for label, fn, moved in (("copy", lambda: b.copy_(a), 2 * n * 4), ("read", lambda: s.copy_(a.sum()), n * 4)):
            for _ in range(3):
                fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            row[f"{label}_gbs"] = round(moved / (ms * 1e-3) / 1e9, 1)
        out.append(row)
        del a, b
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-waves", type=int, default=None)
    ap.add_argument("--path", choices=("streaming", "persistent"), default="streaming",
                    help="n=256 forward: per-pass launches, or the single persistent XCD-local launch")
    ap.add_argument("--group-waves", type=int, default=2)
    ap.add_argument("--blocks-per-cu", type=int, default=3)
    ap.add_argument("--serving-requests", type=int, default=32)
    ap.add_argument("--serving-qps", type=float, default=60.0)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-gradient", action="store_true")
    ap.add_argument("--only-forward", action="store_true",
                    help="time the headline forward only (no detector, graph, gradient, potential, serving or CPU "
                         "legs): the run rocprofv3 counter passes profile")
    ap.add_argument("--workload", choices=("config2", "config1", "serving", "gradient", "gradient256", "membw",
                                           "potential", "grid",
                                           "calib"),
                    default="config2",
                    help="config2: 256^2 x 64 slices from a 1024^2 slab, 4096 positions (the headline); "
                         "config1: 128^2 x 20 slices, 256 positions (the reference's CPU-runnable case)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    import torch

    from ptyrodactyl_amd import engine, physics, synth

    if args.workload == "serving":
        res = serving_bench(args.serving_requests, args.serving_qps)
        print(json.dumps({"metric": "4D-STEM patterns/s, mixed request stream",
                          "value": res["saturated"]["batched"]["patterns_per_s"],
                          "unit": "patterns/s", "n_gpus": 1, "higher_is_better": True, "serving": res}), flush=True)
        return 0

    if args.workload == "potential":
        res = potential_bench(max(args.steps, 1))
        print(json.dumps({"metric": "potential slices/s, 1024^2 x 64 slab", "value": res["slices_per_s"],
                          "unit": "slices/s", "higher_is_better": True, "n_gpus": 1, "dtype": "f64",
                          "data": "synthetic", "potential_builder": res}), flush=True)
        return 0
    if args.workload == "grid":
        res = grid_bench(max(args.steps, 1))
        print(json.dumps({"metric": "4D-STEM patterns/s on grids other than square powers of two",
                          "value": res[0]["patterns_per_s"], "unit": "patterns/s", "n_gpus": 1,
                          "higher_is_better": True, "dtype": "complex64", "data": "synthetic", "grid": res}),
              flush=True)
        return 0
    if args.workload == "calib":
        print(json.dumps({"workload": "counter_calibration", "plan": calibration_kernels()}), flush=True)
        return 0
    if args.workload == "membw":
        res = membw_bench()
        print(json.dumps({"metric": "achievable HBM bandwidth (torch copy / reduction)",
                          "value": max(r["copy_gbs"] for r in res), "unit": "GB/s", "n_gpus": 1,
                          "higher_is_better": True, "membw": res}), flush=True)
        return 0

    if args.workload == "gradient256":
        res = gradient256_bench(args.steps, args.warmup)
        print(json.dumps({"metric": "gradient patterns/s at 256^2 x 32 slices", "value": res["patterns_per_s"],
                          "unit": "patterns/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": res["ms_per_step"], "higher_is_better": True, "gradient": res}), flush=True)
        return 0
    if args.workload == "gradient":
        res = gradient_bench(args.steps, args.warmup)
        print(json.dumps({"metric": "gradient patterns/s at 128^2 x 16 slices", "value": res["patterns_per_s"],
                          "unit": "patterns/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": res["ms_per_step"], "higher_is_better": True, "gradient": res}), flush=True)
        return 0

    t0 = time.perf_counter()
    w = synth.config2() if args.workload == "config2" else synth.config1()
    print(f"# workload built in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    dev = torch.device("cuda", 0)
    slab = torch.from_numpy(w.slab_t).to(dev)
    pk = torch.fft.fft2(torch.from_numpy(w.probe[None]).to(dev)).to(torch.complex64).contiguous()
    weights = torch.ones(1, dtype=torch.float32, device=dev)
    origins = torch.from_numpy(w.origins).to(dev)
    shifts = torch.from_numpy(w.shifts_px).to(dev)
    prop = physics.propagator_factors(w.n, w.slice_thickness, w.voltage_kv, w.calib, device=dev)
    p_count = len(w.origins)
    out = torch.empty((p_count, w.n, w.n), dtype=torch.float32, device=dev)

    persistent = args.path == "persistent" and w.n == 256
    if persistent:
        lib = engine._native.load()
        state = torch.zeros(int(lib.pty_persistent_state_bytes()) // 4, dtype=torch.int32, device=dev)
        pws = torch.empty(int(lib.pty_persistent_workspace_bytes(w.n, args.group_waves)), dtype=torch.uint8,
                          device=dev)
        grid = torch.cuda.get_device_properties(dev).multi_processor_count * args.blocks_per_cu

    def step():
        if persistent:
            engine.multislice_forward_persistent(slab, pk, weights, origins, shifts, prop, patterns_out=out,
                                                 group_waves=args.group_waves, grid=grid, state=state,
                                                 workspace=pws, check=False)
        else:
            engine.multislice_forward(slab, pk, weights, origins, shifts, prop, patterns_out=out,
                                      max_batch_waves=args.batch_waves)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t_start = time.perf_counter()
    for i in range(args.steps):
        starts[i].record()
        step()
        ends[i].record()
        print(f"# step {i} issued", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    if persistent and int(state[0].item()) != 0:
        print("persistent forward reported an error flag", file=sys.stderr, flush=True)
        return 1
    dev_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    ms_per_step = wall * 1000.0 / args.steps
    value = p_count * args.steps / wall
    kernel_ms = float(np.mean(dev_ms))
    bytes_step = algorithmic_bytes(w.n, w.num_slices) * p_count
    achieved = bytes_step / (kernel_ms * 1e-3) / 1e9
    batch = args.batch_waves or engine.DEFAULT_BATCH_WAVES
    traffic, traffic_source = measured_traffic(w.name, "persistent" if persistent else "streaming", batch)
    flops_step = algorithmic_flops(w.n, w.num_slices) * p_count
    result = {
        "metric": "4D-STEM patterns/s at 256^2 x 64 slices" if args.workload == "config2" else
                  "4D-STEM patterns/s at 128^2 x 20 slices",
        "value": round(value, 2),
        "unit": "patterns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64",
        "data": "synthetic (seeded perovskite-like crystal, generator in ptyrodactyl_amd/synth.py)",
        "config": {"workload": w.name, "window": w.n, "slices": w.num_slices, "positions": p_count,
                   "slab": list(w.slab_t.shape[1:]), "voltage_kv": w.voltage_kv, "aperture_mrad": w.aperture_mrad,
                   "defocus_ang": w.defocus_ang, "cs_um": w.c3_ang * 1e3,
                   "path": "persistent" if persistent else "streaming"},
        "latency_ms": {"p50": round(float(np.percentile(dev_ms, 50)), 3),
                       "p99": round(float(np.percentile(dev_ms, 99)), 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": bytes_step,
                     "traffic_gbs": round(traffic / (kernel_ms * 1e-3) / 1e9, 1) if traffic else None,
                     "traffic_source": traffic_source,
                     "kernel": "multislice forward (all launches of one step)", "kernel_ms": round(kernel_ms, 3),
                     "fp32_tflops": round(flops_step / (kernel_ms * 1e-3) / 1e12, 2),
                     "fp32_frac": round(flops_step / (kernel_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 4),
                     "valu_issue": measured_valu(w.name, "persistent" if persistent else "streaming", batch,
                                                 kernel_ms)},
    }
    if args.only_forward:
        print(json.dumps(result), flush=True)
        return 0
    if args.workload == "config2" and not persistent:
        result["virtual_detectors"] = detectors_bench(w, slab, pk, weights, origins, shifts, prop, args.steps)
        result["small_request_graph"] = captured_small_bench(w, slab, pk, weights, origins, shifts, prop)
        print(f"# small request {result['small_request_graph']}", file=sys.stderr, flush=True)
        print(f"# detectors {result['virtual_detectors']}", file=sys.stderr, flush=True)
    if not args.no_gradient and args.workload == "config2":
        result["gradient"] = gradient_bench(max(args.steps, 3), 1)
        print(f"# gradient {result['gradient']}", file=sys.stderr, flush=True)
    if args.workload == "config2" and not persistent:
        result["potential_builder"] = potential_bench(max(min(args.steps, 3), 1))
        print(f"# potential builder {result['potential_builder']}", file=sys.stderr, flush=True)
    if not args.no_serving and args.workload == "config2":
        del slab, out
        torch.cuda.empty_cache()
        result["serving"] = serving_bench(args.serving_requests, args.serving_qps)
    if not args.no_cpu_baseline:
        cores = len(os.sched_getaffinity(0))
        result["cpu_baseline"] = cpu_baseline(w, args.cpu_sample, cores)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
