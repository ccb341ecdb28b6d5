"""Request loop: CPU checks of the request model and GPU checks that every
result is the same alone and in the mix, and that admission respects the
HBM budget."""

import numpy as np
import pytest
import torch

from ptyrodactyl_amd import serving


def test_request_stream_mix_and_arrivals():
    reqs = serving.make_request_stream(400, qps=50.0, seed=1)
    frac = np.mean([r.kind == "gradient" for r in reqs])
    assert 0.22 < frac < 0.38
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0)
    assert 400 / arr[-1] == pytest.approx(50.0, rel=0.2)
    assert {r.n for r in reqs} == {64, 128, 256}
    assert {r.slices for r in reqs} == {8, 16, 32, 64}
    assert {r.positions for r in reqs} == {16, 256, 1024, 4096}


def test_request_bytes_grow_with_size():
    a = serving.Request(0, "forward", 64, 8, 16, 0)
    b = serving.Request(1, "forward", 256, 64, 4096, 0)
    g = serving.Request(2, "gradient", 256, 64, 4096, 0)
    assert 0 < serving.request_bytes(a) < serving.request_bytes(b) < serving.request_bytes(g)


def _mix():
    specs = [("forward", 64, 8, 16), ("gradient", 64, 8, 16), ("forward", 128, 16, 64), ("gradient", 128, 8, 40),
             ("forward", 256, 8, 24), ("gradient", 256, 8, 12), ("forward", 64, 8, 300), ("forward", 128, 16, 20),
             ("gradient", 64, 16, 33), ("forward", 256, 8, 9)]
    return [serving.Request(i, k, n, s, p, seed=10 + i, arrival=0.001 * i) for i, (k, n, s, p) in enumerate(specs)]


@pytest.mark.gpu
def test_results_same_alone_and_in_mix():
    loop = serving.RequestLoop(hbm_budget_bytes=256 << 20, chunk_positions={64: 128, 128: 48, 256: 16},
                               keep_inputs=True)
    mixed = _mix()
    stats = loop.run(mixed, realtime=False)
    assert stats["requests"] == len(mixed)
    plain = serving.PlainExecutor()
    for r in mixed:
        alone = plain.run_one(serving.Request(r.rid, r.kind, r.n, r.slices, r.positions, r.seed), r.inputs)
        print(f"request {r.rid} {r.kind} n={r.n} S={r.slices} P={r.positions}", flush=True)
        assert torch.equal(r.patterns, alone.patterns), f"patterns differ for request {r.rid}"
        if r.kind == "gradient":
            assert r.loss == pytest.approx(alone.loss, rel=1e-9)
            for got, ref in ((r.grad_slab, alone.grad_slab), (r.grad_probe_k, alone.grad_probe_k)):
                err = float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref))
                assert err < 1e-5, f"gradient rel-L2 {err:.2e} for request {r.rid}"


def _bench_sized_mix():
    """Requests of the bench mix's larger kinds: S in {32, 64}, P in {1024, 4096}, every window size."""
    specs = [("forward", 128, 32, 1024), ("forward", 64, 64, 4096), ("gradient", 256, 32, 4096),
             ("forward", 256, 32, 4096), ("gradient", 128, 64, 1024), ("forward", 64, 32, 1024),
             ("gradient", 64, 64, 4096), ("forward", 256, 64, 1024)]
    return [serving.Request(i, k, n, s, p, seed=40 + i, arrival=0.001 * i) for i, (k, n, s, p) in enumerate(specs)]


@pytest.mark.gpu
def test_results_same_alone_and_in_mix_at_bench_sizes():
    """The bench loop's own chunking (8192 / 4096 / 512 positions per step at n = 64 / 128 / 256) over
    S = 32 / 64 and P = 1024 / 4096 requests: every request's patterns are bitwise those of the request run
    alone; losses to 1e-9 and gradients to rel-L2 1e-5."""
    loop = serving.RequestLoop(keep_inputs=True)
    mixed = _bench_sized_mix()
    stats = loop.run(mixed, realtime=False)
    assert stats["requests"] == len(mixed)
    for r in mixed:
        print(f"request {r.rid} {r.kind} n={r.n} S={r.slices} P={r.positions}", flush=True)
        _assert_same_as_alone(r)
        r.inputs = r.patterns = r.grad_slab = r.grad_probe_k = None


@pytest.mark.gpu
def test_admission_respects_hbm_budget():
    reqs = _mix()
    budget = max(serving.request_bytes(r) for r in reqs) + 1
    loop = serving.RequestLoop(hbm_budget_bytes=budget)
    stats = loop.run(reqs, realtime=False)
    assert stats["requests"] == len(reqs)
    # requests were admitted one at a time when the budget only fits one of the largest
    admits = sorted(r.t_admit for r in reqs)
    assert admits[-1] > admits[0]


@pytest.mark.gpu
def test_adjoint_batched_matches_per_request():
    """pty_adjoint_batched (descriptors addressing two requests' slabs, atomic accumulation) agrees with the
    per-request gathered adjoint."""
    from ptyrodactyl_amd import engine
    n, s_count, k = 64, 5, 4
    loop = serving.RequestLoop()
    reqs = [serving.Request(0, "gradient", n, s_count, 24, seed=3), serving.Request(1, "gradient", n, s_count, 40, seed=4)]
    acts = [loop._admit(r, 0.0) for r in reqs]
    torch.cuda.synchronize()
    items = [(a, 0, a.req.positions) for a in acts]
    total = sum(a.req.positions for a in acts)
    nc = (s_count - 1) // k
    ckpt = torch.empty((total, nc, n, n), dtype=torch.complex64, device="cuda")
    engine.forward_batched(n, s_count, loop._descs(items, True, [a.req.patterns for a in acts]), loop.base, ckpt, k)
    gis = [serving.loss_and_grad_intensity(a.req.patterns, a.inp.measured)[1] for a in acts]
    engine.adjoint_batched(n, s_count, loop._descs(items, True, gis), loop.base, ckpt, k)
    one = torch.ones(1, dtype=torch.float32, device="cuda")
    k0 = 0
    for a, gi in zip(acts, gis):
        inp, p = a.inp, a.req.positions
        ref_slab, ref_pk = engine.multislice_adjoint(inp.slab, inp.probe_k[None], one, inp.origins, inp.shifts,
                                                     inp.prop[:n].contiguous(), gi, ckpt[k0:k0 + p], k)
        k0 += p
        for got, ref in ((a.req.grad_slab, ref_slab), (a.req.grad_probe_k, ref_pk[0])):
            err = float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref))
            print(f"request {a.req.rid}: batched vs per-request adjoint rel-L2 {err:.2e}", flush=True)
            assert err < 1e-5


def _assert_same_as_alone(r):
    alone = serving.PlainExecutor().run_one(serving.Request(r.rid, r.kind, r.n, r.slices, r.positions, r.seed),
                                             r.inputs)
    assert torch.equal(r.patterns, alone.patterns), f"patterns differ for request {r.rid}"
    if r.kind == "gradient":
        assert r.loss == pytest.approx(alone.loss, rel=1e-9)
        for got, ref in ((r.grad_slab, alone.grad_slab), (r.grad_probe_k, alone.grad_probe_k)):
            err = float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref))
            assert err < 1e-5, f"gradient rel-L2 {err:.2e} for request {r.rid}"


@pytest.mark.gpu
def test_unbuildable_request_fails_alone():
    """A request the loop cannot build (unsupported window) fails with its error; the rest complete."""
    reqs = _mix()[:5]
    bad = serving.Request(99, "forward", 48, 8, 16, seed=1, arrival=0.0015)
    loop = serving.RequestLoop(hbm_budget_bytes=256 << 20, chunk_positions={64: 128, 128: 48, 256: 16},
                               keep_inputs=True)
    stats = loop.run(reqs + [bad], realtime=False)
    assert stats["requests"] == 5 and stats.get("failed") == 1
    assert bad.status == "failed" and "48" in bad.error
    for r in reqs:
        assert r.status == "done"
        _assert_same_as_alone(r)


class _FaultyLoop(serving.RequestLoop):
    """Raises from the host whenever a launch would include request 3 (a forced launch error)."""

    def _step_items(self, key, items, completed):
        if any(a.req.rid == 3 for a, _, _ in items):
            # the other chunks of the batch are enqueued first, as a real mid-batch failure would leave them
            super()._step_items(key, [it for it in items if it[0].req.rid != 3], completed)
            raise ValueError("forced launch error for request 3")
        super()._step_items(key, items, completed)


@pytest.mark.gpu
def test_failing_step_is_isolated_without_double_counting():
    reqs = _mix()
    loop = _FaultyLoop(hbm_budget_bytes=256 << 20, chunk_positions={64: 128, 128: 48, 256: 16}, keep_inputs=True)
    stats = loop.run(reqs, realtime=False)
    assert stats.get("failed") == 1 and stats["requests"] == len(reqs) - 1
    assert reqs[3].status == "failed" and "forced launch error" in reqs[3].error
    for r in reqs:
        if r.rid != 3:
            assert r.status == "done"
            _assert_same_as_alone(r)


@pytest.mark.gpu
def test_oversize_rejected_cancelled_dropped_and_metrics_recorded():
    reqs = _mix()[:4]
    big = serving.Request(50, "gradient", 256, 64, 4096, seed=3, arrival=0.0)
    doomed = serving.Request(51, "forward", 64, 8, 16, seed=4, arrival=0.002)
    seen = []
    loop = serving.RequestLoop(hbm_budget_bytes=max(serving.request_bytes(r) for r in reqs) * 4,
                               metrics_sink=seen.append)
    loop.cancel(51)
    stats = loop.run(reqs + [big, doomed], realtime=False)
    assert big.status == "rejected" and "budget" in big.error
    assert doomed.status == "cancelled"
    assert stats["requests"] == 4 and stats.get("rejected") == 1 and stats.get("cancelled") == 1
    assert seen and seen == loop.metrics and seen[-1]["active"] == 0 and seen[-1]["hbm_in_use"] == 0


@pytest.mark.gpu
def test_queued_steps_give_the_results_of_synchronised_steps():
    """Steps queued ahead of the GPU (max_steps_ahead 2, the default) against a loop that waits for every step
    (max_steps_ahead 0): every request's patterns are bitwise the same.  Requests retire, and free their HBM budget,
    at different loop iterations in the two runs, so a gradient request may be cut into different chunks: losses
    agree to 1e-9 and gradients to rel-L2 1e-5, the bounds of the alone-versus-mix test."""
    chunks = {64: 128, 128: 48, 256: 16}
    runs = []
    for ahead in (0, 2):
        reqs = _mix()
        loop = serving.RequestLoop(hbm_budget_bytes=256 << 20, chunk_positions=chunks, max_steps_ahead=ahead)
        stats = loop.run(reqs, realtime=False)
        assert stats["requests"] == len(reqs)
        assert all(r.status == "done" for r in reqs)
        runs.append(reqs)
    for a, b in zip(*runs):
        print(f"request {a.rid} {a.kind} n={a.n} S={a.slices} P={a.positions}", flush=True)
        assert torch.equal(a.patterns, b.patterns), f"patterns differ for request {a.rid}"
        if a.kind == "gradient":
            assert a.loss == pytest.approx(b.loss, rel=1e-9)
            for got, ref in ((a.grad_slab, b.grad_slab), (a.grad_probe_k, b.grad_probe_k)):
                err = float(torch.linalg.norm(got - ref) / torch.linalg.norm(ref))
                assert err < 1e-5, f"gradient rel-L2 {err:.2e} for request {a.rid}"
