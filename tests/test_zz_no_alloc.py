"""pty_forward_noalloc (the entry point that allocates nothing) against pty_forward.

Both run the same kernels; the no-allocation variant keeps its streaming waves in the library's
static device scratch instead of a caller workspace, in batches that fit it.  Batching does not
change any wave's arithmetic, so the patterns and amplitudes must be bitwise equal.  Two calls in
a row, the second with a smaller input, exercise the shared scratch that keeps the first call's
contents.  Oversized requests must fail with an error before any launch.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

DEV = "cuda"


def _operands(n, s_count, p_count, modes, seed):
    from ptyrodactyl_amd import physics
    rng = np.random.default_rng(seed)
    hs, ws = n + 40, n + 24
    slab = (rng.uniform(0.95, 1.0, (s_count, hs, ws)) * np.exp(1j * rng.uniform(0, 0.4, (s_count, hs, ws))))
    probes = (rng.standard_normal((modes, n, n)) + 1j * rng.standard_normal((modes, n, n))) / n
    origins = np.stack([rng.integers(0, hs, p_count), rng.integers(0, ws, p_count)], 1).astype(np.int32)
    shifts = rng.uniform(-3.0, 3.0, (p_count, 2))
    weights = rng.uniform(0.2, 1.0, modes)
    return (torch.tensor(slab.astype(np.complex64), device=DEV),
            torch.fft.fft2(torch.tensor(probes, device=DEV)).to(torch.complex64).contiguous(),
            torch.tensor(weights / weights.sum(), dtype=torch.float32, device=DEV),
            torch.tensor(origins, device=DEV), torch.tensor(shifts, device=DEV),
            physics.propagator_factors(n, 1.5, 300.0, 0.1, device=DEV))


def _call_noalloc(lib, ops, want_amplitude):
    from ptyrodactyl_amd import _native
    slab, pk, wt, ot, st, prop = ops
    s_count, hs, ws = slab.shape
    m_count, n, _ = pk.shape
    p_count = ot.shape[0]
    pat = torch.full((p_count, n, n), float("nan"), dtype=torch.float32, device=DEV)
    amp = torch.full((p_count, m_count, n, n), float("nan"), dtype=torch.complex64, device=DEV) \
        if want_amplitude else None
    status = lib.pty_forward_noalloc(n, s_count, p_count, m_count, _native.ptr(slab), hs, ws, _native.ptr(pk),
                                     _native.ptr(wt), _native.ptr(ot), _native.ptr(st), _native.ptr(prop),
                                     _native.ptr(prop), _native.ptr(pat), _native.ptr(amp) if amp is not None else 0,
                                     _native.stream_handle(torch.device(DEV)))
    _native.check(status, "pty_forward_noalloc")
    return pat, amp


def _call_with_workspace(lib, ops, want_amplitude):
    from ptyrodactyl_amd import _native
    slab, pk, wt, ot, st, prop = ops
    s_count, hs, ws = slab.shape
    m_count, n, _ = pk.shape
    p_count = ot.shape[0]
    nbytes = int(lib.pty_forward_workspace_bytes(n, p_count, m_count, 0))
    work = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=DEV)
    pat = torch.empty((p_count, n, n), dtype=torch.float32, device=DEV)
    amp = torch.empty((p_count, m_count, n, n), dtype=torch.complex64, device=DEV) if want_amplitude else None
    status = lib.pty_forward(n, s_count, p_count, m_count, _native.ptr(slab), hs, ws, _native.ptr(pk),
                             _native.ptr(wt), _native.ptr(ot), _native.ptr(st), _native.ptr(prop), _native.ptr(prop),
                             _native.ptr(pat), _native.ptr(amp) if amp is not None else 0, _native.ptr(work), nbytes,
                             _native.stream_handle(torch.device(DEV)))
    _native.check(status, "pty_forward")
    return pat, amp


@pytest.mark.parametrize("n,s_count,modes,p_first,p_second", [
    (256, 5, 1, 300, 7),     # 300 waves > the 128 that fit the scratch: three batches, then one
    (512, 3, 2, 40, 3),      # 2 modes per position, 16 positions per batch
    (128, 4, 1, 50, 9),      # LDS-resident: no scratch used
])
def test_noalloc_equals_workspace_forward_on_two_calls(n, s_count, modes, p_first, p_second):
    from ptyrodactyl_amd import _native
    lib = _native.load()
    for call, p_count in enumerate((p_first, p_second)):
        print(f"noalloc n={n} S={s_count} M={modes} P={p_count} (call {call + 1})", flush=True)
        ops = _operands(n, s_count, p_count, modes, seed=1000 * n + p_count)
        want_amp = call == 0
        pat_a, amp_a = _call_noalloc(lib, ops, want_amp)
        pat_b, amp_b = _call_with_workspace(lib, ops, want_amp)
        torch.cuda.synchronize()
        assert torch.isfinite(pat_a).all()
        assert torch.equal(pat_a, pat_b)
        if want_amp:
            assert torch.equal(amp_a, amp_b)


def test_noalloc_rejects_requests_beyond_the_scratch():
    from ptyrodactyl_amd import _native
    lib = _native.load()
    ops = _operands(256, 1, 1, 1, seed=5)
    slab, _, wt, ot, st, prop = ops
    # n = 4096: one wave is 128 MiB, above the 64 MiB scratch; rejected on the host, nothing launched
    big = torch.zeros((1, 4096, 4096), dtype=torch.complex64, device=DEV)
    pk = torch.zeros((1, 4096, 4096), dtype=torch.complex64, device=DEV)
    prop_big = torch.zeros(4096, dtype=torch.complex64, device=DEV)
    pat = torch.empty((1, 4096, 4096), dtype=torch.float32, device=DEV)
    status = lib.pty_forward_noalloc(4096, 1, 1, 1, _native.ptr(big), 4096, 4096, _native.ptr(pk), _native.ptr(wt),
                                     _native.ptr(ot), _native.ptr(st), _native.ptr(prop_big), _native.ptr(prop_big),
                                     _native.ptr(pat), 0, _native.stream_handle(torch.device(DEV)))
    assert status == _native.PTY_ERR_INVALID
    assert b"static scratch" in lib.pty_last_error()
