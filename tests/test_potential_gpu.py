"""GPU parity of the independent-atom potential builders (``pty_projected_potential``,
``pty_atom_potential_grid``, ``pty_potential_slices_fourier``) against ``oracle/potential_oracle.py`` and the
reference's captured values (``tests/test_ptyrodactyl/test_multislice/test_atom_potentials.py``).

Tolerances: projected potentials and single-atom grids are FP64 evaluations of the same formulas as the
oracle (rel 1e-12; only rounding order differs).  Slices: the oracle accumulates every atom in float32 as the
reference does, the GPU in float64, so they agree to float32 rounding (rel-L2 2e-7, max 2e-6 of the peak)."""

import json
from pathlib import Path

import numpy as np
import pytest
import torch

import ptyrodactyl_amd as pa
from ptyrodactyl_amd import form_factors as ff
from oracle import potential_oracle as po

pytestmark = pytest.mark.gpu

ROWS = json.loads((Path(__file__).parent / "golden" / "form_factor_rows.json").read_text())
LOB = {int(k): np.asarray(v) for k, v in ROWS["lobato"].items()}
KIR = {int(k): np.asarray(v) for k, v in ROWS["kirkland"].items()}
TABLES = {"lobato": LOB, "kirkland": KIR}


@pytest.fixture(autouse=True)
def fixture_tables():
    saved = dict(ff._TABLES)
    ff.set_form_factor_table("lobato", LOB)
    ff.set_form_factor_table("kirkland", KIR)
    yield
    ff._TABLES.update(saved)


def _rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("kind", ["lobato", "kirkland"])
@pytest.mark.parametrize("z", [1, 6, 14, 79])
def test_projected_potential_matches_oracle(kind, z):
    r = np.concatenate([[0.0], np.geomspace(1e-6, 25.0, 400)])
    got = pa.projected_atom_potential(z, torch.as_tensor(r, device="cuda"), parameterization=kind).cpu().numpy()
    want = po.projected(kind, TABLES[kind][z], r)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-300)
    if kind == "lobato":
        p = pa.create_lobato_parameters(LOB[z][0::2], LOB[z][1::2])
        np.testing.assert_allclose(pa.lobato_projected_potential(p, torch.as_tensor(r)).cpu().numpy(), want,
                                   rtol=1e-12)
    else:
        row = KIR[z]
        p = pa.create_kirkland_parameters(row[0:6:2], row[1:6:2], row[6::2], row[7::2])
        np.testing.assert_allclose(pa.kirkland_projected_potential(p, torch.as_tensor(r)).cpu().numpy(), want,
                                   rtol=1e-12)


@pytest.mark.parametrize("z, total, peak", [(6, 3304.630500664369, 599.0660669230273),
                                            (14, 7467.50633667953, 1309.8209272295387),
                                            (79, 16835.066119859865, 6296.83419348897)])
def test_kirkland_single_atom_matches_the_captured_reference(z, total, peak):
    """``test_atom_potentials.py:32-64`` (the reference asserts 2e-15 on its own evaluation)."""
    v = pa.single_atom_potential(z, 0.2, (16, 16), supersampling=2, parameterization="kirkland").cpu().numpy()
    np.testing.assert_allclose(v.sum(), total, rtol=1e-13)
    np.testing.assert_allclose(v.max(), peak, rtol=1e-13)


@pytest.mark.parametrize("kind, z, shape, ss, center", [("lobato", 14, (16, 16), 2, None),
                                                        ("lobato", 79, (33, 20), 3, (0.37, -0.21)),
                                                        ("kirkland", 6, (25, 31), 1, (1.0, 0.5)),
                                                        ("lobato", 1, (64, 64), 4, None)])
def test_single_atom_potential_matches_oracle(kind, z, shape, ss, center):
    ctr = None if center is None else torch.tensor(center, dtype=torch.float64)
    got = pa.single_atom_potential(z, 0.15, shape, ctr, ss, parameterization=kind).cpu().numpy()
    want = po.single_atom_potential(kind, TABLES[kind][z], 0.15, shape, (0.0, 0.0) if center is None else center,
                                    ss)
    np.testing.assert_allclose(got, want, rtol=1e-12)


@pytest.mark.parametrize("kind, z, shape, ss, center", [("lobato", 14, (16, 16), 2, (0.13, -0.17)),
                                                        ("lobato", 79, (33, 20), 3, (0.37, -0.21)),
                                                        ("kirkland", 6, (25, 31), 1, (1.0, 0.5)),
                                                        ("kirkland", 14, (64, 48), 4, (-0.4, 0.25))])
def test_single_atom_center_gradient_matches_oracle(kind, z, shape, ss, center):
    """Autograd through ``single_atom_potential`` in the centre (``pty_atom_potential_grid_center_grad``)
    against the oracle's chain rule: the same FP64 formulas, rel 1e-10."""
    print(f"centre gradient {kind} Z={z} shape={shape} ss={ss}", flush=True)
    rng = np.random.default_rng(z)
    weights = rng.standard_normal(shape)
    ctr = torch.tensor(center, dtype=torch.float64, device="cuda", requires_grad=True)
    v = pa.single_atom_potential(z, 0.15, shape, ctr, ss, parameterization=kind)
    (v * torch.as_tensor(weights, device="cuda")).sum().backward()
    want = po.single_atom_center_grad(kind, TABLES[kind][z], 0.15, shape, center, ss, weights)
    np.testing.assert_allclose(ctr.grad.cpu().numpy(), want, rtol=1e-10, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("kind", ["lobato", "kirkland"])
def test_single_atom_center_gradient_reference_check(kind):
    """``test_atom_potentials.py:91-151`` through the drop-in: Si, 0.2 A, 16 x 16, supersampling 2, centre
    (0.13, -0.17), loss mean(weights * V); the shift changes the potential, and grad . (0.6, -0.8) equals
    central differences of the GPU forward at rtol 2e-5, atol 1e-7."""
    axis = torch.linspace(-1.0, 1.0, 16, dtype=torch.float64, device="cuda")
    yw, xw = torch.meshgrid(axis, axis, indexing="ij")
    weights = 0.37 * xw - 0.61 * yw + 0.23 * xw * yw

    def loss(c):
        return torch.mean(weights * pa.single_atom_potential(14, 0.2, (16, 16), c, 2, parameterization=kind))

    center = torch.tensor([0.13, -0.17], dtype=torch.float64, device="cuda")
    centered = pa.single_atom_potential(14, 0.2, (16, 16), torch.zeros(2, dtype=torch.float64, device="cuda"), 2,
                                        parameterization=kind)
    shifted = pa.single_atom_potential(14, 0.2, (16, 16), center, 2, parameterization=kind)
    assert float(torch.linalg.norm(shifted - centered) / torch.linalg.norm(centered)) > 1e-3
    c = center.clone().requires_grad_(True)
    (grad,) = torch.autograd.grad(loss(c), c)
    assert torch.all(torch.isfinite(grad)) and float(torch.linalg.norm(grad)) > 1e-6
    d = torch.tensor([0.6, -0.8], dtype=torch.float64, device="cuda")
    step = 1e-5
    fd = (loss(center + step * d) - loss(center - step * d)) / (2.0 * step)
    np.testing.assert_allclose(float(grad @ d), float(fd), rtol=2e-5, atol=1e-7)


def test_center_gradient_entry_point_rejects_a_short_workspace():
    from ptyrodactyl_amd import _native
    lib = _native.load()
    rows = torch.zeros((1, 12), dtype=torch.float64, device="cuda")
    rows[0, :10] = torch.as_tensor(TABLES["lobato"][14], dtype=torch.float64)
    g = torch.ones((1, 8, 8), dtype=torch.float64, device="cuda")
    out = torch.empty((1, 2), dtype=torch.float64, device="cuda")
    need = int(lib.pty_atom_center_grad_workspace_bytes(1))
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    st = lib.pty_atom_potential_grid_center_grad(0, 1, _native.ptr(rows), 0.2, 8, 8, 1, None, _native.ptr(g),
                                                 _native.ptr(out), _native.ptr(ws), need - 1,
                                                 _native.stream_handle(torch.device("cuda", 0)))
    assert st != 0 and b"workspace" in lib.pty_last_error()


def test_default_parameterization_is_lobato():
    """``test_atom_potentials.py:715-761``."""
    a = pa.single_atom_potential(14, 0.2, (16, 16), supersampling=2)
    b = pa.single_atom_potential(14, 0.2, (16, 16), supersampling=2, parameterization="lobato")
    assert torch.equal(a, b)
    crystal = pa.create_crystal_data([[0.5, 0.5, 0.2], [1.0, 1.0, 0.7]], [6, 14], lattice=np.eye(3) * 2.0)
    s1 = pa.crystal_potential_slices(crystal, 0.25, padding=0.25, supersampling=1, grid_shape=(4, 4, 1))
    s2 = pa.crystal_potential_slices(crystal, 0.25, padding=0.25, supersampling=1, grid_shape=(4, 4, 1),
                                     parameterization="lobato")
    assert torch.equal(s1.slices, s2.slices) and tuple(s1.slices.shape) == (2, 2, 1)


@pytest.mark.parametrize("v", [0.0, 1.0, 2.0, 3.0, 0.5, 1.3, -0.7])
def test_bessel_kv_matches_oracle(v):
    x = np.concatenate([np.geomspace(1e-3, 2.0, 50), np.linspace(2.0001, 30.0, 50)])
    np.testing.assert_allclose(pa.bessel_kv(v, torch.as_tensor(x)).cpu().numpy(), po.bessel_kv(v, x), rtol=1e-12)
    assert abs(float(pa.bessel_kv(0.0, torch.tensor(0.5, dtype=torch.float64))) - 0.9244190) <= 1e-5


def _crystal_case(seed, n_atoms, span, elements, dz_span):
    rng = np.random.default_rng(seed)
    pos = np.column_stack([rng.uniform(0, span, n_atoms), rng.uniform(0, span, n_atoms),
                           rng.uniform(0, dz_span, n_atoms)])
    return pos, rng.choice(elements, n_atoms)


@pytest.mark.parametrize("kind, seed, n_atoms, span, pixel, pad, ss, grid, reps", [
    ("lobato", 0, 12, 4.0, 0.2, 2.0, 2, None, (1, 1, 1)),
    ("kirkland", 1, 20, 5.0, 0.15, 1.0, 3, None, (1, 1, 1)),
    ("lobato", 2, 6, 3.0, 0.25, 1.0, 2, (40, 36, 3), (2, 2, 1)),
    ("kirkland", 3, 9, 2.5, 0.3, 0.6, 1, None, (2, 1, 2)),
])
def test_crystal_potential_slices_match_oracle(kind, seed, n_atoms, span, pixel, pad, ss, grid, reps):
    pos, nums = _crystal_case(seed, n_atoms, span, [1, 6, 14, 79], 2.6)
    lattice = np.asarray([[3.1, 0.0, 0.0], [0.4, 2.9, 0.0], [0.0, 0.0, 1.7]])
    crystal = pa.create_crystal_data(pos, nums, lattice=lattice)
    got = pa.crystal_potential_slices(crystal, pixel, 1.0, repeats=reps, padding=pad, supersampling=ss,
                                      grid_shape=grid, parameterization=kind)
    want, _ = po.crystal_potential_slices(kind, TABLES[kind], pos, nums, lattice, pixel, 1.0, reps, pad, ss, grid)
    g = got.slices.cpu().numpy()
    assert g.shape == want.shape and got.slice_thickness == 1.0 and got.calib == pixel
    print(f"{kind} seed {seed}: shape {g.shape} rel-L2 {_rel_l2(g, want):.2e}")
    assert _rel_l2(g, want) < 2e-7
    assert np.abs(g - want).max() < 2e-6 * np.abs(want).max()


@pytest.mark.parametrize("kind, seed, n_atoms, span, pixel, pad, ss, grid, reps", [
    ("lobato", 0, 12, 4.0, 0.2, 2.0, 2, None, (1, 1, 1)),
    ("kirkland", 1, 20, 5.0, 0.15, 1.0, 3, None, (1, 1, 1)),
    ("lobato", 2, 6, 3.0, 0.25, 1.0, 2, (40, 36, 3), (2, 2, 1)),
    ("kirkland", 3, 9, 2.5, 0.3, 0.6, 1, None, (2, 1, 2)),
    ("lobato", 4, 300, 12.0, 0.1, 1.0, 2, None, (1, 1, 1)),
])
def test_crystal_slices_position_gradient_matches_oracle(kind, seed, n_atoms, span, pixel, pad, ss, grid, reps):
    """Autograd through ``crystal_potential_slices`` in the atom positions (``pty_potential_slices_shift_grad``,
    grid origin through amin) against the oracle's gradient, which central differences pin: FP64 on both sides,
    rel 1e-9.  Repeats (2, 2, 1) and (2, 1, 2) put ties at the minimum (lattice copies), split evenly on both
    sides."""
    print(f"slice position gradient {kind} seed {seed} atoms {n_atoms} repeats {reps}", flush=True)
    pos, nums = _crystal_case(seed, n_atoms, span, [1, 6, 14, 79], 2.6)
    lattice = np.asarray([[3.1, 0.0, 0.0], [0.4, 2.9, 0.0], [0.0, 0.0, 1.7]])
    p = torch.tensor(pos, dtype=torch.float64, device="cuda", requires_grad=True)
    crystal = pa.create_crystal_data(p, nums, lattice=lattice)
    got = pa.crystal_potential_slices(crystal, pixel, 1.0, repeats=reps, padding=pad, supersampling=ss,
                                      grid_shape=grid, parameterization=kind)
    g = np.random.default_rng(seed + 100).standard_normal(tuple(got.slices.shape))
    (got.slices * torch.as_tensor(g, device="cuda")).sum().backward()
    want = po.crystal_potential_slices_position_grad(kind, TABLES[kind], pos, nums, lattice, pixel, g, 1.0, reps,
                                                     pad, ss, grid)
    err = _rel_l2(p.grad.cpu().numpy(), want)
    print(f"  rel-L2 {err:.2e}", flush=True)
    assert err < 1e-9
    assert torch.all(p.grad[:, 2] == 0)


def test_reference_slice_occupancy_and_atomic_number_ordering():
    """``test_atom_potentials.py:811-1010``: atoms at z = 0, 1.5, 3 fill slices 0, 1, 3 and leave 2 empty;
    heavier atoms give higher peaks."""
    c = pa.create_crystal_data([[0.0, 0.0, 0.0], [2.0, 2.0, 1.5], [1.0, 1.0, 3.0]], [1, 6, 14])
    s = pa.crystal_potential_slices(c, 0.2, 1.0, grid_shape=(50, 50, 4), parameterization="kirkland").slices
    sums = s.sum(dim=(0, 1)).cpu()
    assert bool((sums[[0, 1, 3]] > 0).all()) and abs(float(sums[2])) < 1e-9 * float(sums.abs().max())
    peaks = []
    for z in (1, 6, 14, 79):
        c = pa.create_crystal_data([[0.0, 0.0, 0.0]], [z])
        peaks.append(float(pa.crystal_potential_slices(c, 0.1, 1.0, grid_shape=(81, 81, 1),
                                                       parameterization="kirkland").slices.max()))
    assert peaks == sorted(peaks) and len(set(peaks)) == 4


def test_slab_scale_sum_and_integer_shift_properties():
    """At a 256^2 x 8-slice slab with ~400 atoms (oracle too slow here): with no padding every slice's sum equals
    the sum of its atoms' centred potentials, and atoms on integer pixels give the point-reflected rolls of the
    centred potential."""
    pix = 0.1
    rng = np.random.default_rng(11)
    ij = rng.integers(0, 256, (400, 2))
    z = rng.uniform(0.0, 7.99, 400)
    nums = rng.choice([6, 14, 79], 400)
    pos = np.column_stack([ij[:, 1] * pix, ij[:, 0] * pix, z])
    pos[0, :2] = 0.0                     # pins x_min = y_min = 0
    pos[1, :2] = 255 * pix
    c = pa.create_crystal_data(pos, nums)
    out = pa.crystal_potential_slices(c, pix, 1.0, padding=0.0, supersampling=2, grid_shape=(256, 256, 8),
                                      parameterization="lobato").slices
    kern = {int(e): pa.single_atom_potential(int(e), pix, (256, 256), supersampling=2) for e in (6, 14, 79)}
    sidx = np.floor((z - z.min()) / 1.0).astype(int)
    want = torch.zeros((256, 256, 8), dtype=torch.float64, device="cuda")
    ii = np.rint(pos[:, 1] / pix).astype(int)
    jj = np.rint(pos[:, 0] / pix).astype(int)
    for a in range(400):
        want[:, :, sidx[a]] += torch.roll(kern[int(nums[a])], (-(ii[a] - 128), -(jj[a] - 128)), dims=(0, 1))
    for s in range(8):
        tot = sum(float(kern[int(e)].sum()) for e in nums[sidx == s])
        np.testing.assert_allclose(float(out[:, :, s].sum()), tot, rtol=1e-11)
    err = float((out - want).abs().max() / want.abs().max())
    print(f"256^2 x 8 slab, 400 atoms: max |slices - rolls| / peak = {err:.2e}")
    assert err < 1e-12


def test_stem4d_gradient_wrt_atom_positions_matches_oracle_pipeline():
    """Atoms -> crystal_potential_slices -> stem_4d -> loss, differentiated end to end on the GPU (shift gradient
    of the slices, exp(i sigma V), the HIP adjoint) against central differences of the FP64 oracle pipeline
    (float64-accumulated slices, complex128 multislice).  Tolerance: rel-L2 1e-5 over all coordinates (the
    complex64 gradient bar; measured 5.3e-7)."""
    from oracle import ms_oracle as o
    n = 32
    pos0 = np.array([[3.1, 3.3, 0.2], [4.7, 3.9, 0.4], [3.8, 5.3, 1.3], [5.9, 5.1, 1.6]])
    nums = np.array([14, 6, 79, 14])
    lattice = np.eye(3) * 12.0
    kw = dict(repeats=(1, 1, 1), padding=0.5, supersampling=2, grid_shape=(36, 36, 2))
    dev = torch.device("cuda", 0)
    mic = pa.create_microscope_config(voltage_kv=200.0, aperture_mrad=25.0, defocus_ang=10.0, probe_shape=(n, n))
    probe = pa.make_probe(mic, pa.create_detector_config(real_space_calib_ang=0.25))
    beam = pa.create_probe_modes(probe[..., None].to(dev), torch.tensor([1.0]), 0.25)
    scan = np.array([[10.0, 12.0], [14.5, 9.5], [12.0, 16.5]])
    det = pa.create_detector_config(real_space_calib_ang=0.25, scan_positions_px=torch.tensor(scan))
    wts = np.random.default_rng(21).standard_normal((3, n, n))

    p = torch.tensor(pos0, dtype=torch.float64, device=dev, requires_grad=True)
    slices = pa.crystal_potential_slices(pa.create_crystal_data(p, nums, lattice=lattice), 0.25, 1.0,
                                         parameterization="lobato", **kw)
    assert tuple(slices.slices.shape) == (n, n, 2)
    out = pa.stem_4d(slices, beam, mic, det)
    loss = torch.sum(out.data.double() * torch.tensor(wts, device=dev))
    (g,) = torch.autograd.grad(loss, p)
    g = g.cpu().numpy()

    probe_np = probe.numpy()

    def oracle_loss(pos):
        pot, _ = po.crystal_potential_slices("lobato", LOB, pos, nums, lattice, 0.25, 1.0, accumulate=np.float64,
                                             **kw)
        pats = o.stem_4d(pot, 1.0, 0.25, probe_np[..., None], np.array([1.0]), 0.25, 200.0, scan, 0.25)["data"]
        return float(np.sum(pats * wts))

    h = 1e-4
    fd = np.zeros_like(pos0)
    for a in range(pos0.shape[0]):
        for c in range(2):
            dp = np.zeros_like(pos0)
            dp[a, c] = h
            fd[a, c] = (oracle_loss(pos0 + dp) - oracle_loss(pos0 - dp)) / (2 * h)
    err = _rel_l2(g, fd)
    print(f"d loss / d atom positions: GPU vs oracle FD rel-L2 {err:.2e}", flush=True)
    assert err < 1e-5
    assert np.all(g[:, 2] == 0.0)
