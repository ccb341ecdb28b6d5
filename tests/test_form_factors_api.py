"""Host-side form-factor API: coefficient tables in the reference's CSV format, carriers, validation, and the
closed-form form factors (torch) against the oracle.  No GPU needed."""

import json
from pathlib import Path

import numpy as np
import pytest
import torch

import ptyrodactyl_amd as pa
from ptyrodactyl_amd import form_factors as ff
from ptyrodactyl_amd.types import RuntimeCheckError
from oracle import potential_oracle as po

ROWS = json.loads((Path(__file__).parent / "golden" / "form_factor_rows.json").read_text())
LOB = {int(k): np.asarray(v) for k, v in ROWS["lobato"].items()}
KIR = {int(k): np.asarray(v) for k, v in ROWS["kirkland"].items()}


@pytest.fixture
def fixture_tables():
    saved = dict(ff._TABLES)
    ff.set_form_factor_table("lobato", LOB)
    ff.set_form_factor_table("kirkland", KIR)
    yield
    ff._TABLES.update(saved)


def _synthetic_tables(tmp_path: Path):
    """Both CSV files in the reference's format: the fixture rows for H, C, Si, Au and synthetic positive rows
    for the other elements."""
    rng = np.random.default_rng(0)
    lob = rng.uniform(0.1, 2.0, (103, 10))
    kir = rng.uniform(0.1, 2.0, (103, 12))
    for z in LOB:
        lob[z - 1] = LOB[z]
        kir[z - 1] = KIR[z]
    grouped = np.concatenate([lob[:, 0::2], lob[:, 1::2]], axis=1)
    lines = ["Z,element,a1,a2,a3,a4,a5,b1,b2,b3,b4,b5"]
    lines += [f"{z},X{z}," + ",".join(repr(float(v)) for v in grouped[z - 1]) for z in range(1, 104)]
    (tmp_path / "Lobato_van_Dyck.csv").write_text("\n".join(lines) + "\n")
    np.savetxt(tmp_path / "Kirkland_Potentials.csv", kir, delimiter=",", fmt="%.17e")
    return lob, kir


def test_tables_load_from_a_directory_in_the_reference_format(tmp_path, monkeypatch):
    saved = dict(ff._TABLES)
    try:
        lob, kir = _synthetic_tables(tmp_path)
        ff.load_form_factor_tables(tmp_path)
        np.testing.assert_array_equal(pa.lobato_potentials().numpy(), lob)
        np.testing.assert_array_equal(pa.kirkland_potentials().numpy(), kir)
        np.testing.assert_array_equal(pa.lobato_potentials()[13].numpy(), LOB[14])
        ff._TABLES.update({"lobato": None, "kirkland": None})
        monkeypatch.setenv("PTYRODACTYL_FORM_FACTOR_DIR", str(tmp_path))
        np.testing.assert_array_equal(pa.kirkland_potentials().numpy(), kir)
    finally:
        ff._TABLES.update(saved)


def test_table_validation_messages(tmp_path):
    saved = dict(ff._TABLES)
    try:
        with pytest.raises(ValueError, match=r"Expected Kirkland CSV shape \(103, 12\)"):
            ff.set_form_factor_table("kirkland", np.ones((10, 12)))
        bad = np.ones((103, 10))
        bad[5, 3] = -1.0
        with pytest.raises(ValueError, match="Lobato scale coefficients must be positive"):
            ff.set_form_factor_table("lobato", bad)
        with pytest.raises(ValueError, match="parameterization must be 'lobato' or 'kirkland'"):
            ff.set_form_factor_table("peng", np.ones((103, 10)))
        with pytest.raises(FileNotFoundError):
            ff.load_form_factor_tables(tmp_path)
    finally:
        ff._TABLES.update(saved)


def test_missing_rows_and_tables_raise(fixture_tables, monkeypatch):
    with pytest.raises(RuntimeError, match="no lobato coefficients installed for atomic number 8"):
        ff.coefficient_row(8, "lobato")
    monkeypatch.setattr(ff, "_installed_luggage", lambda: None)
    monkeypatch.delenv("PTYRODACTYL_FORM_FACTOR_DIR", raising=False)
    monkeypatch.setitem(ff._TABLES, "lobato", None)
    with pytest.raises(RuntimeError, match="the lobato coefficient table is not loaded"):
        ff.coefficient_row(6, "lobato")


@pytest.mark.parametrize("bad, msg", [(True, "atom_no must be an integer between 1 and 103"),
                                      (0, "atom_no must be between 1 and 103"),
                                      (104, "atom_no must be between 1 and 103")])
def test_atomic_number_validation(fixture_tables, bad, msg):
    """``test_form_factors.py:159-200``."""
    with pytest.raises(ValueError, match=msg):
        pa.atomic_form_factor(bad, torch.zeros(3, dtype=torch.float64))


def test_parameter_carriers_validate():
    with pytest.raises(ValueError, match=r"amplitudes must have shape \(5,\)"):
        pa.create_lobato_parameters(np.ones(4), np.ones(5))
    with pytest.raises(RuntimeCheckError, match="scales must be strictly positive"):
        pa.create_lobato_parameters(np.ones(5), -np.ones(5))
    with pytest.raises(RuntimeCheckError, match="gaussian_amplitudes must contain only finite values"):
        pa.create_kirkland_parameters(np.ones(3), np.ones(3), [1.0, np.nan, 1.0], np.ones(3))
    with pytest.raises(ValueError, match="parameterization must be 'lobato' or 'kirkland'"):
        pa.projected_atom_potential(6, torch.ones(2, dtype=torch.float64), parameterization="doyle")


@pytest.mark.parametrize("z", [1, 6, 14, 79])
def test_form_factors_match_the_oracle(fixture_tables, z):
    q = 2.0 * np.pi * np.linspace(0.0, 6.0, 97)
    np.testing.assert_allclose(pa.atomic_form_factor(z, torch.as_tensor(q)).numpy(), po.lobato_form_factor(LOB[z], q),
                               rtol=1e-14)
    np.testing.assert_allclose(pa.atomic_form_factor(z, torch.as_tensor(q), parameterization="kirkland").numpy(),
                               po.kirkland_form_factor(KIR[z], q), rtol=1e-14)


def test_hydrogen_anchor_and_bandlimited_peaks(fixture_tables):
    """``test_form_factors.py:42-81``."""
    np.testing.assert_allclose(float(pa.atomic_form_factor(1, torch.tensor(0.0, dtype=torch.float64))), pa.form_factors.A_BOHR,
                               rtol=1e-9)
    for z, g, volts in [(6, 2.0, 417.0), (14, 3.0, 1513.0), (79, 4.0, 10125.0)]:
        assert abs(float(pa.lobato_bandlimited_peak(z, g)) - volts) <= 0.51
    with pytest.raises(ValueError, match="g_max must be positive and finite"):
        pa.lobato_bandlimited_peak(6, 0.0)
    with pytest.raises(RuntimeCheckError, match="g_max must be positive and finite"):
        pa.lobato_bandlimited_peak(6, torch.tensor(float("inf")))


def test_crystal_data_factory():
    """``crystal_types.py:244-400``."""
    c = pa.create_crystal_data([[0.0, 0.0, 0.0]], [6])
    assert torch.equal(c.lattice, torch.eye(3, dtype=torch.float64)) and c.atomic_numbers.dtype == torch.int32
    with pytest.raises(ValueError, match=r"positions must have shape \(N, 3\)"):
        pa.create_crystal_data([[0.0, 0.0]], [6])
    with pytest.raises(ValueError, match=r"atomic_numbers must have shape \(N,\)"):
        pa.create_crystal_data([[0.0, 0.0, 0.0]], [6, 8])
    with pytest.raises(RuntimeCheckError, match="positions contain non-finite values"):
        pa.create_crystal_data([[np.nan, 0.0, 0.0]], [6])
    with pytest.raises(RuntimeCheckError, match="atomic_numbers must be non-negative"):
        pa.create_crystal_data([[0.0, 0.0, 0.0]], [-1])
    with pytest.raises(ValueError, match=r"lattice must have shape \(3, 3\)"):
        pa.create_crystal_data([[0.0, 0.0, 0.0]], [6], lattice=np.eye(2))


@pytest.mark.parametrize("seed, pixel, pad, grid, reps", [(0, 0.2, 2.0, None, (1, 1, 1)),
                                                          (2, 0.25, 1.0, (40, 36, 3), (2, 2, 1)),
                                                          (3, 0.3, 0.6, None, (2, 1, 2))])
def test_slice_plan_with_the_fourier_sum_reproduces_the_oracle(seed, pixel, pad, grid, reps):
    """Host half of ``crystal_potential_slices`` on the CPU: the plan's grid, atom order, shifts and segments,
    evaluated with the Fourier-space sum the GPU kernel computes (numpy here), equal the oracle's per-atom
    Fourier shifts."""
    from ptyrodactyl_amd.atom_potentials import plan_slices
    rng = np.random.default_rng(seed)
    pos = np.column_stack([rng.uniform(0, 3, 9), rng.uniform(0, 3, 9), rng.uniform(0, 2.6, 9)])
    nums = rng.choice([1, 6, 14, 79], 9)
    lattice = np.asarray([[3.1, 0.0, 0.0], [0.4, 2.9, 0.0], [0.0, 0.0, 1.7]])
    plan = plan_slices(pa.create_crystal_data(pos, nums, lattice=lattice), pixel, 1.0, reps, pad, grid, "cpu")
    h, w, t = plan.height, plan.width, plan.elements.numel()
    kern_k = [np.fft.fft2(po.single_atom_potential("lobato", LOB[int(z)], pixel, (h, w), supersampling=2))
              for z in plan.elements.tolist()]
    ky = np.fft.fftfreq(h).reshape(-1, 1)
    kx = np.fft.fftfreq(w).reshape(1, -1)
    sh, seg = plan.shifts.numpy(), plan.segments.numpy()
    got = np.zeros((h, w, plan.num_slices))
    for s in range(plan.num_slices):
        f = np.zeros((h, w), complex)
        for e in range(t):
            a0, a1 = seg[s * t + e], seg[s * t + e + 1]
            sf = sum(np.exp(2j * np.pi * (kx * sh[a, 0] + ky * sh[a, 1])) for a in range(a0, a1))
            f += kern_k[e] * sf
        got[:, :, s] = np.real(np.fft.ifft2(f))
    c = plan.crop
    got = got[c:h - c, c:w - c]
    want, (x_min, y_min, ww, hh) = po.crystal_potential_slices("lobato", LOB, pos, nums, lattice, pixel, 1.0, reps,
                                                               pad, 2, grid)
    assert (hh, ww) == (h, w) and np.isclose(x_min, plan.x_min) and np.isclose(y_min, plan.y_min)
    assert got.shape == want.shape
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 2e-7
