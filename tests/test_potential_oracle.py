"""Pins of the independent-atom potential oracle (``oracle/potential_oracle.py``).

Values from the reference's own tests (paths relative to the reference root):
``tests/test_ptyrodactyl/test_multislice/test_atom_potentials.py`` (captured
Kirkland C/Si/Au single-atom sums and maxima, K_v values, slicing cases) and
``test_form_factors.py`` (published band-limited Lobato peaks, f_H(0) = a_0,
parameterisation agreement).  The coefficient rows are the fixture
``tests/golden/form_factor_rows.json``.
"""

import json
from pathlib import Path

import numpy as np
import pytest
from scipy import integrate, special

from oracle import potential_oracle as po

ROWS = json.loads((Path(__file__).parent / "golden" / "form_factor_rows.json").read_text())
LOB = {int(k): np.asarray(v) for k, v in ROWS["lobato"].items()}
KIR = {int(k): np.asarray(v) for k, v in ROWS["kirkland"].items()}
TABLES = {"lobato": LOB, "kirkland": KIR}


def test_fixture_rows_match_the_reference_tests_anchors():
    """The rows the reference's tests hard-code (``test_form_factor_data.py:66-170``)."""
    assert LOB[1][0] == 0.00647384848835291 and LOB[1][9] == 2.76511897642927
    assert LOB[14][0] == 2.87189142611612 and LOB[79][9] == 0.00037927766747767
    assert KIR[1][0] == 0.0355221981 and KIR[79][11] == 1.3849048


@pytest.mark.parametrize("z, total, peak", [(6, 3304.630500664369, 599.0660669230273),
                                            (14, 7467.50633667953, 1309.8209272295387),
                                            (79, 16835.066119859865, 6296.83419348897)])
def test_kirkland_single_atom_matches_the_captured_reference(z, total, peak):
    """``test_atom_potentials.py:32-64``: 16x16 grid, 0.2 A pixels, supersampling 2."""
    v = po.single_atom_potential("kirkland", KIR[z], 0.2, (16, 16), supersampling=2)
    np.testing.assert_allclose(v.sum(), total, rtol=2e-15)
    np.testing.assert_allclose(v.max(), peak, rtol=2e-15)


@pytest.mark.parametrize("z, g, volts", [(6, 2.0, 417.0), (6, 3.0, 712.0), (6, 4.0, 1028.0), (14, 2.0, 823.0),
                                         (14, 3.0, 1513.0), (14, 4.0, 2238.0), (79, 2.0, 3327.0),
                                         (79, 3.0, 6534.0), (79, 4.0, 10125.0)])
def test_bandlimited_peaks_match_the_published_table(z, g, volts):
    assert abs(po.lobato_bandlimited_peak(LOB[z], g) - volts) <= 0.51


def test_hydrogen_forward_form_factor_is_the_bohr_radius():
    np.testing.assert_allclose(po.lobato_form_factor(LOB[1], np.asarray(0.0)), po.A_BOHR, rtol=1e-9)


@pytest.mark.parametrize("z", [1, 6, 14, 79])
def test_lobato_bethe_normalisation(z):
    a, b = po.lobato_split(LOB[z])
    np.testing.assert_allclose(np.sum(a / b), z / (2.0 * np.pi ** 2 * po.A_BOHR), rtol=1e-8)


def test_abramowitz_stegun_branches_within_published_error_bounds():
    """A&S 9.8.5-9.8.8 error bounds (on the normalised forms) against scipy's K0, K1."""
    xs = np.linspace(1e-3, 2.0, 2001)
    e5 = np.abs(po.k0_small(xs) - special.k0(xs))
    e7 = np.abs(xs * po.k1_small(xs) - xs * special.k1(xs))
    xl = np.linspace(2.0, 60.0, 2001)
    e6 = np.abs(np.sqrt(xl) * np.exp(xl) * (po.k0_large(xl) - special.k0(xl)))
    e8 = np.abs(np.sqrt(xl) * np.exp(xl) * (po.k1_large(xl) - special.k1(xl)))
    assert e5.max() < 1e-8 and e7.max() < 8e-9 and e6.max() < 1.9e-7 and e8.max() < 2.2e-7
    assert e5.max() > 1e-10 and e6.max() > 1e-9  # the approximations, not exact K's
    np.testing.assert_allclose(po.i0(xs), special.i0(xs), rtol=1e-15)
    np.testing.assert_allclose(po.i1(xs), special.i1(xs), rtol=1e-15)


@pytest.mark.parametrize("x, want, tol", [(0.1, 2.4270690, 1e-5), (0.2, 1.7527038, 1e-5), (0.5, 0.9244190, 1e-5),
                                          (1.0, 0.4210244, 1e-5), (1.5, 0.2138056, 1e-2), (2.0, 0.1138938, 5e-3),
                                          (3.0, 0.0347395, 1e-3), (5.0, 0.0036911, 1e-3), (10.0, 0.0000778, 1e-4)])
def test_k0_reference_values(x, want, tol):
    """``test_atom_potentials.py:161-185``."""
    assert abs(float(po.bessel_kv(0.0, np.asarray(x))) - want) <= tol


@pytest.mark.parametrize("v, x, want", [(0.5, 0.5, 1.0750476), (0.5, 1.0, 0.4610685), (0.5, 2.0, 0.1199377),
                                        (0.5, 5.0, 0.0037766), (1.0, 0.5, 1.6564411), (1.0, 1.0, 0.6019072),
                                        (1.0, 2.0, 0.1398659), (1.0, 5.0, 0.0040446), (2.0, 1.0, 1.6248389),
                                        (2.0, 2.0, 0.2537598), (2.0, 5.0, 0.0053089)])
def test_kv_reference_values(v, x, want):
    """``test_atom_potentials.py:322-350``."""
    assert abs(float(po.bessel_kv(v, np.asarray(x))) - want) <= 1e-5


def test_kv_branches_against_scipy():
    xs = np.linspace(0.05, 2.0, 40)
    np.testing.assert_allclose(po.bessel_kv(1.3, xs), special.kv(1.3, xs), rtol=1e-12)   # reflection + I_v series
    np.testing.assert_allclose(po.bessel_kv(-0.7, xs), special.kv(-0.7, xs), rtol=1e-12)
    np.testing.assert_allclose(po.bessel_kv(3.0, xs), special.kv(3.0, xs), rtol=1e-6)    # recurrence from A&S
    xl = np.linspace(8.0, 40.0, 20)
    np.testing.assert_allclose(po.bessel_kv(1.3, xl), special.kv(1.3, xl), rtol=1e-5)    # 5-term asymptotic


def _hankel(v_of_r, g):
    """2 pi int_0^inf v(r) J0(2 pi g r) r dr (the 2-D Fourier transform of a radial function)."""
    f = lambda r: v_of_r(np.asarray([r]))[0] * special.j0(2.0 * np.pi * g * r) * r  # noqa: E731
    pieces = [0.0, 1e-4, 0.05, 0.3, 1.0, 3.0, 8.0, 30.0]
    return 2.0 * np.pi * sum(integrate.quad(f, a, b, limit=400, epsabs=0.0, epsrel=1e-11)[0]
                             for a, b in zip(pieces[:-1], pieces[1:]))


@pytest.mark.parametrize("kind, z", [("lobato", 6), ("lobato", 79), ("kirkland", 14), ("kirkland", 79)])
def test_projected_potentials_are_the_fourier_transform_of_the_form_factor(kind, z):
    """v_z(r) = C int f(g) exp(2 pi i g.r) d^2g, C = 47.87801 V A^2 (Lobato) or 2 pi 0.5292 14.4 (Kirkland);
    exact K0/K1 in place of the A&S evaluation, so the check is on the formula's terms and constants."""
    if kind == "lobato":
        row, c = LOB[z], po.MOTT_BETHE
        v = lambda r: po.lobato_projected_potential(row, r, special.k0, special.k1)  # noqa: E731
    else:
        row, c = KIR[z], 2.0 * np.pi * po.KIRKLAND_A0 * po.KIRKLAND_EK
        v = lambda r: po.kirkland_projected_potential(row, r, special.k0)  # noqa: E731
    for g in (0.0, 0.3, 1.0, 2.0):
        got = _hankel(v, g)
        want = c * po.form_factor(kind, row, np.asarray(2.0 * np.pi * g))
        np.testing.assert_allclose(got, want, rtol=2e-7)


@pytest.mark.parametrize("z", [6, 14, 79])
def test_parameterisations_agree(z):
    """``test_form_factors.py:317-360``."""
    q = 2.0 * np.pi * np.linspace(0.0, 4.0, 201)
    np.testing.assert_allclose(po.kirkland_form_factor(KIR[z], q), po.lobato_form_factor(LOB[z], q), rtol=6e-3)
    r = np.geomspace(0.05, 1.5, 201)
    np.testing.assert_allclose(po.projected("kirkland", KIR[z], r), po.projected("lobato", LOB[z], r), rtol=3e-2)


def test_single_atom_potential_is_the_block_average_of_the_fine_grid():
    ss, h, w, p = 3, 5, 6, 0.3
    v = po.single_atom_potential("lobato", LOB[14], p, (h, w), center=(0.1, -0.2), supersampling=ss)
    for i in range(h):
        for j in range(w):
            acc = 0.0
            for a in range(ss):
                for b in range(ss):
                    y = (i * ss + a - (h * ss) // 2) * (p / ss)
                    x = (j * ss + b - (w * ss) // 2) * (p / ss)
                    acc += po.projected("lobato", LOB[14], np.asarray([np.sqrt((x - 0.1) ** 2 + (y + 0.2) ** 2
                                                                               + 1e-10)]))[0]
            np.testing.assert_allclose(v[i, j], acc / ss ** 2, rtol=1e-14)


@pytest.mark.parametrize("coords, numbers, dz, want_slices, want_order", [
    ([[1.0, 2.0, 0.0], [3.0, 4.0, 1.0], [5.0, 6.0, 2.0]], [6, 8, 14], 1.0, [0, 1, 2], [6, 8, 14]),
    ([[1, 1, .1], [2, 2, .2], [3, 3, 1.5], [4, 4, .3], [5, 5, 1.8]], [1, 2, 3, 4, 5], 1.0, [0, 0, 0, 1, 1],
     [1, 2, 4, 3, 5]),
    ([[1, 1, 0], [2, 2, .1], [3, 3, 2.5], [4, 4, 2.6], [5, 5, 5]], [79] * 5, 1.0, [0, 0, 2, 2, 5], [79] * 5),
    ([[1.5, 2.5, 3.5]], [26], 1.0, [0], [26]),
])
def test_slice_atoms_reference_cases(coords, numbers, dz, want_slices, want_order):
    """``test_atom_potentials.py:385-527``."""
    x, y, s, num = po.slice_atoms(np.asarray(coords, np.float64), np.asarray(numbers), dz)
    assert s.tolist() == want_slices and num.tolist() == want_order


def test_integer_pixel_atom_is_the_point_reflected_roll_of_the_centred_potential():
    """An atom at integer pixel p lands at 2c - p (c = grid centre): the reference's +2 pi i phase shifts the
    centred potential by -(p - c)."""
    pix, pad = 0.25, 1.0
    pos = np.asarray([[1.0, 2.0, 0.0], [3.0, 1.5, 0.2]])
    out, (x_min, y_min, w, h) = po.crystal_potential_slices("lobato", LOB, pos, np.asarray([14, 14]), np.eye(3),
                                                            pix, padding=pad, supersampling=2,
                                                            grid_shape=(24, 20, 1))
    kern = po.single_atom_potential("lobato", LOB[14], pix, (h, w), supersampling=2)
    want = np.zeros((h, w))
    for x, y in pos[:, :2]:
        px, py = int(round((x - x_min) / pix)), int(round((y - y_min) / pix))
        want += np.roll(kern, (-(py - h // 2), -(px - w // 2)), axis=(0, 1))
    crop = round(pad / pix)
    np.testing.assert_allclose(out[..., 0], want[crop:h - crop, crop:w - crop], rtol=2e-6, atol=1e-3)


def test_slices_keep_every_atoms_integral_and_sum_by_linearity():
    """Sum over the uncropped grid = sum of the atoms' centred potentials; the per-atom shifts equal one
    structure-factor sum per element (the form the GPU builder evaluates)."""
    rng = np.random.default_rng(5)
    pos = np.column_stack([rng.uniform(0, 3, 7), rng.uniform(0, 3, 7), rng.uniform(0, 1.9, 7)])
    num = np.asarray([6, 14, 6, 79, 14, 6, 79])
    out, (x_min, y_min, w, h) = po.crystal_potential_slices("kirkland", KIR, pos, num, np.eye(3), 0.2,
                                                            padding=0.0, supersampling=2, grid_shape=(18, 16, 2))
    kern = {z: po.single_atom_potential("kirkland", KIR[z], 0.2, (h, w), supersampling=2) for z in (6, 14, 79)}
    _, _, sidx, snum = po.slice_atoms(pos, num, 1.0)
    for s in range(2):
        want = sum(kern[int(z)].sum() for z in snum[sidx == s])
        np.testing.assert_allclose(out[..., s].sum(), want, rtol=1e-6)
    ky = np.fft.fftfreq(h).reshape(-1, 1)
    kx = np.fft.fftfreq(w).reshape(1, -1)
    x, y, sidx, snum = po.slice_atoms(pos, num, 1.0)
    for s in range(2):
        tot = np.zeros((h, w), complex)
        for z in (6, 14, 79):
            sel = (sidx == s) & (snum == z)
            sf = sum(np.exp(2j * np.pi * (kx * ((xa - x_min) / 0.2 - w / 2) + ky * ((ya - y_min) / 0.2 - h / 2)))
                     for xa, ya in zip(x[sel], y[sel]))
            tot += np.fft.fft2(kern[z]) * sf
        np.testing.assert_allclose(out[..., s], np.real(np.fft.ifft2(tot)), rtol=1e-5, atol=1e-3)


def test_repeats_equal_the_explicitly_listed_supercell():
    lat = np.asarray([[2.0, 0.0, 0.0], [0.3, 2.5, 0.0], [0.0, 0.0, 1.5]])
    base = np.asarray([[0.1, 0.2, 0.0], [1.0, 1.1, 0.7]])
    nums = np.asarray([6, 14])
    p1, n1 = po.tile_repeats(base, nums, lat, (2, 3, 2))
    listed = [base[k] + np.asarray([i, j, l]) @ lat for i in range(2) for j in range(3) for l in range(2)
              for k in range(2)]
    np.testing.assert_allclose(p1, np.asarray(listed), rtol=0, atol=1e-15)
    assert n1.tolist() == [6, 14] * 12


@pytest.mark.parametrize("kind", ["lobato", "kirkland"])
@pytest.mark.parametrize("z", [1, 14, 79])
def test_projected_dr_matches_central_differences(kind, z):
    row = TABLES[kind][z]
    r = np.geomspace(1e-3, 12.0, 300)
    scale = 2.0 * np.pi / np.sqrt(row[1:10:2]) if kind == "lobato" else 2.0 * np.pi * np.sqrt(row[1:6:2])
    # keep clear of the A&S branch point x = 2, where the two approximations differ by ~1e-8
    r = r[np.abs(r[:, None] * scale[None] - 2.0).min(axis=1) > 1e-3]
    h = 1e-6 * r
    fd = (po.projected(kind, row, r + h) - po.projected(kind, row, r - h)) / (2 * h)
    got = po.projected_dr(kind, row, r)
    np.testing.assert_allclose(got, fd, rtol=2e-6, atol=1e-9 * np.abs(fd).max())
    assert np.all(got < 0)  # the screened potential falls off monotonically


@pytest.mark.parametrize("kind", ["lobato", "kirkland"])
def test_single_atom_center_grad_matches_the_reference_check(kind):
    """The reference's case (Si, 0.2 A, 16 x 16, supersampling 2, centre (0.13, -0.17), loss mean(weights * V))
    along its direction (0.6, -0.8): gradient . direction against central differences at rtol 2e-5."""
    row = TABLES[kind][14]
    axis = np.linspace(-1.0, 1.0, 16)
    yw, xw = np.meshgrid(axis, axis, indexing="ij")
    weights = 0.37 * xw - 0.61 * yw + 0.23 * xw * yw
    center = np.array([0.13, -0.17])
    loss = lambda c: np.mean(weights * po.single_atom_potential(kind, row, 0.2, (16, 16), c, 2))  # noqa: E731
    grad = po.single_atom_center_grad(kind, row, 0.2, (16, 16), center, 2, weights / weights.size)
    d = np.array([0.6, -0.8])
    step = 1e-5
    fd = (loss(center + step * d) - loss(center - step * d)) / (2 * step)
    assert np.linalg.norm(grad) > 1e-6
    np.testing.assert_allclose(grad @ d, fd, rtol=2e-5, atol=1e-7)
    # both components separately, and a second centre off the grid's symmetry lines
    for c in (center, np.array([-0.31, 0.07])):
        g = po.single_atom_center_grad(kind, row, 0.2, (16, 16), c, 2, weights / weights.size)
        for i in range(2):
            e = np.eye(2)[i]
            fd_i = (loss(c + step * e) - loss(c - step * e)) / (2 * step)
            np.testing.assert_allclose(g[i], fd_i, rtol=2e-5, atol=1e-7)


def _crystal_case(seed):
    rng = np.random.default_rng(seed)
    pos = np.column_stack([rng.uniform(0.0, 3.0, 6), rng.uniform(0.0, 2.5, 6), rng.uniform(0.0, 1.9, 6)])
    nums = np.array([6, 14, 6, 14, 14, 6])
    lattice = np.array([[3.4, 0.35, 0.0], [0.2, 3.1, 0.0], [0.0, 0.0, 2.0]])
    rows = {6: LOB[6], 14: LOB[14]}
    return rng, pos, nums, lattice, rows


@pytest.mark.parametrize("repeats", [(1, 1, 1), (2, 1, 1)])
def test_crystal_slices_position_grad_matches_central_differences(repeats):
    """dL/dpositions of L = sum G * slices against central differences of the FP64-accumulated oracle slices, for
    every coordinate of three atoms, among them the atoms that set x_min and y_min (whose shift moves the grid
    origin); lattice copies add into their source atom; z has no gradient."""
    rng, pos, nums, lattice, rows = _crystal_case(11)
    kw = dict(slice_thickness=1.0, repeats=repeats, padding=1.0, supersampling=2)

    def slices(p):
        return po.crystal_potential_slices("lobato", rows, p, nums, lattice, 0.25, accumulate=np.float64, **kw)[0]

    g = rng.standard_normal(slices(pos).shape)
    grad = po.crystal_potential_slices_position_grad("lobato", rows, pos, nums, lattice, 0.25, g, **kw)
    atoms = sorted({int(np.argmin(pos[:, 0])), int(np.argmin(pos[:, 1])), 2})
    h = 1e-5
    for a in atoms:
        for c in range(3):
            dp = np.zeros_like(pos)
            dp[a, c] = h
            fd = (np.sum(g * slices(pos + dp)) - np.sum(g * slices(pos - dp))) / (2 * h)
            np.testing.assert_allclose(grad[a, c], fd, rtol=1e-6, atol=1e-8 * np.abs(grad).max())
    assert np.all(grad[:, 2] == 0.0) and np.abs(grad[:, :2]).max() > 1e-3
