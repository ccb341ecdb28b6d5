"""Electron form factors and projected independent-atom potentials.

Mirrors ``src/ptyrodactyl/multislice/form_factors.py:1-584`` (Lobato--Van
Dyck and Kirkland parameterisations, Lobato the default), the coefficient
carriers of ``src/ptyrodactyl/types/form_factor_types.py`` and the table
loaders of ``src/ptyrodactyl/inout/form_factor_data.py:42-230``.

Form factors are torch expressions on the caller's device; projected
potentials run on the GPU in the HIP kernel ``pty_projected_potential``
(FP64, Bessel K0/K1 evaluated as the reference evaluates them).

Coefficient tables: the published Lobato & Van Dyck (2014) and Kirkland
(2010) tables are data the reference bundles as
``inout/luggage/Lobato_van_Dyck.csv`` and ``Kirkland_Potentials.csv``.  This
package reads them in that format at run time from, in order, the directory
given to :func:`load_form_factor_tables`, the directory in the environment
variable ``PTYRODACTYL_FORM_FACTOR_DIR``, or the ``inout/luggage`` folder of
an installed ``ptyrodactyl`` package (located without importing it).
:func:`set_form_factor_table` installs a table, or some of its rows, directly.
"""

from __future__ import annotations

import importlib.util
import math
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Dict, Optional

import numpy as np
import torch

from . import _native
from .types import _as_tensor, error_if

ELEMENT_COUNT = 103
MOTT_BETHE_VOLT_ANGSTROM_SQ = 47.87801
A_BOHR = 0.529177210903
PARAMETERIZATIONS = ("lobato", "kirkland")
_CODES = {"lobato": 0, "kirkland": 1}
_WIDTH = {"lobato": 10, "kirkland": 12}
_FILES = {"lobato": "Lobato_van_Dyck.csv", "kirkland": "Kirkland_Potentials.csv"}
_TABLES: Dict[str, Optional[np.ndarray]] = {"lobato": None, "kirkland": None}


# ---------------------------------------------------------------------------
# Coefficient carriers (types/form_factor_types.py)
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class LobatoParameters:
    """Five Lobato amplitude / scale pairs (scales in A^2)."""

    amplitudes: torch.Tensor
    scales: torch.Tensor


@dataclass(frozen=True)
class KirklandParameters:
    """Three Lorentzian and three Gaussian Kirkland terms."""

    lorentzian_amplitudes: torch.Tensor
    lorentzian_scales: torch.Tensor
    gaussian_amplitudes: torch.Tensor
    gaussian_scales: torch.Tensor


def _vector(values, name: str, length: int) -> torch.Tensor:
    t = _as_tensor(values, torch.float64)
    if tuple(t.shape) != (length,):
        raise ValueError(f"{name} must have shape ({length},)")
    error_if(~torch.isfinite(t).all(), f"{name} must contain only finite values")
    return t


def _scales(values, name: str, length: int) -> torch.Tensor:
    t = _vector(values, name, length)
    error_if((t <= 0).any(), f"{name} must be strictly positive")
    return t


def create_lobato_parameters(amplitudes, scales) -> LobatoParameters:
    return LobatoParameters(_vector(amplitudes, "amplitudes", 5), _scales(scales, "scales", 5))


def create_kirkland_parameters(lorentzian_amplitudes, lorentzian_scales, gaussian_amplitudes,
                               gaussian_scales) -> KirklandParameters:
    return KirklandParameters(_vector(lorentzian_amplitudes, "lorentzian_amplitudes", 3),
                              _scales(lorentzian_scales, "lorentzian_scales", 3),
                              _vector(gaussian_amplitudes, "gaussian_amplitudes", 3),
                              _scales(gaussian_scales, "gaussian_scales", 3))


# ---------------------------------------------------------------------------
# Tables (inout/form_factor_data.py)
# ---------------------------------------------------------------------------

def _installed_luggage() -> Optional[Path]:
    try:
        spec = importlib.util.find_spec("ptyrodactyl")
    except (ImportError, ValueError):
        return None
    if spec is None or not spec.submodule_search_locations:
        return None
    for loc in spec.submodule_search_locations:
        d = Path(loc) / "inout" / "luggage"
        if (d / _FILES["lobato"]).exists() or (d / _FILES["kirkland"]).exists():
            return d
    return None


def _validated(table: np.ndarray, parameterization: str) -> np.ndarray:
    width = _WIDTH[parameterization]
    name = "Lobato" if parameterization == "lobato" else "Kirkland"
    if table.shape != (ELEMENT_COUNT, width):
        raise ValueError(f"Expected {name} CSV shape {(ELEMENT_COUNT, width)}, got {table.shape}")
    rows = ~np.isnan(table).all(axis=1)
    if not np.all(np.isfinite(table[rows])):
        raise ValueError(f"{name} CSV contains non-finite values")
    if np.any(table[rows][:, 1::2] <= 0):
        raise ValueError(f"{name} scale coefficients must be positive")
    return table


def _read_csv(path: Path, parameterization: str) -> np.ndarray:
    if parameterization == "kirkland":
        return _validated(np.loadtxt(path, delimiter=",", dtype=np.float64), "kirkland")
    numbers = np.loadtxt(path, delimiter=",", dtype=np.int64, skiprows=1, usecols=(0,))
    grouped = np.loadtxt(path, delimiter=",", dtype=np.float64, skiprows=1, usecols=range(2, 12))
    if not np.array_equal(numbers, np.arange(1, ELEMENT_COUNT + 1)):
        raise ValueError("Lobato CSV must contain atomic numbers 1 through 103")
    inter = np.empty((ELEMENT_COUNT, 10), np.float64)
    inter[:, 0::2] = grouped[:, :5]
    inter[:, 1::2] = grouped[:, 5:]
    return _validated(inter, "lobato")


def load_form_factor_tables(directory=None) -> None:
    """Read both coefficient tables (the reference's CSV format) from ``directory``, else from
    ``$PTYRODACTYL_FORM_FACTOR_DIR``, else from an installed ``ptyrodactyl`` package."""
    candidates = [directory, os.environ.get("PTYRODACTYL_FORM_FACTOR_DIR"), _installed_luggage()]
    for d in candidates:
        if d is None:
            continue
        d = Path(d)
        found = False
        for kind in PARAMETERIZATIONS:
            if (d / _FILES[kind]).exists():
                _TABLES[kind] = _read_csv(d / _FILES[kind], kind)
                found = True
        if found:
            return
        if directory is not None and d == Path(directory):
            raise FileNotFoundError(f"no {_FILES['lobato']} or {_FILES['kirkland']} in {d}")


def set_form_factor_table(parameterization: str, table) -> None:
    """Install a (103, 10) Lobato or (103, 12) Kirkland table, or a mapping {atomic number: row} of some rows
    (the other rows stay unavailable).  Lobato rows are a1, b1, ..., a5, b5; Kirkland rows a1, b1, a2, b2, a3,
    b3, c1, d1, c2, d2, c3, d3."""
    _check_parameterization(parameterization)
    if isinstance(table, dict):
        full = np.full((ELEMENT_COUNT, _WIDTH[parameterization]), np.nan)
        for z, row in table.items():
            _checked_python_atom(int(z))
            full[int(z) - 1] = np.asarray(row, np.float64)
        _TABLES[parameterization] = _validated(full, parameterization)
    else:
        _TABLES[parameterization] = _validated(np.array(table, dtype=np.float64), parameterization)


def _table(parameterization: str) -> np.ndarray:
    if _TABLES[parameterization] is None:
        load_form_factor_tables()
    t = _TABLES[parameterization]
    if t is None:
        raise RuntimeError(f"the {parameterization} coefficient table is not loaded: call "
                           f"load_form_factor_tables(directory) with the folder holding {_FILES[parameterization]} "
                           f"(ptyrodactyl/inout/luggage), set PTYRODACTYL_FORM_FACTOR_DIR, or "
                           f"set_form_factor_table(...)")
    return t


def lobato_potentials() -> torch.Tensor:
    """(103, 10) float64 Lobato table, rows a1, b1, ..., a5, b5."""
    return torch.as_tensor(_table("lobato").copy())


def kirkland_potentials() -> torch.Tensor:
    """(103, 12) float64 Kirkland table."""
    return torch.as_tensor(_table("kirkland").copy())


def _check_parameterization(parameterization: str) -> None:
    if parameterization not in _CODES:
        raise ValueError("parameterization must be 'lobato' or 'kirkland'")


def _checked_python_atom(atom_no) -> int:
    if isinstance(atom_no, bool):
        raise ValueError("atom_no must be an integer between 1 and 103")
    if isinstance(atom_no, torch.Tensor):
        error_if((atom_no < 1) | (atom_no > ELEMENT_COUNT), "atom_no must be between 1 and 103")
        return int(atom_no)
    if not 1 <= int(atom_no) <= ELEMENT_COUNT:
        raise ValueError("atom_no must be between 1 and 103")
    return int(atom_no)


def coefficient_row(atom_no, parameterization: str = "lobato") -> np.ndarray:
    """The element's coefficient row, padded to 12 doubles (the kernels' layout)."""
    _check_parameterization(parameterization)
    z = _checked_python_atom(atom_no)
    row = _table(parameterization)[z - 1]
    if np.isnan(row).all():
        raise RuntimeError(f"no {parameterization} coefficients installed for atomic number {z}")
    out = np.zeros(12, np.float64)
    out[:row.shape[0]] = row
    return out


def _load_lobato_parameters(atom_no) -> LobatoParameters:
    row = coefficient_row(atom_no, "lobato")
    return create_lobato_parameters(row[0:10:2], row[1:10:2])


def _load_kirkland_parameters(atom_no) -> KirklandParameters:
    row = coefficient_row(atom_no, "kirkland")
    amp, sc = row[0::2], row[1::2]
    return create_kirkland_parameters(amp[:3], sc[:3], amp[3:], sc[3:])


def _row_of(params) -> np.ndarray:
    row = np.zeros(12, np.float64)
    if isinstance(params, LobatoParameters):
        row[0:10:2] = params.amplitudes.detach().cpu().numpy()
        row[1:10:2] = params.scales.detach().cpu().numpy()
    else:
        row[0:6:2] = params.lorentzian_amplitudes.detach().cpu().numpy()
        row[1:6:2] = params.lorentzian_scales.detach().cpu().numpy()
        row[6:12:2] = params.gaussian_amplitudes.detach().cpu().numpy()
        row[7:12:2] = params.gaussian_scales.detach().cpu().numpy()
    return row


# ---------------------------------------------------------------------------
# Form factors (torch on the caller's device)
# ---------------------------------------------------------------------------

def lobato_form_factor(params: LobatoParameters, q) -> torch.Tensor:
    """f(g) = sum_i a_i (2 + b_i g^2) / (1 + b_i g^2)^2, g = q / 2 pi (A)."""
    q = _as_tensor(q, torch.float64)
    g2 = torch.square(q / (2.0 * math.pi))
    s = g2[..., None] * params.scales.to(q.device)
    return torch.sum(params.amplitudes.to(q.device) * (2.0 + s) / torch.square(1.0 + s), dim=-1)


def kirkland_form_factor(params: KirklandParameters, q) -> torch.Tensor:
    """f(g) = sum_i a_i / (g^2 + b_i) + sum_i c_i exp(-d_i g^2)."""
    q = _as_tensor(q, torch.float64)
    g2 = torch.square(q / (2.0 * math.pi))[..., None]
    dev = q.device
    lor = params.lorentzian_amplitudes.to(dev) / (g2 + params.lorentzian_scales.to(dev))
    gau = params.gaussian_amplitudes.to(dev) * torch.exp(-params.gaussian_scales.to(dev) * g2)
    return torch.sum(lor, dim=-1) + torch.sum(gau, dim=-1)


# ---------------------------------------------------------------------------
# Projected potentials (HIP)
# ---------------------------------------------------------------------------

def _gpu(device=None) -> torch.device:
    if device is not None and torch.device(device).type == "cuda":
        return torch.device(device)
    if not torch.cuda.is_available():
        raise RuntimeError("the potential kernels need a GPU; none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def projected_potential_from_row(row: np.ndarray, parameterization: str, r) -> torch.Tensor:
    """Projected potential (V A) at radii ``r`` (A) of one coefficient row, on the GPU."""
    r = _as_tensor(r, torch.float64)
    error_if(~torch.isfinite(r).all() | (r < 0).any(), "r must contain finite nonnegative radial distances")
    dev = _gpu(r.device)
    rr = r.to(dev).contiguous()
    out = torch.empty_like(rr)
    row_t = torch.as_tensor(row, dtype=torch.float64).to(dev)
    lib = _native.load()
    with torch.cuda.device(dev):
        status = lib.pty_projected_potential(_CODES[parameterization], _native.ptr(row_t), rr.numel(),
                                             _native.ptr(rr), _native.ptr(out), _native.stream_handle(dev))
    _native.check(status, "pty_projected_potential")
    return out


def lobato_projected_potential(params: LobatoParameters, r) -> torch.Tensor:
    """Eq. 16 of Lobato & Van Dyck (2014): 2 pi C_MB sum_i [a_i/b_i K0(2 pi r/sqrt b_i)
    + a_i pi r / b_i^1.5 K1(2 pi r/sqrt b_i)] (``form_factors.py:323-372``)."""
    return projected_potential_from_row(_row_of(params), "lobato", r)


def kirkland_projected_potential(params: KirklandParameters, r) -> torch.Tensor:
    """4 pi^2 a0 e sum a_i K0(2 pi r sqrt b_i) + 2 pi^2 a0 e sum c_i/d_i exp(-pi^2 r^2/d_i) with the legacy
    a0 e = 0.5292 * 14.4 (``form_factors.py:375-439``)."""
    return projected_potential_from_row(_row_of(params), "kirkland", r)


def atomic_form_factor(atom_no, q, *, parameterization: str = "lobato") -> torch.Tensor:
    _check_parameterization(parameterization)
    if parameterization == "lobato":
        return lobato_form_factor(_load_lobato_parameters(atom_no), q)
    return kirkland_form_factor(_load_kirkland_parameters(atom_no), q)


def projected_atom_potential(atom_no, r, *, parameterization: str = "lobato") -> torch.Tensor:
    _check_parameterization(parameterization)
    return projected_potential_from_row(coefficient_row(atom_no, parameterization), parameterization, r)


def lobato_bandlimited_peak(atom_no, g_max) -> torch.Tensor:
    """On-nucleus peak (V) of the Lobato potential band-limited at ``g_max`` (1/A):
    4 pi C_MB sum_i a_i b_i^-1.5 [X_i - X_i / (2 (1 + X_i^2)) - atan(X_i) / 2], X_i = g_max sqrt(b_i)."""
    if isinstance(g_max, (int, float)) and not isinstance(g_max, bool) and (not math.isfinite(g_max)
                                                                             or g_max <= 0):
        raise ValueError("g_max must be positive and finite")
    g = _as_tensor(g_max, torch.float64)
    error_if(~torch.isfinite(g) | (g <= 0), "g_max must be positive and finite")
    p = _load_lobato_parameters(atom_no)
    x = g * torch.sqrt(p.scales)
    anti = x - x / (2.0 * (1.0 + torch.square(x))) - 0.5 * torch.atan(x)
    return 4.0 * math.pi * MOTT_BETHE_VOLT_ANGSTROM_SQ * torch.sum(p.amplitudes * torch.pow(p.scales, -1.5) * anti)


__all__ = ["A_BOHR", "KirklandParameters", "LobatoParameters", "MOTT_BETHE_VOLT_ANGSTROM_SQ", "atomic_form_factor",
           "coefficient_row", "create_kirkland_parameters", "create_lobato_parameters", "kirkland_form_factor",
           "kirkland_potentials", "kirkland_projected_potential", "load_form_factor_tables",
           "lobato_bandlimited_peak", "lobato_form_factor", "lobato_potentials", "lobato_projected_potential",
           "projected_atom_potential", "set_form_factor_table"]
