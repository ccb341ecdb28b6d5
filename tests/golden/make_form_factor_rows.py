"""Writes tests/golden/form_factor_rows.json: the Lobato--Van Dyck (2014) and
Kirkland (2010) electron-scattering coefficient rows of H, C, Si and Au as the
reference bundles them (``src/ptyrodactyl/inout/luggage/*.csv``, read by
``src/ptyrodactyl/inout/form_factor_data.py:42-160``), in the reference's
in-memory layouts:

* Lobato: ``a1, b1, a2, b2, ..., a5, b5`` (the CSV's grouped a/b columns
  interleaved, ``form_factor_data.py:150-153``);
* Kirkland: ``a1, b1, a2, b2, a3, b3, c1, d1, c2, d2, c3, d3`` (CSV column
  order).

These four elements are the ones the reference's own tests pin values for
(``tests/test_ptyrodactyl/test_multislice/test_atom_potentials.py:32-64``,
``test_form_factors.py:42-81``, ``test_inout/test_form_factor_data.py:114-185``).
The rows are test inputs; the library reads a full table at run time from a
path the caller supplies.

Run from the repository root in a checkout that has the reference at
/root/reference:  python tests/golden/make_form_factor_rows.py
"""

import json
from pathlib import Path

import numpy as np

LUGGAGE = Path("/root/reference/src/ptyrodactyl/inout/luggage")
ELEMENTS = (1, 6, 14, 79)
OUT = Path(__file__).resolve().parent / "form_factor_rows.json"


def main():
    lob = np.loadtxt(LUGGAGE / "Lobato_van_Dyck.csv", delimiter=",", dtype=np.float64, skiprows=1,
                     usecols=range(2, 12))
    kir = np.loadtxt(LUGGAGE / "Kirkland_Potentials.csv", delimiter=",", dtype=np.float64)
    assert lob.shape == (103, 10) and kir.shape == (103, 12)
    inter = np.empty_like(lob)
    inter[:, 0::2] = lob[:, :5]
    inter[:, 1::2] = lob[:, 5:]
    rows = {"source": "src/ptyrodactyl/inout/luggage (Lobato & Van Dyck 2014; Kirkland 2010)",
            "lobato": {str(z): [float(v) for v in inter[z - 1]] for z in ELEMENTS},
            "kirkland": {str(z): [float(v) for v in kir[z - 1]] for z in ELEMENTS}}
    OUT.write_text(json.dumps(rows, indent=1) + "\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main()
