"""CATHY output files: format-exact against the files the reference wrote.

tests/golden/weil_outputs.npz holds the text of the weil run's psi and sw
(first two output times) and mbeconv (235 steps) as the reference wrote
them.  Parsing them and writing the values back through
pycathy_wrapper_amd/outputs.py must give the same bytes: header lines,
FORTRAN 1PEw.d fields, line breaks.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from pycathy_wrapper_amd import outputs as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "weil_outputs.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLD)


@pytest.mark.parametrize("name", ["psi", "sw"])
def test_nodal_files_byte_identical(golden, name, tmp_path):
    txt = str(golden[name])
    fn = tmp_path / name
    fn.write_text(txt)
    blocks = O.read_nodal_text(str(fn))
    assert len(blocks) == 2 and all(len(v) == 7056 for _, _, v in blocks)
    out = (O.SW_HEADER if name == "sw" else "") + "".join(O.nodal_block(a, b, c) for a, b, c in blocks)
    assert out == txt


def test_mbeconv_byte_identical(golden):
    lines = str(golden["mbeconv"]).splitlines(keepends=True)
    assert O.mbeconv_header(float(lines[0].split("=")[1])) == "".join(lines[:3])
    keys = ["NSTEP", "DELTAT", "TIME", "ITER", "AVG", "STORE1", "STORE2", "DSTORE", "CDSTOR", "VIN", "CVIN", "VOUT",
            "CVOUT", "VINOUT", "VTOT", "ERRAS", "ERREL", "CERRAS", "CAERAS"]
    rows = lines[3:]
    assert len(rows) == 235
    for ln in rows:
        st = dict(zip(keys, (float(v) for v in ln.split())))
        st["NITERT"] = st["AVG"] * st["ITER"]
        assert O.mbeconv_row(st) == ln


def test_fortran_e_edge_cases():
    assert O.fortran_e(0.0, 15, 6) == "   0.000000E+00"
    assert O.fortran_e(-0.0, 15, 6) == "  -0.000000E+00"
    assert O.fortran_e(1.0, 16, 8) == "  1.00000000E+00"
    assert O.fortran_e(-1.234567e-5, 15, 6) == "  -1.234567E-05"
    assert O.fortran_e(1.7e-100, 15, 6) == "   1.700000-100"   # Ew.d: three-digit exponent drops the E
    assert O.fortran_e(2.5e120, 13, 5) == "  2.50000+120"
    assert O.fortran_e(1e300, 9, 3) == "1.000+300"
    assert O.fortran_e(123.456, 10, 3) == " 1.235E+02"


def test_sidecar_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    a, b = rng.normal(size=17), rng.normal(size=17)
    with O.CathyOutputs(str(tmp_path)) as out:
        out.nodal(0, 0.0, a, b)
        out.nodal(7, 12.5, b, a)
    recs = O.read_sidecar(str(tmp_path / "psi.f64"))
    assert [(r[0], r[1]) for r in recs] == [(0, 0.0), (7, 12.5)]
    assert np.array_equal(recs[0][2], a) and np.array_equal(recs[1][2], b)
    blocks = O.read_nodal_text(str(tmp_path / "sw"))
    assert np.allclose(blocks[1][2], a, rtol=5e-7, atol=0)
