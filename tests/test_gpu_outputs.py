"""A full run on the device writing CATHY's output files (run.run_flow):
the full-precision sidecar against the reference run at its saved steps,
the text files against the sidecar (same values at 7 digits, the
reference's formats), mbeconv against the reference run's statistics."""
from __future__ import annotations

import os

import numpy as np
import pytest

from pycathy_wrapper_amd import configs as CF
from pycathy_wrapper_amd import outputs as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hillslope_small_ref.npz")


def test_run_flow_outputs_match_reference(tmp_path):
    from pycathy_wrapper_amd import gpu as G
    from pycathy_wrapper_amd.run import run_flow
    g = np.load(GOLD)
    saved = [int(v) for v in g["steps"]]
    names = [str(v) for v in g["stat_names"]]
    res = run_flow(CF.hillslope_small(), str(tmp_path), G.ORDER_COLUMN3, output_steps=saved)
    assert res["rc"] == int(g["rc_last"]) and res["steps"] == len(g["stats"])
    recs = O.read_sidecar(str(tmp_path / "psi.f64"))
    swrecs = O.read_sidecar(str(tmp_path / "sw.f64"))
    assert [r[0] for r in recs] == [0] + saved
    for (nstep, t, psi), (_, _, sw) in zip(recs[1:], swrecs[1:]):
        i = saved.index(nstep)
        ref = g["psi"][i]
        assert np.max(np.abs(psi - ref)) <= 1e-6 * np.max(np.abs(ref))
        assert np.max(np.abs(sw - g["sw"][i])) <= 1e-6
        assert t == g["stats"][nstep - 1, names.index("TIME")]
    # the text files are the sidecar values in the reference's formats
    txt = (tmp_path / "psi").read_text()
    assert txt == "".join(O.nodal_block(n, t, v) for n, t, v in recs)
    sw_txt = (tmp_path / "sw").read_text()
    assert sw_txt == O.SW_HEADER + "".join(O.nodal_block(n, t, v) for n, t, v in swrecs)
    # mbeconv: one row per step, 19 columns, values of the reference run
    rows = np.loadtxt(str(tmp_path / "mbeconv"), skiprows=2, usecols=range(19))  # pyCATHY's reader call
    assert rows.shape == (len(g["stats"]), 19)
    st = g["stats"]
    col = lambda nm: st[:, names.index(nm)]
    assert np.array_equal(rows[:, 0], col("NSTEP"))
    assert np.array_equal(rows[:, 3], col("ITER"))
    for j, nm in ((1, "DELTAT"), (2, "TIME"), (5, "STORE1"), (7, "DSTORE"), (9, "VIN"), (11, "VOUT")):
        assert np.allclose(rows[:, j], col(nm), rtol=1e-3, atol=1e-12), nm
