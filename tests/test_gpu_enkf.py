"""EnKF analysis on the device (cathy_gpu_enkf_update) against the CPU
restatement of pyCATHY's enkf_analysis (oracle/enkf_oracle.py), on six
perturbed-parameter members held as flow handles, and the forecast that
continues from the analysis states.

Tolerance: the device forms x_j + sum_k (x_k - xbar) W(k,j) with fused
multiply-adds, the restatement X + (X' S^T/(N-1)) COV^-1 D; the two differ
by rounding, bounded here by 1e-11 of the largest |psi|.
"""
from __future__ import annotations

import numpy as np
import pytest

import enkf_oracle as EO
from pycathy_wrapper_amd import configs as CF
from pycathy_wrapper_amd import enkf as EK

pytestmark = pytest.mark.gpu


def _members(n, steps):
    """n members advanced `steps` steps of 10 s each; also the largest
    |mass-balance error| of those steps, per member."""
    from pycathy_wrapper_amd import gpu as G
    fl, err = [], []
    for sd in range(n):
        f = G.FlowGPU(CF.hillslope_small_member(sd), G.ORDER_COLUMN3)
        f.set_dt_schedule([10.0] * steps)  # every member at the same time
        e = 0.0
        for _ in range(steps):
            rc, st = f.advance()
            assert rc == 0 and st["BACK"] == 0
            e = max(e, abs(st["ERRAS"]), abs(st["VIN"] + st["VOUT"]) * 1e-9)
        fl.append(f)
        err.append(e)
    return fl, err


@pytest.mark.parametrize("sakov,inflation", [(False, 1.0), (True, 1.0), (False, 1.2)])
def test_enkf_update_matches_oracle_and_forecast_continues(sakov, inflation):
    fl, err = _members(6, 8)
    try:
        X = np.stack([f.psi() for f in fl], axis=1)
        nodes = np.arange(5, X.shape[0], 400)[:12]
        HX = EK.observe(fl, nodes, "psi")
        assert np.array_equal(HX, X[nodes])
        rng = np.random.default_rng(11)
        y = HX.mean(axis=1) + 0.02 * rng.standard_normal(len(nodes))
        # the Sakov form divides by diag(R) alone (no ensemble spread), so
        # it takes a larger R to keep the update physical
        R = np.eye(len(nodes)) * (0.5 ** 2 if sakov else 0.01 ** 2)
        W = EK.analysis_weights(y, R, HX, sakov=sakov)
        EK.enkf_update(fl, W, inflation)
        Xa = np.stack([f.psi() for f in fl], axis=1)
        Xo, _ = EO.enkf_analysis(y, R, np.zeros((0, len(fl))), X, HX, sakov=sakov, inflate_states=inflation)
        assert np.max(np.abs(Xa - Xo)) <= 1e-11 * np.max(np.abs(X))
        assert np.max(np.abs(Xa - X)) > 1e-6  # the analysis moved the states
        # the forecast continues from the analysis: steps converge and the
        # mass balance stays as closed as before the analysis
        for f, e in zip(fl, err):
            for _ in range(3):
                rc, st = f.advance()
                assert rc == 0
                assert abs(st["ERRAS"]) <= 100.0 * e, (st["ERRAS"], e)
    finally:
        for f in fl:
            f.close()


def test_enkf_update_rejects_members_that_do_not_align():
    from pycathy_wrapper_amd import gpu as G
    a = G.FlowGPU(CF.hillslope_small_member(0), G.ORDER_COLUMN3)
    b = G.FlowGPU(CF.hillslope_small_member(1), G.ORDER_NATURAL)
    try:
        with pytest.raises(G.CathyGpuError, match="do not align"):
            EK.enkf_update([a, b], np.eye(2) * 0.1)
        with pytest.raises(G.CathyGpuError, match="twice"):
            EK.enkf_update([a, a], np.eye(2) * 0.1)
    finally:
        a.close()
        b.close()
