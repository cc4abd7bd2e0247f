"""The coupled atmospheric switching on the CPU: the oracle's SWITCH and
ADRSTN restatements (orc_switch, orc_adrstn) against the reference's own
routines (switch.f, adrstn.f in oracle/_ref) on random surface states that
reach every branch: IFATM -1/0/1/2, rainfall and evaporation, infiltration
and exfiltration, the ponding head PL and PNEW in each band against
PONDH_MIN and PMIN (PONDH_MIN 0, above 0, below 0 so the bands overlap, and
1e10), the |ATMDIF| < 1e-14 cut, and the guards of cases G and H.  IFATM,
PNEW, ATMACT, OVFLNOD and PONDING must be identical, bit for bit.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libcathy_ref.so")
ORC_SO = os.path.join(ROOT, "oracle", "liboracle.so")

dp = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
ip = lambda x: x.ctypes.data_as(C.POINTER(C.c_int))


def _need():
    for p in (REF_SO, ORC_SO):
        if not os.path.exists(p):
            pytest.skip(f"{p} not built (make -C oracle)")


def _state(rng, nnod, pmin):
    s = dict(
        ifatm=rng.choice([-1, 0, 1, 2], nnod).astype(np.int32),
        arenod=rng.uniform(0.05, 2.0, nnod),
        pondnod=np.where(rng.random(nnod) < 0.5, 0.0, rng.uniform(0.0, 0.05, nnod)),
        atmpot=rng.normal(0.0, 1e-5, nnod),
        qtr=np.where(rng.random(nnod) < 0.6, 0.0, rng.uniform(0.0, 3e-6, nnod)),
        pnew=rng.uniform(pmin - 3.0, 1.0, nnod),
    )
    bigp = rng.random(nnod) < 0.2
    s["atmpot"][bigp] = rng.normal(0.0, 1e-3, bigp.sum())
    s["atmpot"][rng.random(nnod) < 0.05] = 0.0
    s["atmact"] = np.where(rng.random(nnod) < 0.5, rng.normal(0.0, 1e-5, nnod), s["atmpot"] * rng.uniform(0, 2, nnod))
    big = rng.random(nnod) < 0.2        # fluxes large enough to carry PL below PMIN
    s["atmact"][big] = rng.normal(0.0, 1e-3, big.sum())
    same = rng.random(nnod) < 0.08      # ATMDIF exactly 0 or below the 1e-14 cut
    s["atmact"][same] = s["atmpot"][same] + s["qtr"][same] + rng.choice([0.0, 3e-15], same.sum())
    for v in (0.0, pmin):
        s["pnew"][rng.random(nnod) < 0.05] = v
    return s


def _case_ids(s, deltat, ph, pmin):
    """(IFATM, rain, infil, band of PL or PNEW) of each node, for coverage."""
    dif = s["atmpot"] - (s["atmact"] - s["qtr"])
    dif[np.abs(dif) < 1e-14] = 0.0
    pl = s["pondnod"] + dif * deltat / s["arenod"]
    h = np.where(s["ifatm"] == 0, s["pnew"], pl)
    band = np.select([h >= ph, (h >= 0) & (h < ph), (h > pmin) & (h < 0), h <= pmin], [1, 2, 3, 4], 0)
    return set(zip(s["ifatm"].tolist(), (s["atmpot"] >= 0).tolist(), (s["atmact"] >= 0).tolist(), band.tolist()))


@pytest.mark.parametrize("ph", [0.0, 0.02, -0.5, 1e10])
def test_switch_matches_reference(ph):
    _need()
    import pyoracle as P
    ref = P.RefLib(REF_SO)
    orc = P.Oracle(ORC_SO)
    orc.L.orc_switch.restype = C.c_int
    rng = np.random.default_rng(int(abs(ph) * 1000) + 7)
    pmin = -5.0
    ref.chord[0] = pmin
    seen = set()
    for trial in range(6):
        nnod = 8000
        deltat = float(rng.choice([0.5, 60.0, 3600.0]))
        s = _state(rng, nnod, pmin)
        seen |= _case_ids(s, deltat, ph, pmin)
        r = {k: v.copy() for k, v in s.items()}
        r["ovfl"] = np.full(nnod, -7.0)
        pond = np.zeros(1, dtype=np.int32)
        ref.L.switch_(C.byref(C.c_int(nnod)), ip(r["ifatm"]), ip(pond), C.byref(C.c_double(0.0)),
                      C.byref(C.c_double(deltat)), C.byref(C.c_double(ph)), dp(r["arenod"]), dp(r["pondnod"]),
                      dp(r["atmpot"]), dp(r["atmact"]), dp(r["qtr"]), dp(r["pnew"]), dp(r["ovfl"]))
        o = {k: v.copy() for k, v in s.items()}
        o["ovfl"] = np.full(nnod, -7.0)
        po = orc.L.orc_switch(C.c_int(nnod), ip(o["ifatm"]), C.c_double(deltat), C.c_double(ph), C.c_double(pmin),
                              dp(o["arenod"]), dp(o["pondnod"]), dp(o["atmpot"]), dp(o["atmact"]), dp(o["qtr"]),
                              dp(o["pnew"]), dp(o["ovfl"]))
        assert int(po) == int(pond[0] != 0), trial
        for k in ("ifatm", "pnew", "atmact", "ovfl"):
            assert np.array_equal(o[k], r[k]), (trial, k, np.flatnonzero(o[k] != r[k])[:5])
        changed = (r["ifatm"] != s["ifatm"]).sum()
        assert changed > 0
    # every (IFATM 0/1/2, rain, infiltration) combination met every band it can reach
    for fa in (0, 1, 2):
        for rain in (True, False):
            for inf in (True, False):
                bands = {b for (f, r_, i_, b) in seen if f == fa and r_ == rain and i_ == inf}
                need = {1, 2, 3, 4} if ph < 1e9 else {2, 3, 4}
                if ph <= 0:
                    need = {1, 3, 4}   # band 2 (0 <= h < PONDH_MIN) is empty
                if fa > 0 and rain and not inf:
                    need &= {1, 2}     # ATMDIF > 0, so PL >= PONDNOD >= 0
                assert need <= bands, (fa, rain, inf, bands)


def test_adrstn_matches_reference():
    _need()
    import pyoracle as P
    ref = P.RefLib(REF_SO)
    orc = P.Oracle(ORC_SO)
    rng = np.random.default_rng(11)
    pmin = -5.0
    ref.chord[0] = pmin
    nnod = 5000
    s = _state(rng, nnod, pmin)
    r = {k: v.copy() for k, v in s.items()}
    ref.L.adrstn_(C.byref(C.c_int(nnod)), ip(r["ifatm"]), dp(r["atmpot"]), dp(r["atmact"]), dp(r["pnew"]))
    o = {k: v.copy() for k, v in s.items()}
    orc.L.orc_adrstn(C.c_int(nnod), ip(o["ifatm"]), dp(o["atmpot"]), dp(o["atmact"]), dp(o["pnew"]),
                     C.c_double(pmin))
    for k in ("ifatm", "pnew", "atmact"):
        assert np.array_equal(o[k], r[k]), k
    hit = (s["pnew"] <= pmin) & ((s["atmpot"] > 0) | (s["atmact"] < s["atmpot"]))
    assert hit.sum() > 100 and np.all(r["pnew"][hit] == pmin) and np.all(r["ifatm"][hit] == 0)
    assert np.array_equal(r["pnew"][~hit], s["pnew"][~hit]) and np.array_equal(r["ifatm"][~hit], s["ifatm"][~hit])
    assert np.array_equal(r["atmact"][hit], s["atmpot"][hit])
