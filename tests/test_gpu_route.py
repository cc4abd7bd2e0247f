"""Surface routing substep (ROUTE + ALTEZZE) on the device against the
oracle, whose orc_route / orc_altezze are bit-identical to the reference's
ROUTE / MC / ALTEZZE (test_oracle_route.py).

Tolerance: 1e-12 relative to the largest value of each array.  The device
evaluates MC's pow / sin / atan / tan with the GPU's math library, so the
outflows may differ from the reference's in the last bits; the inflow sums
follow the reference's order and add nothing further.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from pycathy_wrapper_amd import routing as RT
from test_oracle_route import advance, initial_state

pytestmark = pytest.mark.gpu

TOL = 1e-12
KEYS = ("q_in_kkp1", "q_out_kkp1_1", "q_out_kkp1_2", "volume_kkp1", "h_water_kkp1")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _compare(net, dev_st, orc_states, cu_d, ak_d, cu_o, ak_o, tag):
    for m, st in enumerate(orc_states):
        for k in KEYS:
            e = _rel(dev_st[k].reshape(len(orc_states), -1)[m], st[k])
            assert e <= TOL, (tag, m, k, e)
        assert cu_d[m] == pytest.approx(cu_o[m], rel=TOL) and ak_d[m] == pytest.approx(ak_o[m], rel=TOL), (tag, m)


@pytest.mark.parametrize("reservoir,typed,masked,sinks", [(False, False, 0.0, False), (True, False, 0.0, False),
                                                          (True, True, 0.0, False), (False, False, 0.3, False),
                                                          (False, False, 0.3, True)],
                         ids=["channels", "reservoir", "reservoir-typed-altezze", "masked-catchment", "sinks"])
def test_route_substeps_match_oracle(reservoir, typed, masked, sinks):
    """The 20x20 network of the reference pin, one member, six substeps
    (each substep's new levels feed the next); with a reservoir whose pool
    rises through its spill height.  ALTEZZE takes the reference's all-zero
    cell types, or TIPO_R ("typed"); "masked" leaves 30% of the grid outside
    the catchment (NCELL < NROW*NCOL); "sinks" besides sends part of the
    water of the cells beside it diagonally into cells outside the catchment,
    which the device adds after the levels, in QOI_SN order."""
    net = RT.synthetic_network(20, 20, dx=10.0, seed=2, reservoir=reservoir, masked=masked, sinks=sinks)
    import pyoracle as P
    orc = P.Oracle()
    celtype = net.reservoirs["tipo_r"] if typed else None
    g = RT.RouteGPU(net, celtype=celtype)
    say(f"route 20x20: {g.nlevel} levels, at most {g.maxpred} feeders per cell")
    sd = initial_state(net)
    so = {k: v.copy() for k, v in sd.items()}
    ak_d = ak_o = 0.0
    try:
        for sub in range(6):
            cu_d, akd = g.step(30.0, sd | {"ak_max": ak_d})
            cu_o, ako = orc.route(net, 30.0, so | {"ak_max": ak_o}, celtype=celtype)
            _compare(net, sd, [so], cu_d, akd, [cu_o], [ako], sub)
            if reservoir:
                assert sd["h_pool_kkp1"][0] == pytest.approx(so["h_pool_kkp1"][0], rel=TOL, abs=0), sub
            ak_d, ak_o = float(akd[0]), ako
            advance(sd)
            advance(so)
    finally:
        g.close()


@pytest.mark.parametrize("reservoir", [False, True], ids=["channels", "reservoir"])
def test_route_large_ensemble_matches_oracle(reservoir):
    """A 201x201-cell network (40,401 cells) with four members of different
    surface water and flows in one launch, against the oracle member by
    member (with a reservoir: also its pool level)."""
    net = RT.synthetic_network(201, 201, dx=10.0, seed=4, reservoir=reservoir)
    import pyoracle as P
    orc = P.Oracle()
    g = RT.RouteGPU(net)
    say(f"route 201x201: {g.nlevel} levels, at most {g.maxpred} feeders per cell")
    nm = 4
    sd = initial_state(net, nmember=nm, seed=21)
    orc_states = [{k: v[m].copy() for k, v in sd.items()} for m in range(nm)]
    try:
        cu_d, ak_d = g.step(15.0, sd)
        res = [orc.route(net, 15.0, st) for st in orc_states]
        _compare(net, sd, orc_states, cu_d, ak_d, [r[0] for r in res], [r[1] for r in res], "large")
        if reservoir:
            for m, st in enumerate(orc_states):
                assert sd["h_pool_kkp1"][m, 0] == pytest.approx(st["h_pool_kkp1"][0], rel=TOL, abs=0), m
        assert len(set(np.round(cu_d, 12))) == nm
    finally:
        g.close()


def test_route_rejects_unroutable_network():
    """Cells draining to cells before them in QOI_SN (or off the grid) are
    refused at set-up, before anything reaches the device."""
    net = RT.synthetic_network(6, 5, dx=10.0, seed=1)
    bad = RT.Network(**{**net.__dict__, "qoi": net.qoi[::-1].copy()})
    with pytest.raises(RT.CathyGpuError, match="drains"):
        RT.RouteGPU(bad)


def test_route_rejects_routing_lake_cell():
    """A lake cell (TIPO_R >= 1000) that routes water is refused, where the
    reference stops (route.f: 'INCOMPRENSIBILE CASE')."""
    net = RT.synthetic_network(20, 20, dx=10.0, seed=2, reservoir=True)
    lake = int(np.flatnonzero(net.reservoirs["tipo_r"] >= 1000)[0])
    j, i = lake % net.nrow, lake // net.nrow
    w1 = net.w1.copy()
    w1[j, i] = 1.0
    bad = RT.Network(**{**net.__dict__, "w1": w1})
    with pytest.raises(RT.CathyGpuError, match="TIPO_R >= 1000"):
        RT.RouteGPU(bad)
