"""The ensemble configurations on the device.

1. The lockstep ensemble (cathy_gpu_ensemble_*, one launch per kernel for
   all members, a PCG convergence mask per member) on a reduced mixed-size
   ensemble: 4 S + 2 M + 1 L members of configs.ensemble_member (41x41x15,
   101x101x20 and 201x201x30 nodes, each with its own Ks field and prior
   draws of ln Ks, alpha, n and porosity, under the shared gamma hourly
   rain).  Every member's steps equal those of the same member run alone,
   bit for bit (statistics and psi); the S members follow the oracle step
   by step (dt, Picard counts, back-steps, psi to 1e-6, mass balance to
   1e-8); the M and L members have finite psi and Sw in (0, 1].
2. The stress case: a dry start (psi = -5 m everywhere) under 2 h of
   2e-6 m/s rain, an S member in Newton and one in Picard mode, against the
   oracle for their first steps, ERREL included.  The oracle's dry-start
   balance is pinned against the reference's own Newton and Picard
   trajectories in test_oracle_newton.py (the Newton balance error grows
   past 1% there, a property of the reference's head-based Newton system).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from pycathy_wrapper_amd import configs as CF

pytestmark = pytest.mark.gpu

PSI_TOL = 1e-6
MB_TOL = 1e-8


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _rel_inf(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _mb_close(x, ref, scale):
    return abs(x - ref) <= MB_TOL * max(abs(ref), scale)


def _oracle_run(oracle, pr):
    import pyoracle as P
    m = pr.mesh
    return oracle.run_create(dict(tetra=m.tetra, x=m.x, y=m.y, z=m.z, arenod=m.arenod, nnod=m.nnod, nstr=m.nstr,
                                  ntri=m.ntri), pr.soil, pr.psi0, pr.atm_times, pr.atm_vals, pr.hspatm,
                             P.OrcParams(**pr.params.as_dict()))


def _member(size, seed, iopt=1, dry=False):
    pr = CF.ensemble_member(size, seed, params=CF.hillslope_small_params(tmax=86400.0, iopt=iopt))
    if dry:  # the stress case: dry start, 2 h storm, then nothing
        pr.psi0 = np.full(pr.mesh.n, -5.0)
        pr.atm_times = np.array([0.0, 7200.0, 7201.0])
        pr.atm_vals = np.array([2e-6, 2e-6, 0.0])
    return pr


@pytest.mark.parametrize("layout", ["level-major", "chain-major", "tiled"])
def test_lockstep_mixed_ensemble_matches_solo_runs_and_oracle(oracle, layout):
    from pycathy_wrapper_amd import gpu as G
    K = 3
    plan = [("S", 100), ("S", 101), ("S", 102), ("S", 103), ("M", 132), ("M", 133), ("L", 156)]
    probs = [_member(sz, sd) for sz, sd in plan]
    say("lockstep mixed ensemble: " + ", ".join(f"{sz}{sd} ({p.mesh.n} nodes)" for (sz, sd), p in zip(plan, probs)))
    order = {"level-major": G.ORDER_COLUMN3, "chain-major": G.ORDER_COLUMN3C, "tiled": G.ORDER_COLUMN3T}[layout]
    flows = [G.FlowGPU(p, order) for p in probs]
    ens = G.EnsembleGPU(flows)
    try:
        rc, taken, stats = ens.advance(K)
        info = ens.info()
        say(f"lockstep: rc {rc.tolist()} taken {taken.tolist()} {info}")
        assert np.all(rc == 0) and np.all(taken == K)
        psi = [f.psi() for f in flows]
        sw = [f.sw() for f in flows]
    finally:
        ens.close()
        for f in flows:
            f.close()
    # each member alone: the same steps, bit for bit
    for j, (pr, (sz, sd)) in enumerate(zip(probs, plan)):
        say(f"solo {sz}{sd}")
        g = G.FlowGPU(pr, order)
        try:
            for s in range(K):
                rc_g, sg = g.advance()
                assert rc_g == 0
                assert sg == stats[j][s], (sz, sd, s)
            assert np.array_equal(g.psi(), psi[j]), (sz, sd)
        finally:
            g.close()
        assert np.all(np.isfinite(psi[j])) and np.all(sw[j] > 0.0) and np.all(sw[j] <= 1.0)
    # the S members against the oracle
    for j, (pr, (sz, sd)) in enumerate(zip(probs, plan)):
        if sz != "S":
            continue
        run = _oracle_run(oracle, pr)
        scale = 1e-6 * pr.mesh.volnod.sum()
        try:
            for s in range(K):
                rc_o, so = run.step()
                sg = stats[j][s]
                say(f"S{sd} step {s + 1}: dt {sg['DELTAT']:.4g}/{so['DELTAT']:.4g} iter {sg['ITER']:.0f}/"
                    f"{so['ITER']:.0f} back {sg['BACK']:.0f}/{so['BACK']:.0f}")
                assert rc_o == 0
                assert sg["DELTAT"] == so["DELTAT"] and sg["ITER"] == so["ITER"] and sg["BACK"] == so["BACK"]
                for key in ("DSTORE", "VIN", "VOUT", "ERRAS"):
                    assert _mb_close(sg[key], so[key], scale), (sd, s, key, sg[key], so[key])
            assert _rel_inf(psi[j], run.psi()) <= PSI_TOL
        finally:
            run.close()


def test_lockstep_handles_members_that_stop():
    """Members leave the lockstep at different steps while the others go
    on: a small-hillslope member whose run ends at TMAX = 60 s after 5 steps
    (rc 1), and a mixed-ensemble S member of another mesh size, which stops
    with NOBACK after 2 steps; each member's steps are still its solo
    steps, bit for bit."""
    from pycathy_wrapper_amd import gpu as G
    probs = [CF.hillslope_small_member(0, params=CF.hillslope_small_params(tmax=86400.0)),
             CF.hillslope_small_member(1, params=CF.hillslope_small_params(tmax=60.0)),
             CF.hillslope_small_member(2, params=CF.hillslope_small_params(tmax=86400.0)),
             _member("S", 112)]
    flows = [G.FlowGPU(p, G.ORDER_COLUMN3) for p in probs]
    ens = G.EnsembleGPU(flows)
    try:
        rc, taken, stats = ens.advance(8)
        say(f"rc {rc.tolist()} taken {taken.tolist()}")
        assert rc[1] == 1 and taken[1] == 5 and stats[1][-1]["TIME"] == 60.0
        assert rc[0] == 0 and rc[2] == 0 and taken[0] == 8 and taken[2] == 8
        assert rc[3] != 0 and taken[3] < 8
        psi = [f.psi() for f in flows]
    finally:
        ens.close()
        for f in flows:
            f.close()
    for j, pr in enumerate(probs):
        g = G.FlowGPU(pr, G.ORDER_COLUMN3)
        try:
            for s in range(int(taken[j])):
                rc_g, sg = g.advance()
                assert sg == stats[j][s], (j, s)
            assert rc_g == rc[j]
            assert np.array_equal(g.psi(), psi[j]), j
        finally:
            g.close()


@pytest.mark.parametrize("iopt", [2, 1], ids=["newton", "picard"])
def test_stress_dry_start_matches_oracle(oracle, iopt):
    from pycathy_wrapper_amd import gpu as G
    pr = _member("S", 200 + iopt, iopt=iopt, dry=True)
    say(f"stress dry start, IOPT={iopt}, {pr.mesh.n} nodes")
    g = G.FlowGPU(pr, G.ORDER_COLUMN3)
    run = _oracle_run(oracle, pr)
    scale = 1e-6 * pr.mesh.volnod.sum()
    worst = 0.0
    try:
        for k in range(12):
            rc_o, so = run.step()
            rc_g, sg = g.advance()
            say(f"IOPT={iopt} step {k + 1}: dt {sg['DELTAT']:.4g}/{so['DELTAT']:.4g} iter {sg['ITER']:.0f}/"
                f"{so['ITER']:.0f} ERREL {sg['ERREL']:.6e}/{so['ERREL']:.6e}")
            assert rc_g == rc_o == 0
            assert sg["DELTAT"] == so["DELTAT"] and sg["ITER"] == so["ITER"] and sg["BACK"] == so["BACK"]
            assert _rel_inf(g.psi(), run.psi()) <= PSI_TOL
            for key in ("DSTORE", "VIN", "VOUT", "ERRAS"):
                assert _mb_close(sg[key], so[key], scale), (k, key, sg[key], so[key])
            # ERREL = 100 ERRAS / (VIN + VOUT): the balance error itself agrees
            assert abs(sg["ERREL"] - so["ERREL"]) <= 1e-6 * max(1.0, abs(so["ERREL"])), (k, sg["ERREL"], so["ERREL"])
            worst = max(worst, abs(so["ERREL"]))
    finally:
        g.close()
        run.close()
    say(f"IOPT={iopt}: largest |ERREL| {worst:.4g}% over 12 steps")


def test_lockstep_graphs_match_direct_launches():
    """The lockstep rounds replayed from HIP graphs (each round's fixed
    launch sequences captured once) against the same rounds launched
    directly: every step's statistics and the final psi of every member
    bit for bit.  The graphs are replayed hundreds of times with the
    members' own per-step work and synchronisations between replays, and
    captured again when a member leaves (its run ends at TMAX = 60 s)."""
    from pycathy_wrapper_amd import gpu as G
    probs = [_member("S", 120), _member("S", 121), _member("S", 122), _member("M", 140),
             CF.hillslope_small_member(1, params=CF.hillslope_small_params(tmax=60.0))]
    out = []
    for graphs in (False, True):
        flows = [G.FlowGPU(p, G.ORDER_COLUMN3) for p in probs]
        ens = G.EnsembleGPU(flows)
        ens.graphs(graphs)
        try:
            rc, taken, stats = ens.advance(6)
            info = ens.info()
            say(f"graphs {graphs}: rc {rc.tolist()} taken {taken.tolist()} {info}")
            out.append((rc.tolist(), taken.tolist(), stats, [f.psi() for f in flows], info))
        finally:
            ens.close()
            for f in flows:
                f.close()
    (rc0, tk0, st0, ps0, i0), (rc1, tk1, st1, ps1, i1) = out
    assert i0["graph_captures"] == 0 and i1["graph_captures"] >= 6 and i1["rounds"] > 20
    assert rc0 == rc1 and tk0 == tk1 and rc0[4] == 1
    for j in range(len(probs)):
        assert st0[j] == st1[j], j
        assert np.array_equal(ps0[j], ps1[j]), j
