"""The allocation-free EnKF analysis (cathy_gpu_enkf_update_noalloc) against
the allocating entry (cathy_gpu_enkf_update), through the C ABI.

Two identical ensembles of six perturbed-parameter members at a common time:
one is analysed by each entry with the same weights, twice in a row, the
second time on fewer members (a smaller input, so stale staging from the
first call would show).  Both entries run the same kernel on the same
inputs, so the analysis states must agree bit for bit, and so must the
forecast steps that continue from them.  Sizes above the staging array's
64 members are refused.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from pycathy_wrapper_amd import configs as CF
from pycathy_wrapper_amd import enkf as EK

pytestmark = pytest.mark.gpu


def _ensemble(n, steps):
    from pycathy_wrapper_amd import gpu as G
    fl = []
    for sd in range(n):
        f = G.FlowGPU(CF.hillslope_small_member(sd), G.ORDER_COLUMN3)
        f.set_dt_schedule([10.0] * steps)
        for _ in range(steps):
            rc, st = f.advance()
            assert rc == 0
        fl.append(f)
    return fl


def _weights(flows, seed):
    X = np.stack([f.psi() for f in flows], axis=1)
    nodes = np.arange(5, X.shape[0], 400)[:12]
    HX = EK.observe(flows, nodes, "psi")
    y = HX.mean(axis=1) + 0.02 * np.random.default_rng(seed).standard_normal(len(nodes))
    return EK.analysis_weights(y, np.eye(len(nodes)) * 0.01 ** 2, HX)


def test_noalloc_analysis_matches_allocating_entry():
    print("no-alloc analysis: 2 x 6 members of hillslope_small_member, 8 steps of 10 s", flush=True)
    fa, fb = _ensemble(6, 8), _ensemble(6, 8)
    try:
        for a, b in zip(fa, fb):
            assert np.array_equal(a.psi(), b.psi())
        for call, (m, infl) in enumerate(((6, 1.0), (3, 1.2))):
            W = _weights(fa[:m], seed=31 + call)
            before = [f.psi() for f in fa]
            EK.enkf_update(fa[:m], W, infl)
            EK.enkf_update(fb[:m], W, infl, noalloc=True)
            for j, (a, b) in enumerate(zip(fa, fb)):
                pa, pb = a.psi(), b.psi()
                assert np.array_equal(pa, pb), (call, j, float(np.max(np.abs(pa - pb))))
                if j < m:
                    assert np.max(np.abs(pa - before[j])) > 1e-6, (call, j)  # the analysis moved it
                else:
                    assert np.array_equal(pa, before[j]), (call, j)  # members past m untouched
            for a, b in zip(fa[:m], fb[:m]):
                rc_a, sa = a.advance()
                rc_b, sb = b.advance()
                assert rc_a == rc_b == 0 and sa == sb
    finally:
        for f in fa + fb:
            f.close()


def test_noalloc_analysis_refuses_more_than_64_members():
    from pycathy_wrapper_amd import gpu as G
    n = 65
    h = np.zeros(n, dtype=np.int64)
    w = np.zeros(n * n)
    rc = G.lib().cathy_gpu_enkf_update_noalloc(G._i(n), G._p(h), G._p(w), C.byref(C.c_double(1.0)))
    assert rc < 0
    assert "at most 64" in G.lib().cathy_gpu_last_error().decode()
