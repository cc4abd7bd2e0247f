"""Call-order checks of the flow C ABI on the device.

The forcing series' record width follows HSPATM (one value per record, or
one per surface node).  set_forcing stores the series with the width of the
HSPATM in force; init and advance refuse a series whose width no longer
matches (set_params changed HSPATM after it), so the host never reads
records past the end of the series.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from pycathy_wrapper_amd import configs as CF

pytestmark = pytest.mark.gpu


def test_forcing_width_follows_hspatm():
    from pycathy_wrapper_amd import gpu as G
    print("api: hillslope_small, HSPATM changed around the forcing", flush=True)
    pr = CF.hillslope_small(params=CF.hillslope_small_params())
    L = G.lib()
    g = G.FlowGPU(pr, G.ORDER_COLUMN3)       # HSPATM = 1, one value per record
    try:
        rc, st = g.advance()
        assert rc == 0
        g.set_params(pr.params, hspatm=0)     # accepted; the stored series is now stale
        assert L.cathy_gpu_flow3d_advance(C.byref(g.h), G._p(np.zeros(32))) < 0
        assert "HSPATM" in L.cathy_gpu_last_error().decode()
        psi0 = np.ascontiguousarray(pr.psi0)
        assert L.cathy_gpu_flow3d_init(C.byref(g.h), G._p(psi0)) < 0
        # series set again, NNOD values per record: the run starts over and steps
        t = np.array([0.0, 1e6])
        v = np.full(2 * pr.mesh.nnod, pr.atm_vals[0])
        assert L.cathy_gpu_flow3d_set_forcing(C.byref(g.h), G._i(2), G._p(t), G._p(v)) == 0
        assert L.cathy_gpu_flow3d_init(C.byref(g.h), G._p(psi0)) == 0
        rc2, st2 = g.advance()
        assert rc2 == 0
        # the same uniform rain per node: the same first step as the homogeneous series
        assert st2["ITER"] == st["ITER"] and st2["DELTAT"] == st["DELTAT"]
        assert st2["DSTORE"] == pytest.approx(st["DSTORE"], rel=1e-12)
    finally:
        g.close()
