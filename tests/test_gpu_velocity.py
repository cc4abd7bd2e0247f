"""Darcy velocities on the device (cathy_gpu_flow3d_vel3d) against the
oracle's VEL3D / VNOD3D.

Element values: the device contracts each 4-term sum psi_j b_j into fused
multiply-adds while the oracle rounds every product, so the two differ by a
few ulps of the terms: the bound per element is 1e-13 kr_e K VOLUR
sum_j |psi_j b_j| (plus kr_e Kz for WW's gravity term).  Nodal means add the
device's own element values in element order and divide by the incidence
count, exactly as vnod3d.f does, so they are compared bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from pycathy_wrapper_amd import configs as CF
from pyoracle import dptr, iptr
import systems as SY
import velocity as V

pytestmark = pytest.mark.gpu


def _elem_bound(problem, psi, ckrwe):
    m = problem.mesh
    tet = np.asarray(m.tetra)[:, :4]
    volur = 1.0 / np.asarray(m.volu)
    ks = V.elem_perm(problem)
    out = []
    for coef, k, grav in ((m.bi, ks[0], 0.0), (m.ci, ks[1], 0.0), (m.di, ks[2], 1.0)):
        mag = np.sum(np.abs(psi[tet] * np.asarray(coef)), axis=1)
        out.append(1e-13 * ckrwe * k * (volur * mag + grav))
    return out


@pytest.mark.parametrize("ordering", ["natural", "column3"])
def test_velocity_matches_oracle(oracle, ordering):
    from pycathy_wrapper_amd import gpu as G
    pr = V.layered(CF.hillslope_small())
    m = pr.mesh
    f = G.FlowGPU(pr, G.ORDER_NATURAL if ordering == "natural" else G.ORDER_COLUMN3)
    try:
        for _ in range(6):
            f.advance()
        psi, ck = f.psi(), f.ckrw()
        vel = f.velocity()
        vel2 = f.velocity()  # a second call reuses the cached element table
    finally:
        f.close()
    for k in vel:
        assert np.array_equal(vel[k], vel2[k]), k
    tet = np.ascontiguousarray(m.tetra, dtype=np.int32)
    ckrwe = np.zeros(m.nt)
    oracle.L.orc_nodelt(m.nt, iptr(tet), dptr(ck), dptr(ckrwe))
    ou = V.oracle_vel3d(oracle, pr, psi, ckrwe, 1.0 / np.asarray(m.volu))
    bounds = _elem_bound(pr, psi, ckrwe)
    for name, o, bd in zip(("uu", "vv", "ww"), ou, bounds):
        d = np.abs(vel[name] - o)
        assert np.all(d <= bd), (name, float(np.max(d / bd)))
    assert np.any(ou[2] != 0) and np.any(psi >= 0) and np.any(psi < -0.5)
    tp = SY.node_soil(oracle, pr)["tp"]
    on = V.oracle_vnod3d(oracle, pr, tp, vel["uu"], vel["vv"], vel["ww"])
    for name, o in zip(("unod", "vnod", "wnod"), on):
        assert np.array_equal(vel[name], o), name


def test_velocity_linear_head_on_device():
    """set_pnew with a linear total head: every element's flux is
    (-kr Kx a, -kr Ky b, kr Kz g) with kr_e the mean of the device's kr."""
    from pycathy_wrapper_amd import gpu as G
    pr = V.layered(CF.hillslope_small())
    m = pr.mesh
    f = G.FlowGPU(pr, G.ORDER_COLUMN3)
    try:
        f.advance()
        ck = f.ckrw()
        a, b, g = 0.013, -0.021, 0.3
        f.set_pnew(V.linear_head(pr, a, b, g))
        vel = f.velocity()
    finally:
        f.close()
    kr = ck[np.asarray(m.tetra)[:, :4]].sum(axis=1) * 0.25
    kx, ky, kz = V.elem_perm(pr)
    assert np.allclose(vel["uu"], -kr * kx * a, rtol=1e-7, atol=0)
    assert np.allclose(vel["vv"], -kr * ky * b, rtol=1e-7, atol=0)
    assert np.allclose(vel["ww"], kr * kz * g, rtol=1e-7, atol=0)
