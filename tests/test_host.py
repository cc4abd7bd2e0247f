"""Host-side set-up (mesh, node soil tables, initial condition, workloads)
against the oracle, and the documented workload sizes."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from pycathy_wrapper_amd import configs as CF
from pycathy_wrapper_amd import mesh as M
from pycathy_wrapper_amd import soil as SO
from pyoracle import dptr
import systems as SY


@pytest.fixture(scope="module")
def small():
    return CF.hillslope_small()


def test_hillslope_small_sizes(small):
    """21x21 nodes, 10 layers: N, NT, NTERM (upper+diag) and full nnz."""
    m = small.mesh
    assert (m.n, m.nt) == (4851, 24000)
    topol, ja = M.sym_pattern(m.tetra, m.n)
    ft, fj = M.full_pattern(m.tetra, m.n)
    assert (len(ja), len(fj)) == (35301, 65751)


def test_node_params_match_tpnodi(oracle, small):
    """soil.node_params (numpy) == orc_tpnodi + orc_chparm (tpnodi.f)."""
    ns = SY.node_soil(oracle, small)
    hp = SO.node_params(small.mesh, small.soil)
    n = small.mesh.n
    assert np.array_equal(hp["pnodi"], ns["pnodi"])
    assert np.array_equal(hp["snodi"], ns["snodi"])
    assert np.array_equal(hp["vgn"], ns["vg"][1 * n:2 * n])
    assert np.array_equal(hp["vgrmc"], ns["vg"][3 * n:4 * n])
    assert np.array_equal(hp["vgpsat"], ns["vg"][4 * n:5 * n])
    assert np.array_equal(hp["tp"], ns["tp"])


def test_node_params_heterogeneous(oracle):
    """Per-(layer, zone) tables of the heterogeneous workload, reduced size."""
    pr = CF.hillslope_large(cells=12, nstr=5)
    ns = SY.node_soil(oracle, pr)
    hp = SO.node_params(pr.mesh, pr.soil)
    n = pr.mesh.n
    assert np.array_equal(hp["pnodi"], ns["pnodi"])
    assert np.array_equal(hp["vgn"], ns["vg"][1 * n:2 * n])
    assert np.array_equal(hp["vgpsat"], ns["vg"][4 * n:5 * n])


def test_icvdwt_matches_oracle(oracle, small):
    m = small.mesh
    p = np.zeros(m.n)
    oracle.L.orc_icvdwt(m.nnod, m.nstr, C.c_double(1.0), dptr(p), dptr(np.ascontiguousarray(m.z)))
    assert np.array_equal(CF.icvdwt(m, 1.0), p)
    # hydrostatic: psi + z constant down each column
    h = (p + m.z).reshape(m.nstr + 1, m.nnod)
    assert np.allclose(h, h[0][None, :], rtol=0, atol=1e-12)


def test_full_pattern_symmetric_and_lookup(small):
    m = small.mesh
    ft, fj = M.full_pattern(m.tetra, m.n)
    rows = np.repeat(np.arange(m.n), np.diff(ft))
    fwd = set(zip(rows.tolist(), fj.tolist()))
    assert all((c, r) in fwd for r, c in fwd)
    q = M.csr_lookup(ft, fj, rows[:1000], fj[:1000], m.n)
    assert np.array_equal(q, np.arange(1000))


def test_hillslope_large_sizes_reduced():
    """The 1.25M-node workload's builder at reduced size: graded layers
    sum to the base thickness, one zone per surface cell, seeded."""
    a = CF.hillslope_large(cells=16, nstr=6)
    b = CF.hillslope_large(cells=16, nstr=6)
    m = a.mesh
    assert m.n == 17 * 17 * 7 and m.nt == 16 * 16 * 2 * 3 * 6
    assert a.soil["nzone"] == 16 * 16
    assert np.array_equal(a.soil["permx"], b.soil["permx"])
    depth = (m.z[:m.nnod] - m.z[m.nstr * m.nnod:])
    assert np.allclose(depth, 10.0, rtol=0, atol=1e-9)
