"""The oracle pinned against the reference's own Fortran routines.

oracle/_ref/libcathy_ref.so is the reference's FLOW3D closure compiled from
its own sources (oracle/Makefile, flang).  Each test hands the same inputs to
the oracle routine and to the reference routine it restates and requires
bit-identical results (the oracle is built without FMA contraction and
follows the reference's operation order).  Skipped where the reference
library has not been built (it needs the reference tree at build time).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import has_ref
from pycathy_wrapper_amd import configs as CF
from pyoracle import dptr, iptr
import systems as SY

pytestmark = pytest.mark.skipif(not has_ref(), reason="oracle/_ref not built (needs the reference sources)")

I = lambda v: C.byref(C.c_int(int(v)))
D = lambda v: C.byref(C.c_double(float(v)))


@pytest.fixture(scope="module")
def ref():
    import pyoracle
    return pyoracle.RefLib()


@pytest.fixture(scope="module")
def problem():
    return CF.hillslope_small()


@pytest.fixture(scope="module")
def system(oracle, problem):
    return SY.picard_system(oracle, problem, deltat=3600.0)


def _ref_perm(problem, key):
    # PERMX(MAXSTR=15, MAXZON=1) of the reference's CATHY.H, column-major
    from pyoracle import REF_MAXSTR
    out = np.zeros(REF_MAXSTR)
    out[:problem.mesh.nstr] = np.asarray(problem.soil[key])[:, 0]
    return out


def test_chpic0_bit_identical(ref, system):
    """chpic0.f:23-36 with fvgse/fvgkr/fvgdse (van Genuchten, IVGHU=0)."""
    s = system
    n = s["n"]
    ref.set_vg(s["vg"], n)
    ro = {k: np.zeros(n) for k in ("swnew", "swtimep", "sw", "ckrw", "etai", "et1", "et2")}
    ref.L.chpic0_(I(n), dptr(s["psi"]), dptr(s["psi"]), dptr(s["ptimep"]), dptr(ro["swnew"]), dptr(ro["swtimep"]),
                  dptr(s["snodi"]), dptr(s["pnodi"]), dptr(ro["sw"]), dptr(ro["ckrw"]), dptr(ro["etai"]),
                  dptr(ro["et1"]), dptr(ro["et2"]), I(0))
    for k in ro:
        assert np.array_equal(ro[k], s[k]), k
    # both saturated (psi >= 0) and unsaturated nodes are exercised
    assert (s["psi"] >= 0).any() and (s["psi"] < -0.5).any()


def test_nodelt_bit_identical(ref, problem, system):
    m = problem.mesh
    tet1 = np.ascontiguousarray(m.tetra, dtype=np.int32).copy()
    tet1[:, :4] += 1
    out = np.zeros(m.nt)
    ref.L.nodelt_(I(m.nt), iptr(tet1), dptr(system["ckrw"]), dptr(out))
    assert np.array_equal(out, system["ckrwe"])


def test_asspic_bit_identical(ref, problem, system):
    """asspic.f:9-54: stiffness, storage and porosity matrices."""
    m, s = problem.mesh, system
    tet1 = np.ascontiguousarray(m.tetra, dtype=np.int32).copy()
    tet1[:, :4] += 1
    rtj = np.zeros((m.nt, 4, 4), dtype=np.int32)
    for k in range(4):
        for l in range(4):
            v = s["tetja"][:, k * 4 + l]
            ok = v >= 0
            rtj[ok, l, k] = v[ok] + 1
    nterm = s["nterm"]
    r1 = np.zeros(nterm); r2 = np.zeros(nterm); r4 = np.zeros(nterm)
    kx, ky, kz = (_ref_perm(problem, k) for k in ("permx", "permy", "permz"))
    ref.L.asspic_(I(m.nt), I(m.ntri), iptr(tet1), iptr(np.ascontiguousarray(rtj)), dptr(s["lmass"]), dptr(r1),
                  dptr(r2), dptr(r4), dptr(s["et1e"]), dptr(s["pele"]), dptr(kx), dptr(ky), dptr(kz),
                  dptr(s["ckrwe"]), dptr(s["etai"]), dptr(np.ascontiguousarray(m.volu)), dptr(s["volur"]),
                  dptr(np.ascontiguousarray(m.bi)), dptr(np.ascontiguousarray(m.ci)),
                  dptr(np.ascontiguousarray(m.di)))
    assert np.array_equal(r1, s["asm"]["c1"])
    assert np.array_equal(r2, s["asm"]["c2"])
    assert np.array_equal(r4, s["asm"]["c4"])


def test_rhspic_cfmatp_rhsgrv_bit_identical(ref, problem, system):
    """rhspic.f:10-41, cfmatp.f:9-29 and rhsgrv.f:9-33 chained as picard.f does."""
    m, s = problem.mesh, system
    n, nterm = s["n"], s["nterm"]
    top1 = (s["topol"] + 1).astype(np.int32)
    ja1 = (s["ja"] + 1).astype(np.int32)
    c1 = s["asm"]["c1"].copy(); c2 = s["asm"]["c2"].copy(); c4 = s["asm"]["c4"].copy()
    tno = np.zeros(n)
    ref.L.rhspic_(I(n), iptr(ja1), iptr(top1), D(s["deltat"]), dptr(s["psi"]), dptr(s["psi"]), dptr(s["ptimep"]),
                  dptr(s["swnew"]), dptr(s["swtimep"]), dptr(c1), dptr(c2), dptr(c4), dptr(tno))
    assert np.array_equal(tno, s["rhs_rhspic"])
    c5 = np.zeros(nterm)
    ref.L.cfmatp_(I(n), I(nterm), D(s["tetaf"]), D(s["deltat"]), dptr(c1), dptr(c2), dptr(c4), dptr(c5),
                  iptr(top1), dptr(s["et2"]))
    assert np.array_equal(c1, s["coef1"])
    tet1 = np.ascontiguousarray(m.tetra, dtype=np.int32).copy()
    tet1[:, :4] += 1
    ref.L.rhsgrv_(I(m.nt), I(m.ntri), iptr(tet1), dptr(tno), dptr(np.ascontiguousarray(m.di)),
                  dptr(_ref_perm(problem, "permz")), iptr(np.ascontiguousarray(m.ivol, dtype=np.int32)),
                  dptr(s["ckrwe"]))
    assert np.array_equal(tno, s["tnoti"])


def _ref_symslv(ref, topol, ja, coef1, tnoti, noddir1, itmxcg=500, tolcg=1e-10, rmin=1.7e-100):
    n = len(topol) - 1
    nterm = len(ja)
    x = np.zeros(n); cf2 = np.zeros(nterm); scr = np.zeros(5 * n)
    nd = np.ascontiguousarray(noddir1 if noddir1 is not None else np.zeros(1), dtype=np.int32)
    numdir = 0 if noddir1 is None else len(nd)
    it = C.c_int()
    c1 = np.ascontiguousarray(coef1).copy()
    ref.L.symslv_(I(0), I(n), I(nterm), I(numdir), C.byref(it), I(itmxcg), D(tolcg), D(rmin), iptr(nd),
                  iptr((ja + 1).astype(np.int32)), iptr((topol + 1).astype(np.int32)), dptr(x),
                  dptr(np.ascontiguousarray(tnoti)), dptr(c1), dptr(cf2), dptr(scr))
    return x, it.value, cf2


@pytest.mark.parametrize("dirichlet", [False, True])
def test_symslv_bit_identical(oracle, ref, problem, system, dirichlet):
    """solscal-extended.f SYMSLV = INCLDP + PRODDP + GRADDP + RESSYM, with
    and without BCPIC penalty rows."""
    s = system
    if dirichlet:
        rows = np.arange(0, problem.mesh.nnod, 5)  # a fifth of the surface nodes
        c1, tno, nd1 = SY.penalise(s, rows)
        nd0 = nd1 - 1
    else:
        c1, tno, nd1, nd0 = s["coef1"], s["tnoti"], None, None
    x, it, err, res, cf2 = SY.oracle_symslv(oracle, s["topol"], s["ja"], c1, tno, nd0)
    rx, rit, rcf2 = _ref_symslv(ref, s["topol"], s["ja"], c1, tno, nd1)
    assert it == rit and it > 3
    assert np.array_equal(cf2, rcf2)  # the IC(0) factor
    assert np.array_equal(x, rx)
    assert err <= 1e-10


def test_whole_run_bit_identical(oracle, ref, problem):
    """The oracle's time loop vs the reference's FLOW3D driven through the
    same loop (oracle/refrun.py): 40 steps, identical trajectories."""
    import pyoracle as P
    import refrun as R
    m = problem.mesh
    rr = R.RefFlowRun(ref, m, problem.soil, problem.psi0, float(problem.atm_vals[0]), problem.params.as_dict())
    run = oracle.run_create(dict(tetra=m.tetra, x=m.x, y=m.y, z=m.z, arenod=m.arenod, nnod=m.nnod, nstr=m.nstr,
                                 ntri=m.ntri), problem.soil, problem.psi0, problem.atm_times, problem.atm_vals, problem.hspatm,
                            P.OrcParams(**problem.params.as_dict()))
    try:
        for k in range(40):
            rc1, s1 = rr.step()
            rc2, s2 = run.step()
            assert rc1 == rc2
            for key in ("DELTAT", "TIME", "ITER", "NITERT", "BACK", "DSTORE", "VIN", "VOUT", "ERRAS", "ERREL",
                        "STORE1", "STORE2", "CDSTOR"):
                assert s1[key] == s2[key], (k, key, s1[key], s2[key])
            assert np.array_equal(rr.psi(), run.psi()), k
            assert np.array_equal(rr.sw(), run.sw()), k
            if rc1:
                break
    finally:
        run.close()


def _atm_series(nnod, hspatm, seed=3):
    """Rain and evaporation records at uneven times over the first 215 s
    (the run's steps shrink to under a second near t = 200 s as the water
    table falls): per-node rates of both signs (HSPATM = 0) or one rate
    (HSPATM = 1).  The last record falls inside the run, so the reader meets
    END (HTIATM = 1) as well."""
    rng = np.random.default_rng(seed)
    times = np.array([0.0, 40.0, 95.0, 150.0, 190.0, 215.0])
    base = np.array([2e-7, 6e-7, -3e-7, 9e-7, -1e-7, 4e-7])
    if hspatm:
        return times, base
    return times, (base[:, None] + 5e-7 * rng.uniform(-1.0, 1.0, (len(times), nnod))).ravel()


@pytest.mark.parametrize("hspatm,ieto", [(0, 0), (1, 1)])
def test_whole_run_atmbc_file_bit_identical(oracle, ref, problem, tmp_path, hspatm, ieto):
    """ATMONE / ATMNXT / ATMBAK and BKSTEP pinned: the reference's own
    readers take the forcing from an atmbc file, and back-steps run the
    reference's BKSTEP (oracle/refrun.py, oracle/refatm_io.f); the oracle
    takes the same records as arrays.  Spatially variable
    piecewise-linear records (HSPATM = 0, IETO = 0) and homogeneous step
    records (HSPATM = 1, IETO = 1), with rain, evaporation, infiltration
    switching (PMIN just above the initial surface head, so evaporating
    nodes start head-limited) and back-steps; trajectories and IFATM
    bit-identical for 70 steps."""
    import dataclasses
    import pyoracle as P
    import refrun as R
    m = problem.mesh
    prm = dataclasses.replace(problem.params, ieto=ieto, pmin=-0.995, deltat=30.0).as_dict()
    times, vals = _atm_series(m.nnod, hspatm)
    path = str(tmp_path / "atmbc")
    R.write_atmbc(path, times, vals, hspatm, ieto)
    rr = R.RefFlowRun(ref, m, problem.soil, problem.psi0, 0.0, prm, atmbc=path)
    bak_calls = []      # back-steps on which BKSTEP took ATMBAK (TIME <= ATMTIM(2), bkstep.f:115-122)
    ref_bkstep = rr._bkstep_ref

    def counted_bkstep():
        atmtim = rr.a["ATMTIM"][:3].copy()
        ref_bkstep()
        if rr.a["TIME"][0] <= atmtim[1] and atmtim[0] < atmtim[1]:
            bak_calls.append(rr.a["TIME"][0])
    rr._bkstep_ref = counted_bkstep
    run = oracle.run_create(dict(tetra=m.tetra, x=m.x, y=m.y, z=m.z, arenod=m.arenod, nnod=m.nnod, nstr=m.nstr,
                                 ntri=m.ntri), problem.soil, problem.psi0, times, vals, hspatm, P.OrcParams(**prm))
    backs = 0
    switched = set()
    try:
        for k in range(70):
            rc1, s1 = rr.step()
            rc2, s2 = run.step()
            assert rc1 == rc2
            for key in ("DELTAT", "TIME", "ITER", "NITERT", "BACK", "DSTORE", "VIN", "VOUT", "ERRAS", "ERREL",
                        "STORE1", "STORE2", "CDSTOR"):
                assert s1[key] == s2[key], (k, key, s1[key], s2[key])
            assert np.array_equal(rr.psi(), run.psi()), k
            assert np.array_equal(rr.sw(), run.sw()), k
            ifatm = rr.a["IFATM"][:m.nnod].copy()
            assert np.array_equal(ifatm, run.ifatm()), k
            switched.update(np.unique(ifatm).tolist())
            backs += s1["BACK"]
            if rc1:
                break
        assert rr.a["TIME"][0] > times[-1]          # the series ran out inside the run
        assert rr.a["HTIATM"][0] == 1
    finally:
        run.close()
    assert backs > 0 and bak_calls, "no back-step into ATMBAK with two distinct record times"
    assert {0, 1} <= switched, switched


# ---------------------------------------------------- mesh set-up pins --
IP3 = np.array([1, 2, 3, 2, 3, 1, 3, 1, 2], dtype=np.int32)            # cathy_main.f:2490
IP4 = np.array([1, 2, 3, 4, 2, 3, 4, 1, 3, 4, 1, 2, 4, 1, 2, 3], dtype=np.int32)  # cathy_main.f:2491


def _tet1(m):
    t = np.ascontiguousarray(m.tetra, dtype=np.int32).copy()
    t[:, :4] += 1
    return t


def test_volbas_bit_identical(oracle, ref, problem):
    """volbas.f:10-54 (volums.f, basis6.f): basis coefficients, volumes,
    nodal volumes -- the host builder and the oracle against the reference."""
    m = problem.mesh
    n, nt = m.n, m.nt
    ai = np.zeros((nt, 4)); bi = np.zeros((nt, 4)); ci = np.zeros((nt, 4)); di = np.zeros((nt, 4))
    volnod = np.zeros(n); volu = np.zeros(nt); volur = np.zeros(nt); ivol = np.zeros(nt, dtype=np.int32)
    x, y, z = (np.ascontiguousarray(a) for a in (m.x, m.y, m.z))
    ref.L.volbas_(I(n), I(nt), iptr(_tet1(m)), iptr(IP4), dptr(ai), dptr(bi), dptr(ci), dptr(di), dptr(volnod),
                  dptr(volu), dptr(volur), iptr(ivol), dptr(x), dptr(y), dptr(z))
    for got, want in ((bi, m.bi), (ci, m.ci), (di, m.di), (volu, m.volu), (volnod, m.volnod), (ivol, m.ivol)):
        assert np.array_equal(got, want)
    oa = oracle.volbas(m.tetra, x, y, z)
    assert np.array_equal(oa[0], ai) and np.array_equal(oa[6], volur)


def test_strpic_tetpic_bit_identical(oracle, ref, problem):
    """strpic.f:8-104 / tetpic.f:10-31: the symmetric pattern (diagonal
    first) and the element-to-entry map, against the host builder and the
    oracle."""
    from pycathy_wrapper_amd import mesh as M
    m = problem.mesh
    n, nt, n1 = m.n, m.nt, 25
    ja = np.zeros(n1 * n, dtype=np.int32)
    topol = np.zeros(n + 1, dtype=np.int32)
    nterm = C.c_int(0)
    tet = _tet1(m)
    # IMAX: sentinel above every node id for strpic.f's vertex selection sort
    ref.L.strpic_(I(n), C.byref(nterm), iptr(tet), iptr(ja), iptr(topol), I(nt), I(n1), I(2 ** 30))
    nterm = nterm.value
    htop, hja = M.sym_pattern(m.tetra, n)
    assert nterm == len(hja) == 35301
    assert np.array_equal(topol - 1, htop) and np.array_equal(ja[:nterm] - 1, hja)
    tetja = np.zeros((nt, 4, 4), dtype=np.int32)  # Fortran TETJA(4,4,NT): [e][j][i] here
    ref.L.tetpic_(I(nt), iptr(tet), iptr(ja), iptr(topol), iptr(tetja))
    tj = M.tetja_sym(m.tetra, htop, hja, n)
    for idx, (k, l) in enumerate(M.UPPER_PAIRS):
        assert np.array_equal(tetja[:, l, k] - 1, tj[:, idx])
    ot, oj, _ = oracle.pattern(n, m.tetra)
    assert np.array_equal(ot, htop) and np.array_equal(oj, hja)


def test_area2d_bit_identical(oracle, ref, problem):
    """area2d.f / areas.f: a third of each incident triangle's area."""
    m = problem.mesh
    tri = np.ascontiguousarray(m.triang, dtype=np.int32).copy()
    tri[:, :3] += 1
    are = np.zeros(m.nnod)
    x, y = np.ascontiguousarray(m.x[:m.nnod]), np.ascontiguousarray(m.y[:m.nnod])
    ref.L.area2d_(I(m.nnod), I(m.ntri), iptr(tri), iptr(IP3), dptr(are), dptr(x), dptr(y))
    assert np.array_equal(are, m.arenod)
    assert np.array_equal(oracle.area2d(m.nnod, m.triang, x, y), are)


def test_strnew_bit_identical(oracle, ref, problem):
    """strnew.f: the full pattern (all columns of a row, ascending) the
    Newton path and the device's SELL layout are built from."""
    from pycathy_wrapper_amd import mesh as M
    m = problem.mesh
    n, nt, n1 = m.n, m.nt, 40
    ja = np.zeros(n1 * n, dtype=np.int32)
    topol = np.zeros(n + 1, dtype=np.int32)
    nterm = C.c_int(0)
    ref.L.strnew_(I(n), C.byref(nterm), iptr(_tet1(m)), iptr(ja), iptr(topol), I(nt), I(n1), I(2 ** 30))
    ft, fj = M.full_pattern(m.tetra, n)
    assert nterm.value == len(fj) == 65751
    assert np.array_equal(topol - 1, ft) and np.array_equal(ja[:nterm.value] - 1, fj)
    ot, oj, _ = oracle.pattern(n, m.tetra, full=True)
    assert np.array_equal(ot, ft) and np.array_equal(oj, fj)


def test_icvdwt_bit_identical(oracle, ref, problem):
    """icvdwt.f:8-31, the hydrostatic initial head (INDP=4)."""
    from pycathy_wrapper_amd import configs as CFG
    m = problem.mesh
    p = np.zeros(m.n)
    pond = np.zeros(m.nnod)
    z = np.ascontiguousarray(m.z)
    ref.L.icvdwt_(I(m.nnod), I(m.nstr), I(m.n), I(0), I(0), D(1.0), dptr(p), dptr(pond), dptr(z))
    assert np.array_equal(CFG.icvdwt(m, 1.0), p)
    q = np.zeros(m.n)
    oracle.L.orc_icvdwt(m.nnod, m.nstr, C.c_double(1.0), dptr(q), dptr(z))
    assert np.array_equal(q, p)
