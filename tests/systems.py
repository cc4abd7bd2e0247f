"""Picard linear systems assembled by the oracle (TEST INFRASTRUCTURE).

Builds, for a ``configs.Problem``, the quantities one Picard iteration of
picard.f:94-180 forms -- node soil tables (TPNODI/CHPARM), CHPIC0, NODELT,
ASSPIC, RHSPIC, CFMATP, RHSGRV and, optionally, BCPIC's penalty rows -- with
the C oracle, so that tests can hand the same system to the reference's
routines (oracle/_ref), to the oracle's SYMSLV and to the device SYMSLV.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from pycathy_wrapper_amd import mesh as M
from pyoracle import dptr, iptr

RMAX = 1.7e100  # cathy_main.f's RMAX: BCPIC's penalty is 1e-9 * RMAX


def node_soil(orc, pr):
    """TPNODI + CHPARM (tpnodi.f:9-70, chparm.f:11-34)."""
    m, s = pr.mesh, pr.soil
    n, nt = m.n, m.nt
    pn = np.zeros(n); sn = np.zeros(n); vg = np.zeros(11 * n); pel = np.zeros(nt)
    tp = np.zeros(n, dtype=np.int32)
    tet = np.ascontiguousarray(m.tetra, dtype=np.int32)
    f = lambda k: np.ascontiguousarray(s[k], dtype=np.float64)
    rc = orc.L.orc_tpnodi(n, nt, m.ntri, iptr(tet), s["nzone"], dptr(f("poros")), dptr(f("elstor")), dptr(f("vgn")),
                          dptr(f("vgrmc")), dptr(f("vgpsat")), dptr(pn), dptr(sn), dptr(vg), dptr(pel), iptr(tp))
    assert rc == 0
    orc.L.orc_chparm(n, dptr(pn), dptr(vg))
    return dict(pnodi=pn, snodi=sn, vg=vg, pel=pel, tp=tp)


def picard_system(orc, pr, psi: np.ndarray | None = None, deltat: float = 10.0, tetaf: float = 1.0):
    """One Picard system at iterate ``psi`` (PTNEW = PNEW = psi) from
    PTIMEP = psi0, as picard.f:94-150 forms it before BCPIC."""
    m, s = pr.mesh, pr.soil
    n, nt = m.n, m.nt
    ns = node_soil(orc, pr)
    ptimep = np.ascontiguousarray(pr.psi0, dtype=np.float64)
    if psi is None:
        psi = ptimep.copy()
        psi[::7] -= 0.7
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    out = {k: np.zeros(n) for k in ("swnew", "swtimep", "sw", "ckrw", "etai", "et1", "et2")}
    orc.L.orc_chpic0(n, dptr(psi), dptr(psi), dptr(ptimep), dptr(out["swnew"]), dptr(out["swtimep"]),
                     dptr(ns["snodi"]), dptr(ns["pnodi"]), dptr(ns["vg"]), dptr(out["sw"]), dptr(out["ckrw"]),
                     dptr(out["etai"]), dptr(out["et1"]), dptr(out["et2"]))
    tet = np.ascontiguousarray(m.tetra, dtype=np.int32)
    ckrwe = np.zeros(nt); et1e = np.zeros(nt); pele = np.zeros(nt)
    orc.L.orc_nodelt(nt, iptr(tet), dptr(out["ckrw"]), dptr(ckrwe))
    orc.L.orc_nodelt(nt, iptr(tet), dptr(out["et1"]), dptr(et1e))
    orc.L.orc_nodelt(nt, iptr(tet), dptr(ns["pnodi"]), dptr(pele))
    topol, ja = M.sym_pattern(m.tetra, n)
    topol = np.ascontiguousarray(topol, dtype=np.int32)
    ja = np.ascontiguousarray(ja, dtype=np.int32)
    nterm = len(ja)
    tj = M.tetja_sym(m.tetra, topol, ja, n)
    otj = np.full((nt, 16), -1, dtype=np.int32)
    for idx, (k, l) in enumerate(M.UPPER_PAIRS):
        otj[:, k * 4 + l] = tj[:, idx]
    lm = np.zeros(16)
    orc.L.orc_locmas(dptr(lm), int(pr.params.lump))
    perm = [np.ascontiguousarray(s[k], dtype=np.float64) for k in ("permx", "permy", "permz")]
    c1 = np.zeros(nterm); c2 = np.zeros(nterm); c4 = np.zeros(nterm)
    volur = 1.0 / m.volu
    bi, ci, di = (np.ascontiguousarray(a) for a in (m.bi, m.ci, m.di))
    volu = np.ascontiguousarray(m.volu)
    orc.L.orc_asspic(nt, m.ntri, s["nzone"], iptr(tet), iptr(otj), dptr(lm), dptr(c1), dptr(c2), dptr(c4),
                     dptr(et1e), dptr(pele), dptr(perm[0]), dptr(perm[1]), dptr(perm[2]), dptr(ckrwe), dptr(volu),
                     dptr(volur), dptr(bi), dptr(ci), dptr(di))
    asm = dict(c1=c1.copy(), c2=c2.copy(), c4=c4.copy())
    tno = np.zeros(n)
    orc.L.orc_rhspic(n, iptr(ja), iptr(topol), C.c_double(deltat), dptr(psi), dptr(psi), dptr(ptimep),
                     dptr(out["swnew"]), dptr(out["swtimep"]), dptr(c1), dptr(c2), dptr(c4), dptr(tno))
    rhs = tno.copy()
    c5 = np.zeros(nterm)
    orc.L.orc_cfmatp(n, nterm, C.c_double(tetaf), C.c_double(deltat), dptr(c1), dptr(c2), dptr(c4), dptr(c5),
                     iptr(topol), dptr(out["et2"]))
    ivol = np.ascontiguousarray(m.ivol, dtype=np.int32)
    orc.L.orc_rhsgrv(nt, m.ntri, s["nzone"], iptr(tet), dptr(tno), dptr(di), dptr(perm[2]), iptr(ivol),
                     dptr(ckrwe))
    return dict(n=n, nterm=nterm, topol=topol, ja=ja, tetja=otj, lmass=lm, psi=psi, ptimep=ptimep,
                deltat=deltat, tetaf=tetaf, ckrwe=ckrwe, et1e=et1e, pele=pele, volur=volur, asm=asm,
                rhs_rhspic=rhs, coef1=c1, tnoti=tno, **out, **ns)


def penalise(sysd: dict, rows: np.ndarray):
    """BCPIC's Dirichlet treatment (bcpic.f:34-65) on 0-based ``rows``:
    TNOTI = 0 and the diagonal set to 1e-9 * RMAX.  Returns (coef1, tnoti,
    noddir 1-based)."""
    c1 = sysd["coef1"].copy()
    tno = sysd["tnoti"].copy()
    rows = np.asarray(rows, dtype=np.int64)
    c1[sysd["topol"][rows]] = 1e-9 * RMAX
    tno[rows] = 0.0
    return c1, tno, (rows + 1).astype(np.int32)


def oracle_symslv(orc, topol, ja, coef1, tnoti, noddir0=None, itmxcg=500, tolcg=1e-10, rmin=1.7e-100):
    """orc_symslv on 0-based CSR / 0-based NODDIR; returns (x, niter, err, res, factor)."""
    n = len(topol) - 1
    nterm = len(ja)
    nd = np.ascontiguousarray(noddir0 if noddir0 is not None else np.zeros(1), dtype=np.int32)
    numdir = 0 if noddir0 is None else len(nd)
    x = np.zeros(n); cf2 = np.zeros(nterm); scr = np.zeros(5 * n)
    err = C.c_double(); res = C.c_double()
    it = orc.L.orc_symslv(n, nterm, numdir, itmxcg, C.c_double(tolcg), C.c_double(rmin), iptr(nd), iptr(ja),
                          iptr(topol), dptr(x), dptr(np.ascontiguousarray(tnoti)), dptr(np.ascontiguousarray(coef1)),
                          dptr(cf2), dptr(scr), C.byref(err), C.byref(res))
    return x, it, err.value, res.value, cf2


def newton_system(orc, pr, psi: np.ndarray | None = None, deltat: float = 10.0, tetaf: float = 1.0):
    """One Newton system (newton.f: NEWUNS, ASSNEW, RHSNEW, CFMATN,
    RHSGRV) at iterate ``psi`` from PTIMEP = psi0, on the full pattern."""
    m, s = pr.mesh, pr.soil
    n, nt = m.n, m.nt
    ns = node_soil(orc, pr)
    ptimep = np.ascontiguousarray(pr.psi0, dtype=np.float64)
    if psi is None:
        psi = ptimep.copy()
        psi[::7] -= 0.7
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    out = {k: np.zeros(n) for k in ("sw", "ckrw", "etai", "dckrw", "detai")}
    orc.L.orc_chnew0(n, dptr(psi), dptr(ns["snodi"]), dptr(ns["pnodi"]), dptr(ns["vg"]), dptr(out["sw"]),
                     dptr(out["ckrw"]), dptr(out["etai"]), dptr(out["dckrw"]), dptr(out["detai"]))
    tet = np.ascontiguousarray(m.tetra, dtype=np.int32)
    ckrwe = np.zeros(nt); etae = np.zeros(nt)
    orc.L.orc_nodelt(nt, iptr(tet), dptr(out["ckrw"]), dptr(ckrwe))
    orc.L.orc_nodelt(nt, iptr(tet), dptr(out["etai"]), dptr(etae))
    topol, ja = M.full_pattern(m.tetra, n)
    topol = np.ascontiguousarray(topol, dtype=np.int32)
    ja = np.ascontiguousarray(ja, dtype=np.int32)
    nterm = len(ja)
    ia = np.repeat(np.arange(n), np.diff(topol)).astype(np.int32)
    tj = np.zeros((nt, 16), dtype=np.int32)
    for k in range(4):
        for l in range(4):
            tj[:, k * 4 + l] = M.csr_lookup(topol, ja, m.tetra[:, k], m.tetra[:, l], n)
    lm = np.zeros(16)
    orc.L.orc_locmas(dptr(lm), int(pr.params.lump))
    perm = [np.ascontiguousarray(s[k], dtype=np.float64) for k in ("permx", "permy", "permz")]
    c1 = np.zeros(nterm); c2 = np.zeros(nterm); c3 = np.zeros(nterm)
    volur = 1.0 / m.volu
    bi, ci, di = (np.ascontiguousarray(a) for a in (m.bi, m.ci, m.di))
    volu = np.ascontiguousarray(m.volu)
    ivol = np.ascontiguousarray(m.ivol, dtype=np.int32)
    pnew = psi  # PNEW = PTNEW for TETAF = 1
    orc.L.orc_assnew(nt, m.ntri, s["nzone"], iptr(tet), iptr(tj), dptr(lm), dptr(c1), dptr(c2), dptr(c3),
                     dptr(perm[0]), dptr(perm[1]), dptr(perm[2]), dptr(ckrwe), dptr(etae), dptr(psi), dptr(ptimep),
                     dptr(pnew), dptr(out["dckrw"]), dptr(out["detai"]), C.c_double(tetaf), C.c_double(deltat),
                     dptr(volu), dptr(volur), iptr(ivol), dptr(bi), dptr(ci), dptr(di))
    asm = dict(c1=c1.copy(), c2=c2.copy(), c3=c3.copy())
    tno = np.zeros(n)
    orc.L.orc_rhsnew(n, iptr(ja), iptr(topol), C.c_double(deltat), dptr(psi), dptr(pnew), dptr(ptimep), dptr(c1),
                     dptr(c2), dptr(tno))
    rhs = tno.copy()
    orc.L.orc_cfmatn(nterm, C.c_double(tetaf), C.c_double(deltat), dptr(c1), dptr(c2), dptr(c3))
    orc.L.orc_rhsgrv(nt, m.ntri, s["nzone"], iptr(tet), dptr(tno), dptr(di), dptr(perm[2]), iptr(ivol),
                     dptr(ckrwe))
    return dict(n=n, nterm=nterm, topol=topol, ja=ja, ia=ia, tetja=tj, lmass=lm, psi=psi, ptimep=ptimep,
                deltat=deltat, tetaf=tetaf, ckrwe=ckrwe, etae=etae, volur=volur, asm=asm, rhs_rhsnew=rhs,
                coef1=c1, tnoti=tno, **out, **ns)


def penalise_full(sysd: dict, rows: np.ndarray):
    """BCNEW's Dirichlet treatment (bcnew.f) on the full pattern."""
    c1 = sysd["coef1"].copy()
    tno = sysd["tnoti"].copy()
    rows = np.asarray(rows, dtype=np.int64)
    for r in rows:
        q = sysd["topol"][r] + int(np.nonzero(sysd["ja"][sysd["topol"][r]:sysd["topol"][r + 1]] == r)[0][0])
        c1[q] = 1e-9 * RMAX
    tno[rows] = 0.0
    return c1, tno, (rows + 1).astype(np.int32)


def oracle_nsyslv(orc, sysd, coef1, tnoti, noddir0=None, itmxcg=500, tolcg=1e-10, rmin=1.7e-100):
    """orc_nsyslv (ISOLV=0); returns (x, niter, err, res, factor)."""
    n, nterm = sysd["n"], sysd["nterm"]
    nd = np.ascontiguousarray(noddir0 if noddir0 is not None else np.zeros(1), dtype=np.int32)
    numdir = 0 if noddir0 is None else len(nd)
    x = np.zeros(n); c2 = np.zeros(nterm); c3 = np.zeros(nterm); scr = np.zeros(max(nterm, 11 * n))
    err = C.c_double(); res = C.c_double(); resini = C.c_double()
    it = orc.L.orc_nsyslv(n, nterm, numdir, iptr(nd), itmxcg, C.c_double(tolcg), C.c_double(rmin),
                          iptr(sysd["ia"]), iptr(sysd["ja"]), iptr(sysd["topol"]),
                          dptr(np.ascontiguousarray(coef1)), dptr(c2), dptr(c3), dptr(scr), dptr(x),
                          dptr(np.ascontiguousarray(tnoti)), C.byref(err), C.byref(res), C.byref(resini))
    return x, it, err.value, res.value, c2
