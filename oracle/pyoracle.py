"""ctypes bindings of the checker libraries (TEST INFRASTRUCTURE ONLY).

* ``Oracle``  -- oracle/liboracle.so, the C restatement (cathy_oracle.c).
* ``RefLib``  -- oracle/_ref/libcathy_ref.so, the reference's own Fortran
  flow routines built by oracle/Makefile from the vendored CATHY sources.
  Fortran convention: every argument by reference, 1-based indices,
  column-major arrays, common blocks exported as ``<name>_`` symbols.  The
  reference is compiled against its CATHY.H (NODMAX=441, MAXSTR=15, so
  NMAX = 7056, MAXZON = 1, MAXTRM = 52111): meshes given to it must fit.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcathy_ref.so")
REF_LARGE_SO = os.path.join(HERE, "_ref", "large", "libcathy_ref_large.so")
REF_BC_SO = os.path.join(HERE, "_ref", "bc", "libcathy_ref_bc.so")

VG_FIELDS = ["PNOT", "N", "M", "RMC", "PSAT", "N1", "M52", "NR", "MM1", "MR", "PSN"]
VG_NFIELD = len(VG_FIELDS)

# reference CATHY.H sizing (pyCATHY/tests/weil_exemple/my_cathy_prj/src/CATHY.H)
REF_NMAX = 441 * 16
REF_MAXSTR = 15
REF_MAXTRM = 52111
REF_ROWMAX = REF_COLMAX = 20      # CATHY.H of the reference build: MAXCEL = 400, MAXRES = 1
REF_MAXCEL = REF_ROWMAX * REF_COLMAX

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_ip)


class OrcParams(C.Structure):
    _fields_ = [
        ("tetaf", C.c_double), ("lump", C.c_int), ("iopt", C.c_int), ("l2norm", C.c_int),
        ("toluns", C.c_double), ("tolswi", C.c_double), ("ernlmx", C.c_double),
        ("ituns", C.c_int), ("ituns1", C.c_int), ("ituns2", C.c_int),
        ("itmxcg", C.c_int), ("tolcg", C.c_double),
        ("deltat", C.c_double), ("dtmin", C.c_double), ("dtmax", C.c_double), ("tmax", C.c_double),
        ("dtmaga", C.c_double), ("dtmagm", C.c_double), ("dtreds", C.c_double), ("dtredm", C.c_double),
        ("pmin", C.c_double), ("pondh_min", C.c_double), ("scf", C.c_double),
        ("ieto", C.c_int), ("free_drainage", C.c_int),
    ]


NETWORK_REAL = ("w1", "w2", "slope1", "slope2", "epl1", "epl2", "kss1", "kss2", "ws1", "ws2", "b1", "y1", "nrc")


class OrcNetwork(C.Structure):
    _fields_ = [("nrow", C.c_int), ("ncol", C.c_int), ("ldc", C.c_int), ("ncell", C.c_int),
                ("qoi", C.POINTER(C.c_int)), ("pout1", C.POINTER(C.c_int)), ("pout2", C.POINTER(C.c_int))] + \
               [(k, C.POINTER(C.c_double)) for k in NETWORK_REAL] + \
               [("tipo_r", C.POINTER(C.c_int)), ("reservr", C.POINTER(C.c_int)), ("n_ha", C.POINTER(C.c_int)),
                ("maxres", C.c_int)] + \
               [(k, C.POINTER(C.c_double)) for k in ("hres", "ares", "h_soglia", "h_fondo", "elevation")]


class Oracle:
    """The C restatement, loaded once."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run make -C oracle liboracle")
        L = C.CDLL(path)
        self.L = L
        L.orc_fvgse.restype = C.c_double
        L.orc_fvgkr.restype = C.c_double
        L.orc_fvgdse.restype = C.c_double
        L.orc_stormb.restype = C.c_double
        L.orc_ressym.restype = C.c_double
        L.orc_run_create.restype = C.c_void_p
        L.orc_run_create.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, _dp, _dp, C.c_int,
                                     _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int,
                                     C.c_int, _dp, _dp, C.POINTER(OrcParams)]
        L.orc_run_free.argtypes = [C.c_void_p]
        L.orc_run_step.argtypes = [C.c_void_p, _dp]
        for f in ("orc_run_get_psi", "orc_run_get_sw"):
            getattr(L, f).argtypes = [C.c_void_p, _dp]
        L.orc_run_get_ifatm.argtypes = [C.c_void_p, _ip]
        L.orc_run_n.argtypes = [C.c_void_p]
        L.orc_run_nterm.argtypes = [C.c_void_p]

    # --- mesh -------------------------------------------------------------
    def triangoli(self, dem, zone, dx, dy, west=0.0, south=0.0, factor=1.0):
        dem = np.ascontiguousarray(dem, dtype=np.float64)
        nrow, ncol = dem.shape
        zone = np.ascontiguousarray(zone if zone is not None else np.ones_like(dem, dtype=np.int32), dtype=np.int32)
        mx = (nrow + 1) * (ncol + 1)
        x = np.zeros(mx); y = np.zeros(mx); z = np.zeros(mx)
        tri = np.zeros(2 * nrow * ncol * 4, dtype=np.int32)
        nn = C.c_int(); nt = C.c_int()
        rc = self.L.orc_triangoli(nrow, ncol, C.c_double(dx), C.c_double(dy), C.c_double(west),
                                  C.c_double(south), dptr(dem), iptr(zone), C.c_double(factor),
                                  C.byref(nn), C.byref(nt), dptr(x), dptr(y), dptr(z), iptr(tri))
        assert rc == 0, rc
        return nn.value, nt.value, x[: nn.value], y[: nn.value], z[: nn.value], tri[: 4 * nt.value].reshape(-1, 4)

    def gen3d(self, nnod, nstr, triang, zratio, base, ivert=0, x=None, y=None, z=None, depth=None, sort=True):
        ntri = triang.shape[0]
        n = nnod * (nstr + 1)
        X = np.zeros(n); Y = np.zeros(n); Z = np.zeros(n)
        X[:nnod] = x; Y[:nnod] = y; Z[:nnod] = z
        tet = np.zeros(3 * ntri * nstr * 5, dtype=np.int32)
        zr = np.ascontiguousarray(zratio, dtype=np.float64)
        dep = np.ascontiguousarray(depth if depth is not None else np.zeros(nnod), dtype=np.float64)
        self.L.orc_gen3d(nnod, nstr, ntri, ivert, iptr(np.ascontiguousarray(triang, dtype=np.int32)),
                         dptr(zr), C.c_double(base), dptr(dep), dptr(X), dptr(Y), dptr(Z), iptr(tet),
                         1 if sort else 0)
        return X, Y, Z, tet.reshape(-1, 5)

    def volbas(self, tetra, x, y, z):
        nt = tetra.shape[0]; n = x.shape[0]
        ai = np.zeros((nt, 4)); bi = np.zeros((nt, 4)); ci = np.zeros((nt, 4)); di = np.zeros((nt, 4))
        volnod = np.zeros(n); volu = np.zeros(nt); volur = np.zeros(nt); ivol = np.zeros(nt, dtype=np.int32)
        rc = self.L.orc_volbas(n, nt, iptr(np.ascontiguousarray(tetra, dtype=np.int32)), dptr(x), dptr(y), dptr(z),
                               dptr(ai), dptr(bi), dptr(ci), dptr(di), dptr(volnod), dptr(volu), dptr(volur),
                               iptr(ivol))
        assert rc == 0
        return ai, bi, ci, di, volnod, volu, volur, ivol

    def area2d(self, nnod, triang, x, y):
        are = np.zeros(nnod)
        rc = self.L.orc_area2d(nnod, triang.shape[0], iptr(np.ascontiguousarray(triang, dtype=np.int32)),
                               dptr(np.ascontiguousarray(x[:nnod])), dptr(np.ascontiguousarray(y[:nnod])), dptr(are))
        assert rc == 0
        return are

    def pattern(self, n, tetra, full=False):
        cap = 32 * n
        topol = np.zeros(n + 1, dtype=np.int32); ja = np.zeros(cap, dtype=np.int32)
        f = self.L.orc_strnew if full else self.L.orc_strpic
        nterm = f(n, tetra.shape[0], iptr(np.ascontiguousarray(tetra, dtype=np.int32)), iptr(topol), iptr(ja), cap)
        assert nterm > 0
        tetja = np.zeros(tetra.shape[0] * 16, dtype=np.int32)
        g = self.L.orc_tetnew if full else self.L.orc_tetpic
        rc = g(tetra.shape[0], iptr(np.ascontiguousarray(tetra, dtype=np.int32)), iptr(topol), iptr(ja), iptr(tetja))
        assert rc == 0
        return topol, ja[:nterm].copy(), tetja.reshape(-1, 4, 4)

    # --- whole run ----------------------------------------------------------
    def route(self, net, deltat: float, st: dict, celtype=None) -> tuple[float, float]:
        """orc_route + orc_altezze on one member's state dict (the keys of
        pycathy_wrapper_amd.routing.RouteGPU.step, plus h_pool_kk /
        h_pool_kkp1 with reservoirs); arrays updated in place.  ``celtype``:
        ALTEZZE's type array (nrow*ncol, by I_BASIN - 1); None is 0
        everywhere, as the reference's time loop passes it.  Returns
        (CU_MAX, AK_MAX); raises where the reference would stop."""
        keep = [np.ascontiguousarray(a) for a in
                (net.qoi.astype(np.int32), net.pout1.astype(np.int32), net.pout2.astype(np.int32))]
        reals = [np.ascontiguousarray(getattr(net, k), dtype=np.float64) for k in NETWORK_REAL]
        on = OrcNetwork(net.nrow, net.ncol, net.ncol, net.ncell, *[iptr(a) for a in keep], *[dptr(a) for a in reals])
        res = net.reservoirs
        hp0 = hp1 = None
        if res is not None:
            rk = {k: np.ascontiguousarray(res[k], dtype=np.int32) for k in ("tipo_r", "reservr", "n_ha")}
            rd = {k: np.ascontiguousarray(res[k], dtype=np.float64)
                  for k in ("hres", "ares", "h_soglia", "h_fondo", "elevation")}
            on.tipo_r, on.reservr, on.n_ha = iptr(rk["tipo_r"]), iptr(rk["reservr"]), iptr(rk["n_ha"])
            on.maxres = int(res["maxres"])
            on.hres, on.ares, on.h_soglia, on.h_fondo, on.elevation = (
                dptr(rd[k]) for k in ("hres", "ares", "h_soglia", "h_fondo", "elevation"))
            hp0, hp1 = st["h_pool_kk"], st["h_pool_kkp1"]
        cu_ak = np.array([0.0, float(st.get("ak_max", 0.0))])
        self.L.orc_route.restype = C.c_int
        rc = self.L.orc_route(C.byref(on), C.c_double(deltat), *[dptr(st[k]) for k in
                              ("surface_water", "q_in_kk", "q_out_kk1", "q_out_kk2", "q_in_kkp1", "q_out_kkp1_1",
                               "q_out_kkp1_2")], dptr(cu_ak), dptr(hp0) if hp0 is not None else None,
                              dptr(hp1) if hp1 is not None else None)
        if rc != 0:
            raise RuntimeError("orc_route: the reference stops here (reservoir level outside its table, "
                               "or a routed lake cell)")
        ct = None if celtype is None else np.ascontiguousarray(celtype, dtype=np.int32)
        self.L.orc_altezze(C.byref(on), iptr(ct) if ct is not None else None, C.c_int(net.nrow * net.ncol),
                           C.c_double(net.dx), C.c_double(net.dy),
                           C.c_double(deltat),
                           *[dptr(st[k]) for k in ("q_in_kk", "q_in_kkp1", "q_out_kk1", "q_out_kk2",
                                                   "q_out_kkp1_1", "q_out_kkp1_2", "volume_kk", "surface_water",
                                                   "volume_kkp1", "h_water_kkp1")],
                           dptr(hp1) if hp1 is not None else None)
        return float(cu_ak[0]), float(cu_ak[1])

    def run_create(self, mesh, soil, psi0, atm_times, atm_vals, hspatm, params: OrcParams):
        keep = dict(
            tetra=np.ascontiguousarray(mesh["tetra"], dtype=np.int32),
            x=np.ascontiguousarray(mesh["x"]), y=np.ascontiguousarray(mesh["y"]), z=np.ascontiguousarray(mesh["z"]),
            arenod=np.ascontiguousarray(mesh["arenod"]),
            **{k: np.ascontiguousarray(soil[k], dtype=np.float64) for k in
               ("permx", "permy", "permz", "elstor", "poros", "vgn", "vgrmc", "vgpsat")},
            psi0=np.ascontiguousarray(psi0, dtype=np.float64),
            at=np.ascontiguousarray(atm_times, dtype=np.float64),
            av=np.ascontiguousarray(atm_vals, dtype=np.float64),
        )
        h = self.L.orc_run_create(mesh["nnod"], mesh["nstr"], mesh["ntri"], iptr(keep["tetra"]), dptr(keep["x"]),
                                  dptr(keep["y"]), dptr(keep["z"]), dptr(keep["arenod"]), soil["nzone"],
                                  dptr(keep["permx"]), dptr(keep["permy"]), dptr(keep["permz"]), dptr(keep["elstor"]),
                                  dptr(keep["poros"]), dptr(keep["vgn"]), dptr(keep["vgrmc"]), dptr(keep["vgpsat"]),
                                  dptr(keep["psi0"]), len(keep["at"]), hspatm, dptr(keep["at"]), dptr(keep["av"]),
                                  C.byref(params))
        assert h
        return OracleRun(self, h, mesh["nnod"] * (mesh["nstr"] + 1), mesh["nnod"])


STAT_NAMES = ["NSTEP", "DELTAT", "TIME", "ITER", "NITERT", "BACK", "STORE1", "STORE2", "DSTORE", "CDSTOR",
              "VIN", "CVIN", "VOUT", "CVOUT", "VINOUT", "VTOT", "ERRAS", "ERREL", "CERRAS", "CAERAS", "PINF",
              "ITRTOT", "ITLIN", "KLSFAI"]


def oracle_etran(orc: "Oracle", n, nnod, nstr, atmpot, z, psi, vegtype0, tab: dict, scf):
    """orc_etran: QTRANIE(n); vegtype0 0-based; tab holds pcana, pcref,
    pcwlt, zroot, pz, omgc arrays."""
    q = np.zeros(n)
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    keep = [f(atmpot), f(z), f(psi), np.ascontiguousarray(vegtype0, dtype=np.int32)] + \
           [f(np.atleast_1d(tab[k])) for k in ("pcana", "pcref", "pcwlt", "zroot", "pz", "omgc")]
    rc = orc.L.orc_etran(C.c_int(n), C.c_int(nnod), C.c_int(nstr), *[dptr(x) if x.dtype == np.float64 else
                                                                      x.ctypes.data_as(C.POINTER(C.c_int))
                                                                      for x in keep],
                         C.c_double(scf), dptr(q))
    if rc != 0:
        raise ValueError("orc_etran: a root zone reaches the base of its column")
    return q


class OracleRun:
    def __init__(self, orc: Oracle, h, n, nnod):
        self.orc, self.h, self.n, self.nnod = orc, h, n, nnod

    def step(self):
        st = np.zeros(24)
        rc = self.orc.L.orc_run_step(self.h, dptr(st))
        return rc, dict(zip(STAT_NAMES, st.tolist()))

    def psi(self):
        p = np.zeros(self.n); self.orc.L.orc_run_get_psi(self.h, dptr(p)); return p

    def sw(self):
        p = np.zeros(self.n); self.orc.L.orc_run_get_sw(self.h, dptr(p)); return p

    def ifatm(self):
        p = np.zeros(self.nnod, dtype=np.int32); self.orc.L.orc_run_get_ifatm(self.h, iptr(p)); return p

    def close(self):
        if self.h:
            self.orc.L.orc_run_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RefLib:
    """The reference's Fortran routines (by-reference args, 1-based)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run make -C oracle ref")
        self.L = C.CDLL(path)
        # array bounds of the build: the weil CATHY.H, or the sized header's
        # sizes.json beside the library (oracle/mkcathyh.py)
        sz = os.path.join(os.path.dirname(os.path.abspath(path)), "sizes.json")
        if os.path.exists(sz):
            import json
            with open(sz) as f:
                d = json.load(f)
            self.sizes = d
            self.nmax, self.maxstr, self.maxzon, self.maxtrm = d["nmax"], d["maxstr"], d["maxzon"], d["maxtrm"]
            self.npmax, self.nqmax = d.get("npmax", 1), d.get("nqmax", 1)
            self.nsfmax, self.nnsfmx, self.maxveg = d.get("nsfmax", 1), d.get("nnsfmx", 1), d.get("maxveg", 1)
        else:
            self.sizes = {}
            self.nmax, self.maxstr, self.maxzon, self.maxtrm = REF_NMAX, REF_MAXSTR, 1, REF_MAXTRM
            self.npmax = self.nqmax = self.nsfmax = self.nnsfmx = self.maxveg = 1
        self.vg = np.ctypeslib.as_array((C.c_double * (VG_NFIELD * self.nmax)).in_dll(self.L, "vg_"))
        # /CHORD/PMIN,... : PMIN is the first REAL*8 of the block
        self.chord = np.ctypeslib.as_array((C.c_double * 1).in_dll(self.L, "chord_"))

    def set_vg(self, vg_orc: np.ndarray, n: int):
        """Copy an oracle-layout vg (VG_NFIELD*n) into the NMAX-strided /VG/."""
        assert n <= self.nmax
        for f in range(VG_NFIELD):
            self.vg[f * self.nmax: f * self.nmax + n] = vg_orc[f * n:(f + 1) * n]

    def get_vg(self, n: int) -> np.ndarray:
        out = np.zeros(VG_NFIELD * n)
        for f in range(VG_NFIELD):
            out[f * n:(f + 1) * n] = self.vg[f * self.nmax: f * self.nmax + n]
        return out

    def zone_table(self, tab) -> np.ndarray:
        """A (layer, zone) soil table as the reference's (MAXSTR, MAXZON)
        column-major array (PERMX, POROS, VGNCELL, ...)."""
        t = np.asarray(tab, dtype=np.float64)
        nstr, nz = t.shape
        assert nstr <= self.maxstr and nz <= self.maxzon, (t.shape, self.maxstr, self.maxzon)
        out = np.zeros((max(nz, 1), self.maxstr))
        out[:nz, :nstr] = t.T
        return out.reshape(-1)

    def __getattr__(self, name):
        return getattr(self.L, name.lower() + "_")


def byref_i(v):
    return C.byref(C.c_int(int(v)))


def byref_d(v):
    return C.byref(C.c_double(float(v)))
