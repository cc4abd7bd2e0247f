"""A pyCATHY project directory read by the reference's own DATIN (TEST
INFRASTRUCTURE ONLY: the checker of tests/ and the golden maker).

DATIN (datin.f:6-674) reads parm, soil, ic, the DEM and zone rasters, the
root map, dem_parameters and, for ISIMGR = 2, the reservoir files, QOI_SN
and the thirteen DTM rasters of the drainage network; it builds the
surface triangulation (TRIANGOLI), the cell numbering (INDEX_DEM) and the
cell elevations (ASSIGN_DEM).  The reference's main program opens those
files on the units of COMMON /IOUNIT/ from cathy.fnames; here each unit is
given a number and connected by REFUNIT_OPENN (oracle/refatm_io.f) to the
file pyCATHY's cathy.fnames names for it, and the output units DATIN echoes
to are connected to files in a scratch directory.

``RefProject(ref, root)`` calls DATIN once and keeps every argument array;
``rivernetwork()`` and ``surfwater()`` view the COMMON blocks DATIN and
ASSIGN_DEM fill.
"""
from __future__ import annotations

import ctypes as C
import os
import tempfile

import numpy as np

# COMMON /IOUNIT/ in IOUNITS.H order
IOUNIT_NAMES = ("IFN IIN1 IIN2 IIN3 IIN4 IIN5 IIN6 IIN7 IIN8 IIN9 IIN10 IIN11 IIN16 IIN17 IIN18 IIN20 IIN21 IIN22 "
                "IIN23 IIN25 IIN26 IIN27 IIN28 IIN29 IIN30 IIN31 IIN32 IIN33 IIN34 IIN35 IIN36 IIN37 IIN38 IIN39 "
                "IIN50 IIN51 IIN60 IIN61 IIN62 IIN63 IIN64 ITERM IOUT1 IOUT2 IOUT3 IOUT4 IOUT5 IOUT6 IOUT7 IOUT8 "
                "IOUT9 IOUT10 IOUT11 IOUT12 IOUT13 IOUT14 IOUT15 IOUT16 IOUT17 IOUT18 IOUT19 IOUT20 IOUT30 IOUT31 "
                "IOUT32 IOUT36 IOUT40 IOUT41 IOUT42 IOUT43 IOUT44 IOUT50 IOUT51 IOUT57 IOUT80 IOUT81 "
                "IOUTPT").split()

# the files pyCATHY's cathy.fnames connects to DATIN's input units
UNIT_FILES = {
    "IIN1": "input/parm", "IIN3": "input/root_map", "IIN4": "input/soil", "IIN5": "input/ic",
    "IIN10": "prepro/dem", "IIN11": "input/dem_parameters", "IIN17": "input/posizione_serb",
    "IIN18": "input/livelli_iniz_s", "IIN20": "prepro/lakes_map", "IIN21": "prepro/zone",
    "IIN23": "prepro/qoi_a", "IIN25": "prepro/dtm_w_1", "IIN26": "prepro/dtm_w_2",
    "IIN27": "prepro/dtm_p_outflow_1", "IIN28": "prepro/dtm_p_outflow_2", "IIN29": "prepro/dtm_local_slope_1",
    "IIN30": "prepro/dtm_local_slope_2", "IIN31": "prepro/dtm_epl_1", "IIN32": "prepro/dtm_epl_2",
    "IIN33": "prepro/dtm_kSs1_sf_1", "IIN34": "prepro/dtm_kSs1_sf_2", "IIN35": "prepro/dtm_Ws1_sf_1",
    "IIN36": "prepro/dtm_Ws1_sf_2", "IIN37": "prepro/dtm_b1_sf", "IIN38": "prepro/dtm_y1_sf",
    "IIN39": "prepro/dtm_nrc",
}

# DATIN's dummy arguments (datin.f:6-20): i/d/l scalars, I/D arrays
DATIN_ARGS = [
    ("ISIMGR", "i"), ("IVERT", "i"), ("ISP", "i"), ("WTPOSITION", "d"), ("BASE", "d"), ("ZRATIO", "D"),
    ("TRIANG", "I"), ("X", "D"), ("Y", "D"), ("Z", "D"), ("PERMX", "D"), ("PERMY", "D"), ("PERMZ", "D"),
    ("ELSTOR", "D"), ("POROS", "D"), ("VGNCELL", "D"), ("VGRMCCELL", "D"), ("VGPSATCELL", "D"), ("CONTR", "I"),
    ("NODVP", "I"), ("ID_QOUT", "I"), ("TIMPRT", "D"), ("PONDNOD", "D"), ("PTIMEP", "D"), ("TETAF", "d"),
    ("DELTAT", "d"), ("DTMIN", "d"), ("DTMAX", "d"), ("TMAX", "d"), ("DTMAGA", "d"), ("DTMAGM", "d"),
    ("DTREDS", "d"), ("DTREDM", "d"), ("ITUNS", "i"), ("ITUNS1", "i"), ("ITUNS2", "i"), ("TOLUNS", "d"),
    ("TOLSWI", "d"), ("ERNLMX", "d"), ("ITMXCG", "i"), ("TOLCG", "d"), ("KSLOPE", "i"), ("LUMP", "i"),
    ("IPEAT", "i"), ("IVGHU", "i"), ("NLKP", "i"), ("VTKF", "i"), ("IOPT", "i"), ("ISOLV", "i"), ("IPRT1", "i"),
    ("IPRT", "i"), ("IPOND", "i"), ("INDP", "i"), ("NNOD", "i"), ("NTRI", "i"), ("NSTR", "i"), ("NZONE", "i"),
    ("NVEG", "i"), ("N1", "i"), ("NR", "i"), ("NUMVP", "i"), ("NUM_QOUT", "i"), ("NPRT", "i"), ("N", "i"),
    ("NT", "i"), ("ISFONE", "i"), ("ISFCVG", "i"), ("DUPUIT", "i"), ("L2NORM", "i"), ("NLRELX", "i"),
    ("OMEGA", "d"), ("PONDH_MIN", "d"), ("FL3D", "l"), ("SURF", "l"), ("DEM", "l"), ("GRID", "l"), ("NROW", "i"),
    ("NCOL", "i"), ("NCELNL", "i"), ("NCELL", "i"), ("DOSTEP", "i"), ("NCELL_COARSE", "i"), ("DEM_MAP", "D"),
    ("ZONE", "I"), ("LAKES_MAP", "I"), ("INDEX", "I"), ("INDEX_WITH_LAKES", "I"), ("CELL", "I"),
    ("CELLCOL", "I"), ("CELLROW", "I"), ("TP2D", "I"), ("NODI", "I"), ("TIPO_R", "I"), ("RESERVR", "I"),
    ("N_HA", "I"), ("CELLCOL_WL", "I"), ("CELLROW_WL", "I"), ("BASE_MAP", "D"), ("DEPTH", "D"), ("ELTRIA", "D"),
    ("NCOUT", "i"), ("TRAFLAG", "i"), ("TRANSP", "l"), ("VELREC", "i"), ("VEG_TYPE", "I"),
]

# COMMON /RIVERNETWORK/ (RIVERNETWORK.H): name, kind, length in units of the grid (COLMAX*ROWMAX) or 1
RIVER_LAYOUT = [("NORTH", 1), ("SOUTH", 1), ("EAST", 1), ("WEST", 1), ("FACTOR", 1), ("DELTA_X", 1),
                ("DELTA_Y", 1), ("DTM_W_1", "g"), ("DTM_W_2", "g"), ("DTM_LOCAL_SLOPE_1", "g"),
                ("DTM_LOCAL_SLOPE_2", "g"), ("DTM_EPL_1", "g"), ("DTM_EPL_2", "g"), ("DTM_KSS1_SF_1", "g"),
                ("DTM_KSS1_SF_2", "g"), ("DTM_WS1_SF_1", "g"), ("DTM_WS1_SF_2", "g"), ("DTM_B1_SF", "g"),
                ("DTM_Y1_SF", "g"), ("DTM_NRC", "g")]   # then DTM_P_OUTFLOW_1/2 (INTEGER grids) and QOI_SN(MAXCEL)

# COMMON /SURFWATER/ (SURFWATER.H): REAL*8 members, length in MAXCEL / MAXRES units
SURF_LAYOUT = [("AK_MAX", 1), ("AK_MAX_SAV", 1), ("AK_MAX_P", 1), ("CUTRGT", 1), ("CU_MAX", 1),
               ("ELEVATION", "c"), ("ELEVATION_WITH_LAKES", "c"), ("CELLCOARSE", "c"), ("H_WATER_KKP1_SN", "c"),
               ("SURFACE_WATER_SN", "c"), ("Q_IN_KK_SN", "c"), ("Q_IN_KK_SN_SAV", "c"), ("Q_IN_KK_SN_P", "c"),
               ("Q_IN_KKP1_SN", "c"), ("Q_OUT_KK_SN_1", "c"), ("Q_OUT_KK_SN_1_SAV", "c"), ("Q_OUT_KK_SN_2", "c"),
               ("Q_OUT_KK_SN_2_SAV", "c"), ("Q_OUT_KK_SN_1_P", "c"), ("Q_OUT_KK_SN_2_P", "c"),
               ("Q_OUT_KKP1_SN_1", "c"), ("Q_OUT_KKP1_SN_2", "c"), ("VOLUME_KK_SN", "c"), ("VOLUME_KK_SN_SAV", "c"),
               ("VOLUME_KK_SN_P", "c"), ("VOLUME_KKP1_SN", "c"), ("H_POOL_KK_VEC", "r"), ("H_POOL_KK_VEC_SAV", "r"),
               ("H_POOL_KK_VEC_P", "r"), ("H_POOL_KKP1_VEC", "r"), ("HRES", "r30"), ("ARES", "r30"),
               ("H_SOGLIA", "r"), ("H_FONDO", "r")]


def _encode(path: str) -> np.ndarray:
    return np.frombuffer(os.fsencode(path), dtype=np.uint8).astype(np.int32)


class RefProject:
    """DATIN on the project in ``root``; ``a`` holds its arguments."""

    def __init__(self, ref, root: str):
        self.ref = ref
        L = ref.L
        sz = ref.sizes
        assert sz, "the reference build has no sizes.json (oracle/mkcathyh.py)"
        self.rowmax, self.colmax, self.maxcel = sz["rowmax"], sz["colmax"], sz["maxcel"]
        self.maxres = sz.get("maxres", 1)
        big = max(sz["nmax"], sz["ntemax"] * 5, 5 * self.maxcel, (self.rowmax + 1) * (self.colmax + 1),
                  sz["maxstr"] * sz["maxzon"], 4 * 2 * self.maxcel) + 64
        a = {}
        for name, kind in DATIN_ARGS:
            if kind in "il":
                a[name] = np.zeros(1, dtype=np.int32)
            elif kind == "d":
                a[name] = np.zeros(1)
            elif kind == "I":
                a[name] = np.zeros(big, dtype=np.int32)
            else:
                a[name] = np.zeros(big)
        self.a = a
        iou = np.ctypeslib.as_array((C.c_int * len(IOUNIT_NAMES)).in_dll(L, "iounit_"))
        self._tmp = tempfile.mkdtemp(prefix="refprj_")
        units = {}
        for k, nm in enumerate(IOUNIT_NAMES):
            iou[k] = 6 if nm == "ITERM" else 19 + k
            units[nm] = int(iou[k])
        self.units = units
        opened = []
        try:
            for nm, rel in UNIT_FILES.items():
                path = os.path.join(root, rel)
                if not os.path.exists(path):
                    continue
                self._open(units[nm], path, 0)
                opened.append(units[nm])
            for nm in IOUNIT_NAMES:
                if nm.startswith("IOUT"):
                    self._open(units[nm], os.path.join(self._tmp, nm.lower()), 1)
                    opened.append(units[nm])
            L.datin_(*[a[name].ctypes.data_as(C.c_void_p) for name, _ in DATIN_ARGS])
        finally:
            for u in opened:
                L.refunit_closen_(C.byref(C.c_int(u)))

    def _open(self, unit: int, path: str, mode: int) -> None:
        codes = _encode(path)
        ierr = C.c_int(0)
        self.ref.L.refunit_openn_(C.byref(C.c_int(unit)), codes.ctypes.data_as(C.POINTER(C.c_int)),
                                  C.byref(C.c_int(len(codes))), C.byref(C.c_int(mode)), C.byref(ierr))
        if ierr.value != 0:
            raise OSError(f"REFUNIT_OPENN({path}) failed: IOSTAT {ierr.value}")

    def scalar(self, name: str):
        v = self.a[name][0]
        return int(v) if self.a[name].dtype == np.int32 else float(v)

    def grid(self, name: str, dtype=None) -> np.ndarray:
        """A (ROWMAX, *) argument (DEM_MAP, ZONE, LAKES_MAP, INDEX, ...: the
        readers store RAST(IROW, ICOL) with IROW = 1 the file's first row)
        as (NROW, NCOL) in the file's row order."""
        nrow, ncol = self.scalar("NROW"), self.scalar("NCOL")
        v = self.a[name][:self.rowmax * ncol].reshape(ncol, self.rowmax).T[:nrow, :ncol]
        return v.astype(dtype) if dtype else v.copy()

    def rivernetwork(self) -> dict:
        """COMMON /RIVERNETWORK/: the DTM grids as (NROW, NCOL) raster
        order (they are stored (COLMAX, ROWMAX): column index first), the
        two outflow-pointer grids and QOI_SN(NCELL)."""
        g = self.colmax * self.rowmax
        nd = 7 + 13 * g
        blk = (C.c_char * (8 * nd + 4 * (2 * g + self.maxcel))).in_dll(self.ref.L, "rivernetwork_")
        dv = np.frombuffer(blk, dtype=np.float64, count=nd)
        iv = np.frombuffer(blk, dtype=np.int32, count=2 * g + self.maxcel, offset=8 * nd)
        nrow, ncol = self.scalar("NROW"), self.scalar("NCOL")
        # RAST_INPUT_REAL / _INT store RAST(ICOL, IROW) with IROW = 1 the
        # raster's last (southern) row: flip to the file's row order
        grid = lambda v: v.reshape(self.rowmax, self.colmax)[:nrow, :ncol][::-1].copy()
        out, pos = {}, 0
        for name, ln in RIVER_LAYOUT:
            if ln == 1:
                out[name] = float(dv[pos]); pos += 1
            else:
                out[name] = grid(dv[pos:pos + g])
                pos += g
        out["DTM_P_OUTFLOW_1"] = grid(iv[:g])
        out["DTM_P_OUTFLOW_2"] = grid(iv[g:2 * g])
        out["QOI_SN"] = iv[2 * g:2 * g + self.scalar("NCELL")].copy()
        return out

    def surfwater(self) -> dict:
        """COMMON /SURFWATER/ as views (writable)."""
        c, r = self.maxcel, self.maxres
        lens = {1: 1, "c": c, "r": r, "r30": 30 * r}
        tot = sum(lens[ln] for _, ln in SURF_LAYOUT)
        arr = np.ctypeslib.as_array((C.c_double * tot).in_dll(self.ref.L, "surfwater_"))
        out, pos = {}, 0
        for name, ln in SURF_LAYOUT:
            out[name] = arr[pos:pos + lens[ln]]
            pos += lens[ln]
        return out

    def close(self) -> None:
        import shutil
        if self._tmp:
            shutil.rmtree(self._tmp, ignore_errors=True)
            self._tmp = None
