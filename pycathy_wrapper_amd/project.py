"""A pyCATHY project directory as a device run: the input files CATHY's
DATIN and INITAL read, read the same way, into a ``configs.Problem``.

Files (all Fortran list-directed text, read with bcfiles' record reader;
paths relative to the project directory):

* ``input/parm`` -- solver and output parameters (datin.f:80-122, the
  READ sequence of unit IIN1; counts read earlier size later lists);
* ``input/soil`` -- PMIN, IPEAT SCF, CBETA0 CANG, one Feddes line per
  vegetation type, IVGHU, the Huyakorn / Brooks-Corey lines, then
  PERMX PERMY PERMZ ELSTOR POROS VGNCELL VGRMCCELL VGPSATCELL for every
  layer and zone (datin.f:421-518);
* ``input/ic`` -- INDP IPOND, WTPOSITION for INDP 3/4, PTIMEP for INDP
  0/1, PONDNOD for IPOND 1/2 (datin.f:377-403); the initial head is then
  formed as INITAL does (ICVHE for INDP 2, ICVHWT 3, ICVDWT 4;
  inital.f:93-103);
* ``input/atmbc`` -- HSPATM IETO, then TIME and one or NNOD values per
  record (atmone.f:41-58);
* ``input/dem_parameters``, ``prepro/dem``, ``prepro/zone`` (optional),
  ``input/root_map`` (optional) -- the DEM mesh (datin.f:185-232; the
  vegetation map becomes one type per surface node through TRIANGOLI's
  cell-to-node rule, truncated to an integer);
* ``input/nansfdirbc``, ``input/nansfneubc``, ``input/sfbc`` -- the
  boundary series (bcfiles).

A project with ISIMGR = 2 (surface routing coupled) runs with the coupled
atmospheric switching -- SWITCH in CONVER, ADRSTN at step starts and
back-steps (conver.f:61-71, cathy_main.f:2942-2946, bkstep.f:135-139) --
and PONDH_MIN from the parm file.  Where the project holds its drainage
network (prepro/qoi_a, the dtm_* rasters, lakes_map and posizione_serb:
``read_routing``), a ponded surface is routed as the reference's time
loop routes it (``Problem.routing``, cathy_gpu_flow3d_set_routing);
without one, a step that ends ponded stops the run (status 3).  A project
that starts ponded (IPOND != 0) starts with its ponding heads and routes
from the first step (``Problem.routing["pond0"]``); without a drainage
network it is refused.  ``subsurface_only`` runs
such a project as ISIMGR = 1 instead (SWITCH_OLD, PONDH_MIN = 1e10,
datin.f:120).
"""
from __future__ import annotations

import os

import numpy as np

from . import bcfiles as B
from . import configs as CF
from . import mesh as M
from .bcfiles import _ListDirected


def _reader(path: str) -> _ListDirected:
    with open(path) as f:
        return _ListDirected(f.read(), os.path.basename(path))


PARM_LINES = [
    (("iprt1", "ncout", "traflag"), "iii"),
    (("isimgr", "pondh_min", "velrec"), "idi"),
    (("kslope", "tolksl"), "id"),
    (("pkrl", "pkrr", "psel", "pser"), "dddd"),
    (("pdse1l", "pdse1r", "pdse2l", "pdse2r"), "dddd"),
    (("isfone", "isfcvg", "dupuit"), "iii"),
    (("tetaf", "lump", "iopt"), "dii"),
    (("nlrelx", "omega"), "id"),
    (("l2norm", "toluns", "tolswi", "ernlmx"), "iddd"),
    (("ituns", "ituns1", "ituns2"), "iii"),
    (("isolv", "itmxcg", "tolcg"), "iid"),
    (("deltat", "dtmin", "dtmax", "tmax"), "dddd"),
    (("dtmaga", "dtmagm", "dtreds", "dtredm"), "dddd"),
]


def _num(t: str, kind: str):
    return int(t) if kind == "i" else float(t.replace("d", "e").replace("D", "e"))


def _counted(rd: _ListDirected, lead: int, at: int) -> list[str]:
    """One list-directed READ of ``lead`` items followed by as many more as
    its item ``at`` says (e.g. IPRT, VTKF, NPRT, (TIMPRT(I), I=1,NPRT))."""
    toks: list[str] = []
    while len(toks) < lead:
        if rd.eof():
            raise EOFError(f"{rd.name}: file ends inside a counted list")
        toks += rd._tokens(rd.lines[rd.pos])
        rd.pos += 1
    need = lead + int(toks[at])
    while len(toks) < need:
        if rd.eof():
            raise EOFError(f"{rd.name}: file ends inside a counted list")
        toks += rd._tokens(rd.lines[rd.pos])
        rd.pos += 1
    return toks[:need]


def read_parm(path: str) -> dict:
    """input/parm in DATIN's READ order (datin.f:80-122); PONDH_MIN becomes
    1e10 for ISIMGR <= 1 as DATIN sets it (datin.f:120)."""
    rd = _reader(path)
    out = {}
    for names, kinds in PARM_LINES:
        for nm, k, t in zip(names, kinds, rd.read(len(names))):
            out[nm] = _num(t, k)
    t = _counted(rd, 3, 2)                      # IPRT VTKF NPRT (TIMPRT(I),I=1,NPRT)
    out["iprt"], out["vtkf"], out["nprt"] = int(t[0]), int(t[1]), int(t[2])
    out["timprt"] = [_num(v, "d") for v in t[3:]]
    t = _counted(rd, 1, 0)                      # NUMVP (NODVP(I),I=1,NUMVP)
    out["numvp"], out["nodvp"] = int(t[0]), [int(v) for v in t[1:]]
    out["nr"] = int(rd.read_int(1)[0])          # NR, then CONTR(1..NR) if NR != 0
    out["contr"] = [int(v) for v in rd.read_int(out["nr"])] if out["nr"] else []
    t = _counted(rd, 1, 0)                      # NUM_QOUT (ID_QOUT(I),I=1,NUM_QOUT)
    out["num_qout"], out["id_qout"] = int(t[0]), [int(v) for v in t[1:]]
    if out["isimgr"] <= 1:
        out["pondh_min"] = 1.0e10
    return out


def read_soil(path: str, nstr: int, nzone: int, nveg: int = 1) -> dict:
    """input/soil in DATIN's READ order (datin.f:421-518).  Returns the
    scalars, the Feddes table (one entry per vegetation type) and the
    (layer, zone) tables in ``configs.Problem.soil`` form."""
    rd = _reader(path)
    f = lambda k: [float(v) for v in rd.read_float(k)]
    s = {}
    s["pmin"] = f(1)[0]
    ipeat, scf = rd.read(2)
    s["ipeat"], s["scf"] = int(ipeat), float(scf)
    s["cbeta0"], s["cang"] = f(2)
    veg = np.array([f(6) for _ in range(nveg)])
    s["veg"] = {k: veg[:, j].copy() for j, k in enumerate(("pcana", "pcref", "pcwlt", "zroot", "pz", "omgc"))}
    s["ivghu"] = int(rd.read_int(1)[0])
    if s["ivghu"] not in (0,):
        raise ValueError(f"{path}: IVGHU = {s['ivghu']}: the device step has van Genuchten (IVGHU = 0) only")
    if s["ipeat"] != 0:
        raise ValueError(f"{path}: IPEAT = 1 (peat deformation) is not supported")
    s["hu"] = f(5) + f(1) + f(2)          # HUALFA..HUSWR, HUN, HUA HUB
    s["bc"] = f(3)                        # BCBETA BCRMC BCPSAT
    keys = ("permx", "permy", "permz", "elstor", "poros", "vgn", "vgrmc", "vgpsat")
    tab = {k: np.zeros((nstr, nzone)) for k in keys}
    for i in range(nstr):
        for j in range(nzone):
            vals = f(8)
            for k, v in zip(keys, vals):
                tab[k][i, j] = v
    s["table"] = dict(nzone=nzone, **tab)
    return s


def read_ic(path: str, n: int, nnod: int) -> dict:
    """input/ic (datin.f:377-403)."""
    rd = _reader(path)
    indp, ipond = (int(v) for v in rd.read_int(2))
    out = dict(indp=indp, ipond=ipond, wtposition=None, ptimep=None, pondnod=None)
    if indp in (3, 4):
        out["wtposition"] = float(rd.read_float(1)[0])
    if indp == 0:
        out["ptimep"] = np.full(n, float(rd.read_float(1)[0]))
    elif indp == 1:
        out["ptimep"] = rd.read_float(n)
    if ipond == 1:
        out["pondnod"] = np.full(nnod, float(rd.read_float(1)[0]))
    elif ipond == 2:
        out["pondnod"] = rd.read_float(nnod)
    return out


def initial_head(ic: dict, mesh, surf: bool) -> np.ndarray:
    """PTIMEP as DATIN and INITAL form it: ponding heads over the surface
    (datin.f:390-402, only with the surface coupled), then ICVHE (INDP 2),
    ICVHWT (3) or ICVDWT (4) (inital.f:93-103)."""
    n, nnod, nstr = mesh.n, mesh.nnod, mesh.nstr
    z = mesh.z
    ipond = ic["ipond"] if surf else 0
    pt = np.zeros(n) if ic["ptimep"] is None else ic["ptimep"].copy()
    pond = np.zeros(nnod) if ic["pondnod"] is None else ic["pondnod"]
    if ipond in (1, 2):
        sel = pond > 0.0
        pt[:nnod][sel] = pond[sel]
    mult = 0.0 if ipond == 0 else 1.0
    zs = np.tile(z[:nnod], nstr + 1)
    if ic["indp"] == 2:          # icvhe.f: hydrostatic, water table at the surface (plus ponding)
        pt = np.tile(z[:nnod] + mult * pond, nstr + 1) - z
    elif ic["indp"] == 3:        # icvhwt.f: water table WTPOSITION above the base
        pt = ic["wtposition"] - z + np.tile(z[nstr * nnod:], nstr + 1)
    elif ic["indp"] == 4:        # icvdwt.f: water table WTPOSITION below the surface
        pt = zs - z - ic["wtposition"]
    elif ic["indp"] not in (0, 1):
        raise ValueError(f"INDP = {ic['indp']} is not an initial-condition option")
    return pt


def read_atmbc(path: str, nnod: int):
    """input/atmbc (atmone.f:41-58): HSPATM, IETO, record times and values
    (one per record for HSPATM != 0, else NNOD)."""
    rd = _reader(path)
    hspatm, ieto = (int(v) for v in rd.read_int(2))
    w = 1 if hspatm != 0 else nnod
    times, vals = [], []
    while True:
        while not rd.eof() and not rd._tokens(rd.lines[rd.pos]):
            rd.pos += 1
        if rd.eof():
            break
        times.append(float(rd.read_float(1)[0]))
        vals.append(rd.read_float(w))
    if not times:
        raise ValueError(f"{path}: no atmospheric record")
    return hspatm, ieto, np.array(times), np.concatenate(vals)


def read_raster(path: str):
    """A GRASS ASCII raster as the reference's RAST_INPUT_* read it: the six
    header lines (north, south, east, west, rows, cols), then one list-
    directed READ per row, top row first (rast_input_dem.f:48-50): each row
    starts on a new line, takes cols values (across lines if need be) and
    drops what its last line holds beyond them."""
    with open(path) as f:
        lines = f.read().splitlines()
    hdr = {}
    for ln in lines[:6]:
        k, v = ln.split(":", 1)
        hdr[k.strip()] = float(v)
    nrow, ncol = int(hdr["rows"]), int(hdr["cols"])
    out = np.empty((nrow, ncol))
    li = 6
    for r in range(nrow):
        vals = []
        while len(vals) < ncol:
            if li >= len(lines):
                raise ValueError(f"{path}: the file ends at row {r + 1} of the {nrow} x {ncol} its header "
                                 f"declares (the reference's read runs past the end)")
            vals.extend(lines[li].split())
            li += 1
        out[r] = np.array(vals[:ncol], dtype=np.float64)
    return out, hdr


def read_dem_parameters(path: str) -> dict:
    """input/dem_parameters (datin.f:193-205): DELTA_X DELTA_Y, FACTOR,
    DOSTEP, NZONE NSTR N1, IVERT ISP BASE, then ZRATIO(1..NSTR)."""
    toks = []
    for t in open(path).read().split():
        try:
            toks.append(float(t))
        except ValueError:
            break
    dx, dy, factor, dostep, nzone, nstr, n1, ivert, isp, base = toks[:10]
    nstr = int(nstr)
    if len(toks) < 10 + nstr:
        raise ValueError(f"{path}: {len(toks) - 10} layer ratios for NSTR = {nstr}")
    return dict(dx=dx, dy=dy, factor=factor, dostep=int(dostep), nzone=int(nzone), nstr=nstr, n1=int(n1),
                ivert=int(ivert), isp=int(isp), base=base, zratio=np.array(toks[10:10 + nstr]))


ROUTING_FILES = {"qoi_a": "prepro/qoi_a", "lakes_map": "prepro/lakes_map",
                 "posizione_serb": "input/posizione_serb",
                 "w1": "prepro/dtm_w_1", "w2": "prepro/dtm_w_2", "pout1": "prepro/dtm_p_outflow_1",
                 "pout2": "prepro/dtm_p_outflow_2", "slope1": "prepro/dtm_local_slope_1",
                 "slope2": "prepro/dtm_local_slope_2", "epl1": "prepro/dtm_epl_1", "epl2": "prepro/dtm_epl_2",
                 "kss1": "prepro/dtm_kSs1_sf_1", "kss2": "prepro/dtm_kSs1_sf_2", "ws1": "prepro/dtm_Ws1_sf_1",
                 "ws2": "prepro/dtm_Ws1_sf_2", "b1": "prepro/dtm_b1_sf", "y1": "prepro/dtm_y1_sf",
                 "nrc": "prepro/dtm_nrc"}


def read_routing(root: str, dem: np.ndarray, zone: np.ndarray, dx: float, dy: float, dostep: int) -> dict:
    """The drainage network and cell numbering DATIN reads and builds for
    ISIMGR = 2 (datin.f:216-225, 287-370): the reservoir count of
    input/posizione_serb (IIN17; none are supported), prepro/lakes_map
    (RAST_INPUT_LZ; no lakes), QOI_SN from prepro/qoi_a (IIN23: a header
    line, then NCELL values), the thirteen DTM rasters (RAST_INPUT_REAL /
    _INT store RAST(ICOL, IROW) with IROW = 1 the raster's last row: the
    ``routing.Network`` layout), INDEX / INDEX_WITH_LAKES (INDEX_DEM: the
    grid position (I-1)*NCOL + J of a cell with ZONE /= 0 and no lake, else
    0) and CELL (TRIANGOLI, ``mesh.dem_cells``)."""
    from .routing import Network
    f = {k: os.path.join(root, v) for k, v in ROUTING_FILES.items()}
    for k, p in f.items():
        if not os.path.exists(p):
            raise FileNotFoundError(f"project {root}: {ROUTING_FILES[k]} missing (the drainage network)")
    if dostep != 1:
        raise ValueError(f"DOSTEP = {dostep}: the coupled step takes the full-resolution surface only")
    nres = int(_reader(f["posizione_serb"]).read_int(1)[0])
    if nres != 0:
        raise ValueError(f"{nres} reservoirs in input/posizione_serb: the coupled step routes none")
    lakes = read_raster(f["lakes_map"])[0].astype(np.int32)
    if np.any(lakes > 0):
        raise ValueError("prepro/lakes_map holds lakes: the coupled step routes none")
    nrow, ncol = dem.shape
    index_wl = np.where(zone != 0, np.arange(1, nrow * ncol + 1).reshape(nrow, ncol), 0).astype(np.int32)
    index = np.where(lakes <= 0, index_wl, 0).astype(np.int32)
    ncell = int((zone != 0).sum())
    with open(f["qoi_a"]) as fh:
        toks = fh.read().split("\n", 1)[1].split()
    if len(toks) < ncell:
        raise ValueError(f"prepro/qoi_a: {len(toks)} values for NCELL = {ncell}")
    qoi = np.array(toks[:ncell], dtype=np.int64).astype(np.int32)
    grids = {}
    for k in ("w1", "w2", "slope1", "slope2", "epl1", "epl2", "kss1", "kss2", "ws1", "ws2", "b1", "y1", "nrc",
              "pout1", "pout2"):
        r, hdr = read_raster(f[k])
        if r.shape != (nrow, ncol):
            raise ValueError(f"{ROUTING_FILES[k]}: {r.shape} for a {nrow} x {ncol} DEM")
        r = r[::-1].copy()
        grids[k] = r.astype(np.int32) if k.startswith("pout") else r
    net = Network(nrow=nrow, ncol=ncol, qoi=qoi, dx=dx, dy=dy, **grids)
    return dict(nrow=nrow, ncol=ncol, dx=dx, dy=dy, net=net, index=index, index_wl=index_wl, lakes=lakes,
                cell=M.dem_cells(dem), dem=dem, ncell=ncell)


def _exists(*parts) -> str | None:
    p = os.path.join(*parts)
    return p if os.path.exists(p) else None


def load_project(root: str, subsurface_only: bool = False, files: dict | None = None) -> CF.Problem:
    """The project in ``root`` (pyCATHY's layout: input/, prepro/) as a
    Problem for ``gpu.FlowGPU``.  ``files`` may name files elsewhere (keys
    parm, soil, ic, atmbc, dem_parameters, dem, zone, root_map, nansfdirbc,
    nansfneubc, sfbc)."""
    fl = dict(parm=_exists(root, "input", "parm"), soil=_exists(root, "input", "soil"),
              ic=_exists(root, "input", "ic"), atmbc=_exists(root, "input", "atmbc"),
              dem_parameters=_exists(root, "input", "dem_parameters"), dem=_exists(root, "prepro", "dem"),
              zone=_exists(root, "prepro", "zone"), root_map=_exists(root, "input", "root_map"),
              nansfdirbc=_exists(root, "input", "nansfdirbc"), nansfneubc=_exists(root, "input", "nansfneubc"),
              sfbc=_exists(root, "input", "sfbc"))
    fl.update(files or {})
    for k in ("parm", "soil", "ic", "atmbc", "dem_parameters", "dem"):
        if not fl.get(k):
            raise FileNotFoundError(f"project {root}: {k} missing")
    parm = read_parm(fl["parm"])
    if parm["isimgr"] == 3:
        raise ValueError("ISIMGR = 3 (surface only) has no subsurface step")
    if parm["iopt"] not in (1, 2):
        raise ValueError(f"IOPT = {parm['iopt']}")
    if parm["nlrelx"] != 0:
        raise ValueError("NLRELX != 0 (nonlinear relaxation) is not supported")
    if parm["kslope"] != 0:
        raise ValueError("KSLOPE != 0 (chord slopes) is not supported")
    dp = read_dem_parameters(fl["dem_parameters"])
    dem, hdr = read_raster(fl["dem"])
    zone = read_raster(fl["zone"])[0].astype(np.int32) if fl.get("zone") else np.ones(dem.shape, np.int32)
    if zone.shape != dem.shape:
        raise ValueError(f"project {root}: the zone raster is {zone.shape[0]} x {zone.shape[1]}, the DEM "
                         f"{dem.shape[0]} x {dem.shape[1]}")
    m = M.build_mesh(dem, dp["dx"], dp["dy"], dp["nstr"], dp["zratio"], dp["base"], zone=zone, ivert=dp["ivert"],
                     west=hdr["west"], south=hdr["south"], factor=dp["factor"])
    # vegetation per surface node: TRIANGOLI's cell-to-node rule on the map
    if fl.get("root_map"):
        vmap = read_raster(fl["root_map"])[0]
        vnod = M.surface_from_dem(vmap, dp["dx"], dp["dy"], zone, hdr["west"], hdr["south"], dp["factor"])[4]
        veg_type = vnod.astype(np.int64).astype(np.int32)
    else:
        veg_type = np.ones(m.nnod, dtype=np.int32)
    nveg = int(max(1, veg_type.max()))
    soil = read_soil(fl["soil"], dp["nstr"], dp["nzone"], nveg)
    ic = read_ic(fl["ic"], m.n, m.nnod)
    surf = parm["isimgr"] >= 2 and not subsurface_only
    ponded = surf and ic["ipond"] != 0   # inital.f:201-205: PONDING from the start, so step 1 routes
    if ponded and not os.path.exists(os.path.join(root, "prepro", "qoi_a")):
        raise ValueError(f"project {root}: initial ponding (IPOND = {ic['ipond']}) routes from the first step, "
                         f"and the project has no drainage network (prepro/qoi_a)")
    psi0 = initial_head(ic, m, surf)
    hspatm, ieto, at, av = read_atmbc(fl["atmbc"], m.nnod)
    bc_neu = B.read_bc(fl["nansfneubc"], "neumann", m.nnod, m.nstr) if fl.get("nansfneubc") else None
    bc_dir = B.read_bc(fl["nansfdirbc"], "dirichlet", m.nnod, m.nstr) if fl.get("nansfdirbc") else None
    free = bool(bc_neu.free_drainage) if bc_neu is not None else False
    if bc_neu is not None and not free and any(r.nodin2 < 0 for r in bc_neu.records):
        raise ValueError("nansfneubc mixes free-drainage and listed records")
    prm = CF.Params(tetaf=parm["tetaf"], lump=parm["lump"], iopt=parm["iopt"], l2norm=parm["l2norm"],
                    toluns=parm["toluns"], tolswi=parm["tolswi"], ernlmx=parm["ernlmx"], ituns=parm["ituns"],
                    ituns1=parm["ituns1"], ituns2=parm["ituns2"], itmxcg=parm["itmxcg"], tolcg=parm["tolcg"],
                    deltat=parm["deltat"], dtmin=parm["dtmin"], dtmax=parm["dtmax"], tmax=parm["tmax"],
                    dtmaga=parm["dtmaga"], dtmagm=parm["dtmagm"], dtreds=parm["dtreds"], dtredm=parm["dtredm"],
                    pmin=soil["pmin"], pondh_min=1.0e10 if not surf else parm["pondh_min"], scf=soil["scf"],
                    ieto=ieto, free_drainage=int(free))
    # listed series only when they list something (the usual files hold empty records)
    has = lambda s: s is not None and any(len(r.nodes) - (m.nnod if r.nodin2 < 0 else 0) > 0 for r in s.records)
    seep = None
    if fl.get("sfbc"):
        recs = B.read_sfbc(fl["sfbc"], m.n, m.z)
        if any(r.faces for r in recs):
            seep = dict(faces=B.seepage_faces(recs), isfcvg=parm["isfcvg"])
    veg = dict(veg_type=veg_type, **soil["veg"])
    # the drainage network of a coupled run, where the project holds one
    routing = None
    if surf and os.path.exists(os.path.join(root, "prepro", "qoi_a")):
        routing = read_routing(root, dem, zone, dp["dx"], dp["dy"], dp["dostep"])
        if ponded:   # PONDNOD as DATIN read it (datin.f:394-403)
            routing["pond0"] = np.asarray(ic["pondnod"], dtype=np.float64).copy()
    meta = dict(project=os.path.abspath(root), parm=parm, isimgr=parm["isimgr"], subsurface_only=subsurface_only,
                ic=dict(indp=ic["indp"], ipond=ic["ipond"], wtposition=ic["wtposition"]), nveg=nveg)
    return CF.Problem(os.path.basename(os.path.abspath(root)), m, soil["table"], psi0, at, av, prm, 0, meta, hspatm,
                      bc_dir if has(bc_dir) else None, bc_neu if has(bc_neu) else None,
                      veg if parm["iopt"] == 1 else None, seep, surf, routing)
