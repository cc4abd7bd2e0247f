"""Surface routing substep (ISIMGR = 2): ROUTE + ALTEZZE on the device.

The reference routes the surface water of every DEM cell down its drainage
network once per routing substep (route.f:4-234, mc.f:8-63): cells are
visited from the highest to the lowest (QOI_SN), each splitting its inflow
and its share of surface water between a cardinal (W_1) and a diagonal
(W_2) downstream cell with Muskingum-Cunge coefficients, and each outflow is
added to the downstream cell's inflow Q_IN_KKP1_SN.  ALTEZZE
(altezze.f:4-60) then updates every cell's water volume and depth.

``Network`` holds the reference's RIVERNETWORK arrays for one DEM:
DTM_*(I, J) at row J-1, column I-1 of an (nrow, ncol) array (so the flat
index is (I-1) + (J-1)*ncol, the reference's column-major layout with
COLMAX = ncol) and QOI_SN as 1-based I_BASIN = (I-1)*nrow + J.
``RouteGPU`` runs the substep for a batch of ensemble members that share the
network (cathy_gpu_route_* in include/cathy_gpu.h), reservoirs included
(SERBATOIO / CALCOLI).  ALTEZZE's cell types are its own argument: the
reference's time loop passes CELTYPE, zero everywhere (cathy_main.f:3062),
so every cell takes the volume balance unless ``celtype`` says otherwise.
The library rejects networks whose downstream cells do not come later in
QOI_SN, or whose targets leave the grid.

``synthetic_network`` builds a pit-free network on a tilted, perturbed
plane.  It stands in for the DEM preprocessing (not part of this package):
its cells drain toward cell (1, 1), which comes last.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .gpu import CathyGpuError, _check, _i, _p, lib  # noqa: F401  (CathyGpuError re-exported)

DTM_REAL = ("w1", "w2", "slope1", "slope2", "epl1", "epl2", "kss1", "kss2", "ws1", "ws2", "b1", "y1", "nrc")

# D8 direction codes of DTM_P_OUTFLOW_*: code = 5 + 3*dI + dJ (route.f:95-98)
CARDINAL = ((-1, 0), (1, 0), (0, -1), (0, 1))
DIAGONAL = ((-1, -1), (-1, 1), (1, -1), (1, 1))


@dataclass
class Network:
    nrow: int
    ncol: int
    qoi: np.ndarray        # (ncell,) int32, 1-based I_BASIN in routing order
    pout1: np.ndarray      # (nrow, ncol) int32, DTM_P_OUTFLOW_1
    pout2: np.ndarray      # (nrow, ncol) int32, DTM_P_OUTFLOW_2
    w1: np.ndarray         # (nrow, ncol) float64, DTM_W_1 (likewise below)
    w2: np.ndarray
    slope1: np.ndarray
    slope2: np.ndarray
    epl1: np.ndarray
    epl2: np.ndarray
    kss1: np.ndarray
    kss2: np.ndarray
    ws1: np.ndarray
    ws2: np.ndarray
    b1: np.ndarray
    y1: np.ndarray
    nrc: np.ndarray
    dx: float = 1.0
    dy: float = 1.0
    # reservoirs, or None: tipo_r, reservr, elevation (nrow*ncol, by I_BASIN - 1:
    # TIPO_R, RESERVR, ELEVATION), n_ha (numres), hres / ares ((30, maxres):
    # HRES(MAXRES,30) column-major), h_soglia, h_fondo (numres), maxres
    reservoirs: dict | None = None

    @property
    def ncell(self) -> int:
        return int(len(self.qoi))

    @property
    def ngrid(self) -> int:
        return self.nrow * self.ncol


def synthetic_network(nrow: int, ncol: int, dx: float = 10.0, seed: int = 0, reservoir: bool = False,
                      masked: float = 0.0, sinks: bool = False) -> Network:
    """A drainage network on z = s*((I-1) + (J-1))*dx + noise (|noise| <
    s*dx/4, so every cell but (1, 1) has a lower cardinal neighbour and no
    two cells tie).  Each cell sends its water down its steepest cardinal
    and steepest diagonal descents in proportion to their slopes; cells on
    the low edges have no lower diagonal and send all of it cardinally.
    Roughness, width and channel-shape parameters vary by cell.

    With ``reservoir``: one reservoir (RESERVR = 1) whose outlet is the cell
    (I, J) = (ncol/2, nrow/2) (TIPO_R = 1) and whose lake covers the two
    cells above it (TIPO_R = 1000 + the outlet's I_BASIN, no routing of
    their own); a 5-entry height-area table around the outlet's elevation,
    the spill height 0.2 m above it and the bottom 1 m below.

    With ``masked`` > 0: that fraction of the grid, the highest cells, lies
    outside the catchment (ZONE = 0 in the reference's DEM): those cells are
    left out of QOI_SN and route nothing, so NCELL < NROW*NCOL.  No cell
    drains into them, since every drainage path descends -- unless
    ``sinks``: then every routed cell with an outside diagonal neighbour
    sends 30% of its water there (route.f adds such outflow to the cell's
    inflow, which only ALTEZZE uses)."""
    rng = np.random.default_rng(seed)
    s = 0.02
    I = np.arange(1, ncol + 1)[None, :]
    J = np.arange(1, nrow + 1)[:, None]
    z = s * ((I - 1) + (J - 1)) * dx + rng.uniform(-0.24, 0.24, (nrow, ncol)) * s * dx
    ib = (I - 1) * nrow + J                                   # 1-based I_BASIN
    pout1 = np.full((nrow, ncol), 5, dtype=np.int32)
    pout2 = np.full((nrow, ncol), 5, dtype=np.int32)
    sl1 = np.zeros((nrow, ncol))
    sl2 = np.zeros((nrow, ncol))
    for jj in range(nrow):
        for ii in range(ncol):
            for dirs, dist, po, sl in ((CARDINAL, dx, pout1, sl1), (DIAGONAL, dx * np.sqrt(2.0), pout2, sl2)):
                best = 0.0
                for di, dj in dirs:
                    i2, j2 = ii + di, jj + dj
                    if 0 <= i2 < ncol and 0 <= j2 < nrow:
                        drop = (z[jj, ii] - z[j2, i2]) / dist
                        if drop > best:
                            best = drop
                            po[jj, ii] = 5 + 3 * di + dj
                sl[jj, ii] = best
    w1 = np.where(sl1 + sl2 > 0, sl1 / np.maximum(sl1 + sl2, 1e-300), 0.0)
    w2 = np.where(sl1 + sl2 > 0, sl2 / np.maximum(sl1 + sl2, 1e-300), 0.0)
    w1[0, 0], w2[0, 0], pout1[0, 0], sl1[0, 0] = 1.0, 0.0, 2, s   # the outlet leaves through I = 0
    order = np.argsort(-z.ravel(order="C"), kind="stable")
    nmask = int(round(masked * nrow * ncol))
    if nmask:
        out = order[:nmask]                                  # the highest cells
        w1.ravel(order="C")[out] = 0.0
        w2.ravel(order="C")[out] = 0.0
        order = order[nmask:]
    qoi = ib.ravel(order="C")[order].astype(np.int32)
    if sinks and nmask:
        routed = np.zeros(nrow * ncol, dtype=bool)
        routed[qoi - 1] = True
        for jj in range(nrow):
            for ii in range(ncol):
                if not routed[ii * nrow + jj] or ii * nrow + jj + 1 == qoi[-1]:
                    continue
                for di, dj in DIAGONAL:
                    i2, j2 = ii + di, jj + dj
                    if 0 <= i2 < ncol and 0 <= j2 < nrow and not routed[i2 * nrow + j2]:
                        pout2[jj, ii] = 5 + 3 * di + dj
                        w2[jj, ii] = 0.3
                        w1[jj, ii] = 0.7
                        sl2[jj, ii] = max(sl2[jj, ii], 1e-3)
                        break
    f = lambda lo, hi: rng.uniform(lo, hi, (nrow, ncol))
    res = None
    if reservoir:
        io, jo = ncol // 2, nrow // 2                        # 1-based outlet (I, J)
        ib_o = (io - 1) * nrow + jo
        tipo = np.zeros(nrow * ncol, dtype=np.int32)
        reservr = np.zeros(nrow * ncol, dtype=np.int32)
        tipo[ib_o - 1], reservr[ib_o - 1] = 1, 1
        for i, j in ((io + 1, jo), (io, jo + 1)):
            tipo[(i - 1) * nrow + j - 1] = 1000 + ib_o
            reservr[(i - 1) * nrow + j - 1] = 1
            w1[j - 1, i - 1] = w2[j - 1, i - 1] = 0.0
        zo = z[jo - 1, io - 1]
        elev = np.empty(nrow * ncol)
        elev[(ib.ravel(order="C") - 1)] = z.ravel(order="C")
        res = dict(tipo_r=tipo, reservr=reservr, n_ha=np.array([5], dtype=np.int32), maxres=1,
                   hres=(zo + np.array([-2.0, -0.5, 0.5, 1.5, 3.0])).reshape(5, 1),
                   ares=np.array([5.0, 10.0, 20.0, 30.0, 40.0]).reshape(5, 1),
                   h_soglia=np.array([zo + 0.2]), h_fondo=np.array([zo - 1.0]), elevation=elev)
    return Network(nrow=nrow, ncol=ncol, qoi=qoi, pout1=pout1, pout2=pout2, w1=w1, w2=w2,
                   slope1=np.maximum(sl1, 1e-3), slope2=np.maximum(sl2, 1e-3),
                   epl1=np.full((nrow, ncol), dx), epl2=np.full((nrow, ncol), dx * np.sqrt(2.0)),
                   kss1=f(5.0, 40.0), kss2=f(5.0, 40.0), ws1=f(0.5, 5.0), ws2=f(0.5, 5.0),
                   b1=f(0.0, 0.5), y1=f(0.0, 0.3), nrc=np.round(f(1.0, 3.0)), dx=dx, dy=dx, reservoirs=res)


STATE_IN = ("surface_water", "q_in_kk", "q_out_kk1", "q_out_kk2", "volume_kk")
STATE_INOUT = ("q_in_kkp1", "q_out_kkp1_1", "q_out_kkp1_2")
STATE_OUT = ("volume_kkp1", "h_water_kkp1")


class RouteGPU:
    """One drainage network on the device; ``step`` runs ROUTE + ALTEZZE
    for ``nmember`` members at once (arrays [member][I_BASIN-1]).
    ``celtype``: ALTEZZE's type array (nrow*ncol, by I_BASIN - 1), None for
    0 everywhere (the reference's call)."""

    def __init__(self, net: Network, celtype=None):
        self.net = net
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        self._keep = dict(qoi=i32(net.qoi), pout1=i32(net.pout1), pout2=i32(net.pout2),
                          **{k: f64(getattr(net, k)) for k in DTM_REAL})
        k = self._keep
        h = C.c_int64(0)
        _check(lib().cathy_gpu_route_create(
            C.byref(h), _i(net.nrow), _i(net.ncol), _i(net.ncol), _i(net.ncell), _p(k["qoi"]), _p(k["pout1"]),
            _p(k["pout2"]), *[_p(k[n]) for n in DTM_REAL]), "cathy_gpu_route_create")
        self.h = h
        self.numres = 0
        if net.reservoirs is not None:
            res = net.reservoirs
            self._keep.update(tipo_r=i32(res["tipo_r"]), reservr=i32(res["reservr"]), n_ha=i32(res["n_ha"]),
                              hres=f64(res["hres"]), ares=f64(res["ares"]), h_soglia=f64(res["h_soglia"]),
                              h_fondo=f64(res["h_fondo"]), elevation=f64(res["elevation"]))
            self.numres = len(res["h_soglia"])
            _check(lib().cathy_gpu_route_set_reservoirs(
                C.byref(self.h), _p(k["tipo_r"]), _p(k["reservr"]), _i(self.numres), _i(res["maxres"]),
                _p(k["n_ha"]), _p(k["hres"]), _p(k["ares"]), _p(k["h_soglia"]), _p(k["h_fondo"]),
                _p(k["elevation"])), "cathy_gpu_route_set_reservoirs")
        if celtype is not None:
            self._keep["celtype"] = i32(celtype)
            _check(lib().cathy_gpu_route_set_celtype(C.byref(self.h), _p(self._keep["celtype"])),
                   "cathy_gpu_route_set_celtype")
        info = np.zeros(4, dtype=np.int32)
        _check(lib().cathy_gpu_route_info(C.byref(self.h), _p(info)), "cathy_gpu_route_info")
        self.nlevel, self.maxpred = int(info[0]), int(info[1])

    def step(self, deltat: float, state: dict) -> tuple[np.ndarray, np.ndarray]:
        """Advance ``state`` (dict of float64 arrays, (nmember, nrow*ncol) or
        (nrow*ncol,), indexed by I_BASIN - 1) in place: q_in_kkp1 / q_out_kkp1_* /
        volume_kkp1 are read and written, h_water_kkp1 written; with
        reservoirs also h_pool_kk (read) and h_pool_kkp1 (read and written),
        (nmember, numres).  Returns CU_MAX and AK_MAX per member (AK_MAX
        enters as state["ak_max"] if present, else 0)."""
        ncell = self.net.ngrid
        arrs = {}
        for k in STATE_IN + STATE_INOUT + STATE_OUT:
            a = state[k]
            if a.dtype != np.float64 or not a.flags.c_contiguous or a.size % ncell:
                raise ValueError(f"{k}: float64, C-contiguous, a multiple of {ncell} values")
            arrs[k] = a
        nm = arrs["q_in_kk"].size // ncell
        if any(a.size != nm * ncell for a in arrs.values()):
            raise ValueError("every state array needs nmember * ncell values")
        cu_ak = np.zeros((nm, 2))
        cu_ak[:, 1] = np.asarray(state.get("ak_max", 0.0), dtype=np.float64)
        hp = [None, None]
        if self.numres:
            for i, k in enumerate(("h_pool_kk", "h_pool_kkp1")):
                a = state[k]
                if a.dtype != np.float64 or not a.flags.c_contiguous or a.size != nm * self.numres:
                    raise ValueError(f"{k}: float64, C-contiguous, nmember * numres values")
                hp[i] = _p(a)
        _check(lib().cathy_gpu_route_step(
            C.byref(self.h), _i(nm), C.byref(C.c_double(deltat)), C.byref(C.c_double(self.net.dx)),
            C.byref(C.c_double(self.net.dy)), *[_p(arrs[k]) for k in STATE_IN + STATE_INOUT + STATE_OUT],
            _p(cu_ak), hp[0], hp[1]), "cathy_gpu_route_step")
        return cu_ak[:, 0].copy(), cu_ak[:, 1].copy()

    def close(self):
        if self.h.value:
            lib().cathy_gpu_route_destroy(C.byref(self.h))
            self.h = C.c_int64(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
