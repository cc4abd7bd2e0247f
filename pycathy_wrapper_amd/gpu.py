"""ctypes binding of libcathy_gpu.so (the gfx950 FLOW3D / SYMSLV drop-in).

The shared library is built in-tree by ``__graft_entry__.build()`` (make -C
pycathy_wrapper_amd/csrc).  Nothing here falls back to a CPU path: a missing
library or device raises.

``FlowGPU`` wraps the FLOW3D handle of include/cathy_gpu.h with the host
arrays a CATHY run holds (1-based TETRA, column-major soil tables, ...),
built from a ``configs.Problem``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import soil as SO

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcathy_gpu.so")

ORDER_NATURAL = 0
ORDER_COLUMN3 = 1
ORDER_COLUMN3C = 2  # the column ordering with each soil column's rows contiguous
ORDER_COLUMN3T = 3  # the column ordering with rows in tiles of 8 soil columns

STAT_NAMES = ["NSTEP", "DELTAT", "TIME", "ITER", "NITERT", "BACK", "STORE1", "STORE2", "DSTORE", "CDSTOR",
              "VIN", "CVIN", "VOUT", "CVOUT", "VINOUT", "VTOT", "ERRAS", "ERREL", "CERRAS", "CAERAS", "PINF",
              "ITRTOT", "ITLIN", "KLSFAI", "PL2", "FINF", "FL2", "IKMAX", "PICARD", "PCG", "PONDING"]
# stats(32) slots of STAT_NAMES; slot 28 (the step's wall-clock ms) is kept apart
_STAT_SLOTS = list(range(28)) + [29, 30, 31]


def _stat_dict(st) -> dict:
    return {nm: float(st[i]) for nm, i in zip(STAT_NAMES, _STAT_SLOTS)}

_lib = None


class CathyGpuError(RuntimeError):
    pass


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CathyGpuError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.cathy_gpu_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc: int, what: str) -> int:
    if rc < 0:
        msg = lib().cathy_gpu_last_error().decode(errors="replace")
        raise CathyGpuError(f"{what} failed ({rc}): {msg}")
    return rc


def _i(v):
    return C.byref(C.c_int(int(v)))


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def device_info() -> tuple[str, int]:
    buf = C.create_string_buffer(128)
    ln = C.c_int(128)
    cu = C.c_int(0)
    _check(lib().cathy_gpu_device_info(buf, C.byref(ln), C.byref(cu)), "cathy_gpu_device_info")
    return buf.value.decode(), cu.value


class FlowGPU:
    """Device-resident subsurface flow run (one mesh)."""

    def __init__(self, problem, ordering: int = ORDER_COLUMN3):
        m = problem.mesh
        s = problem.soil
        self.problem = problem
        self.n, self.nnod = m.n, m.nnod
        nodep = SO.node_params(m, s)
        kz = SO.kz_nodes(m, s)
        tet = np.ascontiguousarray(m.tetra, dtype=np.int32).copy()
        tet[:, :4] += 1
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        # (nstr, nzone)[layer, zone] -> Fortran PERMX(MAXSTR=nstr, NZONE)
        tab = {k: f64(np.asarray(s[k]).T) for k in ("permx", "permy", "permz")}
        keep = dict(tet=tet, bi=f64(m.bi), ci=f64(m.ci), di=f64(m.di), volu=f64(m.volu),
                    ivol=np.ascontiguousarray(m.ivol, dtype=np.int32), volnod=f64(m.volnod), arenod=f64(m.arenod),
                    snodi=f64(nodep["snodi"]), pnodi=f64(nodep["pnodi"]), vgn=f64(nodep["vgn"]),
                    vgrmc=f64(nodep["vgrmc"]), vgpsat=f64(nodep["vgpsat"]), kz=f64(kz), **tab)
        self._keep = keep
        self.node_params = nodep
        h = C.c_int64(0)
        rc = lib().cathy_gpu_flow3d_create(
            C.byref(h), _i(m.n), _i(m.nt), _i(m.ntri), _i(m.nnod), _i(m.nstr), _i(s["nzone"]), _i(m.nstr),
            _p(keep["tet"]), _p(keep["bi"]), _p(keep["ci"]), _p(keep["di"]), _p(keep["volu"]), _p(keep["ivol"]),
            _p(keep["volnod"]), _p(keep["arenod"]), _p(keep["permx"]), _p(keep["permy"]), _p(keep["permz"]),
            _p(keep["snodi"]), _p(keep["pnodi"]), _p(keep["vgn"]), _p(keep["vgrmc"]), _p(keep["vgpsat"]),
            _p(keep["kz"]), _i(problem.params.lump), _i(ordering))
        _check(rc, "cathy_gpu_flow3d_create")
        self.h = h
        self.set_params(problem.params)
        self.set_forcing(problem.atm_times, problem.atm_vals)
        if getattr(problem, "bc_dir", None) is not None:
            self.set_bc(problem.bc_dir)
        if getattr(problem, "bc_neu", None) is not None:
            self.set_bc(problem.bc_neu)
        if getattr(problem, "veg", None) is not None:
            self.set_vegetation(**problem.veg)
        if getattr(problem, "seepage", None) is not None:
            self.set_seepage(problem.seepage["faces"], problem.seepage.get("isfcvg", 0))
        if getattr(problem, "surf", False):
            self.set_surface(True)
        self.route = None
        if getattr(problem, "routing", None) is not None:
            self.set_routing(problem.routing)
            if problem.routing.get("pond0") is not None:
                self.set_initial_ponding(problem.routing["pond0"])
        self.init(problem.psi0)

    def set_params(self, p, hspatm: int | None = None):
        """IPRM/DPRM of cathy_gpu_flow3d_set_params; HSPATM defaults to the
        problem's (and fixes the record width set_forcing expects)."""
        hsp = int(self.problem.hspatm if hspatm is None else hspatm)
        ip = np.array([p.iopt, p.l2norm, p.ituns, p.ituns1, p.ituns2, p.itmxcg, p.ieto, p.free_drainage, hsp],
                      dtype=np.int32)
        dp = np.array([p.tetaf, p.toluns, p.tolswi, p.ernlmx, p.tolcg, p.deltat, p.dtmin, p.dtmax, p.tmax,
                       p.dtmaga, p.dtmagm, p.dtreds, p.dtredm, p.pmin, p.pondh_min, p.scf], dtype=np.float64)
        _check(lib().cathy_gpu_flow3d_set_params(C.byref(self.h), _p(ip), _p(dp)), "set_params")

    def set_forcing(self, times, vals):
        t = np.ascontiguousarray(times, dtype=np.float64)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        width = 1 if self.problem.hspatm else self.nnod
        if v.size != len(t) * width:
            raise ValueError(f"set_forcing: {v.size} values for {len(t)} records of width {width}")
        _check(lib().cathy_gpu_flow3d_set_forcing(C.byref(self.h), _i(len(t)), _p(t), _p(v)), "set_forcing")

    def set_bc(self, series) -> None:
        """A nansfdirbc / nansfneubc series (bcfiles.BCSeries) for the run;
        call init again afterwards (the constructor does, for the problem's
        own series)."""
        kind = {"dirichlet": 0, "neumann": 1}[series.kind]
        if kind == 1 and series.free_drainage != bool(self.problem.params.free_drainage):
            raise ValueError("set_bc: the Neumann series and the free-drainage parameter disagree")
        t, cnt, nodes, vals = series.explicit(self.nnod)
        self._keep[f"bc{kind}"] = (t, cnt, nodes, vals)
        _check(lib().cathy_gpu_flow3d_set_bc(C.byref(self.h), _i(kind), _i(len(t)), _p(t), _p(cnt), _p(nodes),
                                             _p(vals)), "set_bc")

    def set_seepage(self, faces, isfcvg: int = 0) -> None:
        """Time-invariant seepage faces: a list of 0-based node arrays, each
        face top to bottom (bcfiles.read_sfbc gives them); ISFCVG = 1 makes
        the exit points part of the convergence test.  Call init again
        afterwards."""
        counts = np.array([len(f) for f in faces], dtype=np.int32)
        nodes = (np.concatenate([np.asarray(f, dtype=np.int64) for f in faces]) + 1).astype(np.int32) \
            if faces else np.zeros(0, dtype=np.int32)
        self._keep["sf"] = (counts, nodes)
        _check(lib().cathy_gpu_flow3d_set_seepage(C.byref(self.h), _i(len(faces)), _p(counts), _p(nodes),
                                                  _i(isfcvg)), "set_seepage")

    def bench_spmv(self, nrep: int = 50, flush_mb: float = 1024.0):
        """(cold ms, warm ms) of the PCG SpMV on the current matrix: each
        cold launch follows a flush_mb copy that clears the caches."""
        out = np.zeros(2)
        _check(lib().cathy_gpu_flow3d_bench_spmv(C.byref(self.h), _i(nrep), C.byref(C.c_double(flush_mb)), _p(out)),
               "bench_spmv")
        return float(out[0]), float(out[1])

    def seepage_state(self):
        """(SFEX, SFQ) of the seepage nodes, face by face, top to bottom."""
        m = _check(lib().cathy_gpu_flow3d_get_seepage(C.byref(self.h), None, None), "get_seepage")
        ex, q = np.zeros(m, dtype=np.int32), np.zeros(m)
        _check(lib().cathy_gpu_flow3d_get_seepage(C.byref(self.h), _p(ex), _p(q)), "get_seepage")
        return ex, q

    def set_vegetation(self, veg_type, pcana, pcref, pcwlt, zroot, pz, omgc) -> None:
        """Feddes root uptake (ETRAN): vegetation type per surface node
        (1-based) and the per-type table; call init again afterwards."""
        f64 = lambda a: np.ascontiguousarray(np.atleast_1d(a), dtype=np.float64)
        vt = np.ascontiguousarray(veg_type, dtype=np.int32)
        tab = [f64(a) for a in (pcana, pcref, pcwlt, zroot, pz, omgc)]
        z = f64(self.problem.mesh.z)
        self._keep["veg"] = (vt, tab, z)
        _check(lib().cathy_gpu_flow3d_set_vegetation(C.byref(self.h), _i(len(tab[0])), _p(vt),
                                                     *[_p(a) for a in tab], _p(z)), "set_vegetation")

    def set_surface(self, surf: bool) -> None:
        """The coupled surface of ISIMGR = 2 (SWITCH in CONVER, ADRSTN at
        step starts; advance returns 3 after a step that ends ponded) or the
        subsurface-only switching; call init again afterwards."""
        _check(lib().cathy_gpu_flow3d_set_surface(C.byref(self.h), _i(1 if surf else 0)), "set_surface")

    def set_routing(self, routing: dict) -> None:
        """The routing block of a coupled run (project.read_routing): the
        drainage network goes to a routing handle (routing.RouteGPU) linked
        to this flow handle; call init again afterwards."""
        from .routing import RouteGPU
        m = self.problem.mesh
        if self.route is not None:
            self.route.close()
        self.route = RouteGPU(routing["net"])
        nr, nc = routing["nrow"], routing["ncol"]
        colmaj = lambda g: np.ascontiguousarray(np.asarray(g).T.reshape(-1))   # (I, J) at (I-1) + (J-1)*NROW
        cell = np.zeros((len(routing["cell"]), 5), dtype=np.int32)
        cell[:, :4] = routing["cell"] + 1
        tri = np.ascontiguousarray(m.triang, dtype=np.int32).copy()
        tri[:, :3] += 1
        keep = dict(cell=np.ascontiguousarray(cell.reshape(-1)), indcel=colmaj(routing["index"]).astype(np.int32),
                    indcelwl=colmaj(routing["index_wl"]).astype(np.int32),
                    dem=colmaj(routing["dem"]).astype(np.float64), tri=np.ascontiguousarray(tri.reshape(-1)))
        self._keep["routing"] = keep
        _check(lib().cathy_gpu_flow3d_set_routing(
            C.byref(self.h), C.byref(self.route.h), _i(nr), _i(nc), C.byref(C.c_double(routing["dx"])),
            C.byref(C.c_double(routing["dy"])), _i(len(cell)), _p(keep["cell"]), _p(keep["indcel"]),
            _p(keep["indcelwl"]), _p(keep["dem"]), _i(m.ntri), _p(keep["tri"])), "set_routing")

    def set_initial_ponding(self, pondnod) -> None:
        """Initial ponding heads PONDNOD (IPOND 1 or 2): the run starts
        ponded and its first step routes (needs the routing linked); None
        clears them.  Takes effect at the next init."""
        if pondnod is None:
            _check(lib().cathy_gpu_flow3d_set_initial_ponding(C.byref(self.h), None), "set_initial_ponding")
            return
        p = np.ascontiguousarray(pondnod, dtype=np.float64)
        if p.shape != (self.nnod,):
            raise ValueError("one ponding head per surface node")
        self._keep["pond0"] = p
        _check(lib().cathy_gpu_flow3d_set_initial_ponding(C.byref(self.h), _p(p)), "set_initial_ponding")

    def routing_state(self):
        """(PONDNOD, OVFLNOD, H_WATER_KKP1_SN, routing blocks run)."""
        m = self.problem.mesh
        r = self.problem.routing
        pond, ovfl = np.zeros(m.nnod), np.zeros(m.nnod)
        hw = np.zeros(r["nrow"] * r["ncol"])
        info = np.zeros(2, dtype=np.int32)
        _check(lib().cathy_gpu_flow3d_routing_state(C.byref(self.h), _p(pond), _p(ovfl), _p(hw), _p(info)),
               "routing_state")
        return pond, ovfl, hw, int(info[0])

    def init(self, psi0):
        p = np.ascontiguousarray(psi0, dtype=np.float64)
        _check(lib().cathy_gpu_flow3d_init(C.byref(self.h), _p(p)), "flow3d_init")

    def advance(self):
        st = np.zeros(32)
        rc = _check(lib().cathy_gpu_flow3d_advance(C.byref(self.h), _p(st)), "flow3d_advance")
        self.last_wall_ms = float(st[28])
        return rc, _stat_dict(st)

    def psi(self) -> np.ndarray:
        out = np.zeros(self.n)
        _check(lib().cathy_gpu_flow3d_get_state(C.byref(self.h), _p(out), None, None, None), "get_state")
        return out

    def ifatm(self) -> np.ndarray:
        """IFATM of the surface nodes (-1 another boundary condition, 0
        Neumann, 1/2 Dirichlet atmospheric)."""
        out = np.zeros(self.nnod, dtype=np.int32)
        _check(lib().cathy_gpu_flow3d_get_state(C.byref(self.h), None, None, None, _p(out)), "get_state")
        return out

    def sw(self) -> np.ndarray:
        out = np.zeros(self.n)
        _check(lib().cathy_gpu_flow3d_get_state(C.byref(self.h), None, _p(out), None, None), "get_state")
        return out

    def guess(self, net=None, stat=None, detach: bool = False) -> None:
        """Attach a learned-initial-iterate network (guess.GuessNet) and/or
        set the static column features (guess.static_features)."""
        st = None if stat is None else np.ascontiguousarray(stat, dtype=np.float32)
        if st is not None:
            self._keep["guess_stat"] = st
        h = None
        if net is not None:
            h = C.byref(C.c_int64(net.h.value))
            self._guess_net = net
        elif detach:
            h = C.byref(C.c_int64(0))
            self._guess_net = None
        _check(lib().cathy_gpu_flow3d_guess(C.byref(self.h), h, None if st is None else _p(st)), "flow3d_guess")

    def set_dt_schedule(self, dts) -> None:
        """Replay step sizes: dts[k] is the DELTAT of step k+1 (call after
        construction, before the first advance)."""
        d = np.ascontiguousarray(dts, dtype=np.float64)
        _check(lib().cathy_gpu_flow3d_set_dt_schedule(C.byref(self.h), _i(d.size), _p(d)), "set_dt_schedule")

    def factor_fp64(self, on: bool = True) -> None:
        """FP64 IC(0) factor in the column ordering (FP32 by default; the
        natural ordering is always FP64)."""
        _check(lib().cathy_gpu_flow3d_set_factor_fp64(C.byref(self.h), _i(1 if on else 0)), "set_factor_fp64")

    def spmv_transposed(self, on: bool = True) -> bool:
        """PCG SpMV reading the lower triangle through the transpose (bitwise
        the same results; faster for one member alone, slower when members
        share the GPU).  Returns whether it applies to this mesh."""
        act = C.c_int(0)
        _check(lib().cathy_gpu_flow3d_set_spmv_transposed(C.byref(self.h), _i(1 if on else 0), C.byref(act)),
               "set_spmv_transposed")
        return bool(act.value)

    def pcg_batches(self, first_pct: int = 100, followup: int = 2) -> None:
        """PCG launch batches between convergence polls (results unchanged)."""
        _check(lib().cathy_gpu_flow3d_set_pcg_batches(C.byref(self.h), _i(first_pct), _i(followup)),
               "set_pcg_batches")

    def guess_record(self, on: bool = True) -> None:
        _check(lib().cathy_gpu_flow3d_guess_record(C.byref(self.h), _i(1 if on else 0)), "guess_record")

    def guess_sample(self):
        """Raw features [NNOD][Lp][16] and increments [NNOD][Lp] of the last
        accepted step (recording on)."""
        L = self.problem.mesh.nstr + 1
        Lp = (L + 15) // 16 * 16
        feat = np.zeros((self.nnod, Lp, 16), dtype=np.float32)
        tgt = np.zeros((self.nnod, Lp), dtype=np.float32)
        _check(lib().cathy_gpu_flow3d_guess_sample(C.byref(self.h), _p(feat), _p(tgt)), "guess_sample")
        return feat, tgt

    def set_pnew(self, psi) -> None:
        p = np.ascontiguousarray(psi, dtype=np.float64)
        _check(lib().cathy_gpu_flow3d_set_pnew(C.byref(self.h), _p(p)), "set_pnew")

    def ckrw(self) -> np.ndarray:
        """Nodal kr; after advance() it is kr(PNEW) (chvelo.f)."""
        out = np.zeros(self.n)
        _check(lib().cathy_gpu_flow3d_get_state(C.byref(self.h), None, None, _p(out), None), "get_state")
        return out

    def ifatm(self) -> np.ndarray:
        out = np.zeros(self.nnod, dtype=np.int32)
        _check(lib().cathy_gpu_flow3d_get_state(C.byref(self.h), None, None, None, _p(out)), "get_state")
        return out

    def velocity(self) -> dict:
        """VEL3D + VNOD3D at the current state: element UU/VV/WW (NT) and
        nodal UNOD/VNOD/WNOD (N, CATHY order)."""
        k = self._keep
        nt, n = len(k["volu"]), self.n
        if "volur" not in k:
            k["volur"] = np.ascontiguousarray(1.0 / k["volu"])
        out = {name: np.zeros(nt) for name in ("uu", "vv", "ww")}
        out.update({name: np.zeros(n) for name in ("unod", "vnod", "wnod")})
        _check(lib().cathy_gpu_flow3d_vel3d(C.byref(self.h), _p(k["tet"]), _p(k["permx"]), _p(k["permy"]),
                                            _p(k["permz"]), _p(k["bi"]), _p(k["ci"]), _p(k["di"]), _p(k["volur"]),
                                            _p(k["ivol"]), *(_p(out[nm]) for nm in ("uu", "vv", "ww", "unod", "vnod",
                                                                                   "wnod"))), "vel3d")
        return out

    def counters(self) -> dict:
        out = np.zeros(6)
        _check(lib().cathy_gpu_flow3d_counters(C.byref(self.h), _p(out)), "counters")
        return dict(picard=out[0], pcg=out[1], fallbacks=out[2], solve_ms=out[3], step_ms=out[4])

    def profile(self, enable: bool | None = None) -> tuple[float, int]:
        """SpMV kernel timing: returns (total ms, launches) accumulated since
        the last enable/disable; ``enable`` switches timing on/off and resets."""
        ms = C.c_double(0.0)
        cnt = C.c_int64(0)
        en = None if enable is None else _i(1 if enable else 0)
        _check(lib().cathy_gpu_flow3d_profile(C.byref(self.h), en, C.byref(ms), C.byref(cnt)), "profile")
        return ms.value, cnt.value

    def spmv_bytes(self) -> float:
        b = C.c_double(0.0)
        _check(lib().cathy_gpu_flow3d_spmv_bytes(C.byref(self.h), C.byref(b)), "spmv_bytes")
        return b.value

    def layout(self) -> dict:
        out = np.zeros(6, dtype=np.int64)
        _check(lib().cathy_gpu_flow3d_layout(C.byref(self.h), _p(out)), "layout")
        return dict(n=int(out[0]), nnz=int(out[1]), nnz_sell=int(out[2]), maxrow=int(out[3]),
                    coded=int(out[4]), code_mismatch=int(out[5]))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().cathy_gpu_flow3d_destroy(C.byref(self.h))
            self.h = C.c_int64(0)
        if getattr(self, "route", None) is not None:
            self.route.close()
            self.route = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EnsembleGPU:
    """Flow runs advanced in lockstep (cathy_gpu_ensemble_*): one launch per
    kernel for all members, a convergence mask per member.  ``members``:
    FlowGPU objects (column ordering, Picard) that the ensemble drives while
    ``advance`` runs."""

    def __init__(self, members):
        self.members = list(members)
        h = np.array([m.h.value for m in self.members], dtype=np.int64)
        e = C.c_int64(0)
        _check(lib().cathy_gpu_ensemble_create(C.byref(e), _i(len(h)), _p(h)), "cathy_gpu_ensemble_create")
        self.h = e

    def advance(self, nsteps: int = 1):
        """Every member takes ``nsteps`` accepted steps (or ends its run).
        Returns (rc[m], taken[m], stats) with stats[m][s] the statistics dict
        of member m's step s (s < taken[m]); ``self.wall_ms[m][s]`` is that
        step's wall-clock ms, from its start to its end."""
        nm = len(self.members)
        rc = np.zeros(nm, dtype=np.int32)
        taken = np.zeros(nm, dtype=np.int32)
        st = np.zeros((nm, nsteps, 32))
        _check(lib().cathy_gpu_ensemble_advance(C.byref(self.h), _i(nsteps), _p(rc), _p(taken), _p(st)),
               "cathy_gpu_ensemble_advance")
        stats = [[_stat_dict(st[m, s]) for s in range(int(taken[m]))] for m in range(nm)]
        self.wall_ms = [st[m, :int(taken[m]), 28].copy() for m in range(nm)]
        return rc, taken, stats

    def graphs(self, on: bool = True) -> None:
        """Capture each round's fixed launch sequences as HIP graphs and
        replay them (the same results)."""
        _check(lib().cathy_gpu_ensemble_set_graphs(C.byref(self.h), _i(1 if on else 0)), "ensemble_set_graphs")

    def info(self) -> dict:
        out = np.zeros(5, dtype=np.int64)
        _check(lib().cathy_gpu_ensemble_info(C.byref(self.h), _p(out)), "ensemble_info")
        return dict(members=int(out[0]), rounds=int(out[1]), polls=int(out[2]), pcg_batches=int(out[3]),
                    graph_captures=int(out[4]))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().cathy_gpu_ensemble_destroy(C.byref(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def symslv(topol: np.ndarray, ja: np.ndarray, coef1: np.ndarray, tnoti: np.ndarray, noddir=None,
           itmxcg: int = 500, tolcg: float = 1e-10, rmin: float = 1.7e-100):
    """SYMSLV on the device (solscal-extended.f:4669): ``topol`` (N+1),
    ``ja`` (NTERM) 1-based upper CSR with the diagonal first, ``coef1``
    (NTERM), ``tnoti`` (N), ``noddir`` 1-based Dirichlet rows.  Returns
    (pnew, niter, err, res, rc) with rc = 1 when NITER reached ITMXCG."""
    topol = np.ascontiguousarray(topol, dtype=np.int32)
    ja = np.ascontiguousarray(ja, dtype=np.int32)
    coef1 = np.ascontiguousarray(coef1, dtype=np.float64)
    tnoti = np.ascontiguousarray(tnoti, dtype=np.float64)
    nd = np.ascontiguousarray(noddir if noddir is not None else np.zeros(0), dtype=np.int32)
    n = len(topol) - 1
    pnew = np.zeros(n)
    niter = C.c_int(0)
    err = C.c_double(0.0)
    res = C.c_double(0.0)
    rc = _check(lib().cathy_gpu_symslv(_i(n), _i(len(ja)), _i(len(nd)), C.byref(niter), _i(itmxcg),
                                       C.byref(C.c_double(tolcg)), C.byref(C.c_double(rmin)),
                                       _p(nd) if len(nd) else None, _p(ja), _p(topol), _p(pnew), _p(tnoti),
                                       _p(coef1), C.byref(err), C.byref(res)), "cathy_gpu_symslv")
    return pnew, niter.value, err.value, res.value, rc


def symslv_release() -> None:
    _check(lib().cathy_gpu_symslv_release(), "cathy_gpu_symslv_release")


def calib_copy(nbytes: int, reps: int = 20) -> float:
    """Average ms of a copy kernel reading and writing ``nbytes``."""
    ms = C.c_double(0.0)
    _check(lib().cathy_gpu_calib_copy(C.byref(C.c_int64(int(nbytes))), _i(reps), C.byref(ms)), "calib_copy")
    return ms.value
