"""Synthetic workloads of the subsurface Picard step.

Each builder returns a ``Problem``: mesh, per-(layer, zone) soil tables,
initial pressure head, atmospheric forcing series and the ``input/parm``
solver parameters.  Everything is generated from a seed; nothing is read
from disk.

* ``hillslope_small``  -- 21x21-node DEM (20x20 cells of 5 m) on a 2% planar
  slope with N(0, 0.05^2) noise, 10 uniform layers to 5 m (4,851 nodes,
  24,000 tets); homogeneous van Genuchten soil (n=1.46, alpha=1/m,
  theta_s=0.45, theta_r=0.05, Ks=1e-5 m/s, Ss=1e-4 1/m); water table 1 m
  deep; 1e-7 m/s rain; free drainage at the base; one simulated day.
* ``hillslope_large``  -- 201x201-node DEM (200x200 cells of 2 m) with
  0-20 m spectral (fractal) relief, 30 geometrically graded layers to 10 m
  (1,252,431 nodes, 7.2M tets); lognormal Ks field (geometric mean 1e-5 m/s,
  ln-sigma 1, exponential covariance length 20 m) per (layer, DEM cell)
  zone, alpha and n perturbed +-20% per zone; 6 h of 5e-7 m/s rain then
  drainage; two simulated days.  ``cells`` shrinks the DEM for tests.

Soil values are rounded to 4 significant digits, the precision of the
soil file pyCATHY writes (cathy_tools.py:3130, ``%1.3e``), so a run from
files and a run from these arrays see the same numbers.
"""
from __future__ import annotations

from dataclasses import dataclass, field, asdict

import numpy as np

from . import mesh as M


@dataclass
class Params:
    """input/parm fields used by the subsurface step (datin.f:80-122), the
    soil-file head (PMIN) and the forcing options."""
    tetaf: float = 1.0
    lump: int = 1
    iopt: int = 1
    l2norm: int = 0
    toluns: float = 1e-6
    tolswi: float = 1e30
    ernlmx: float = 1e30
    ituns: int = 10
    ituns1: int = 5
    ituns2: int = 7
    itmxcg: int = 500
    tolcg: float = 1e-10
    deltat: float = 10.0
    dtmin: float = 1.0
    dtmax: float = 3600.0
    tmax: float = 86400.0
    dtmaga: float = 0.0
    dtmagm: float = 1.1
    dtreds: float = 0.0
    dtredm: float = 0.5
    pmin: float = -5.0
    pondh_min: float = 0.0
    scf: float = 0.0
    ieto: int = 0
    free_drainage: int = 1

    def as_dict(self):
        return asdict(self)


@dataclass
class Problem:
    name: str
    mesh: M.Mesh
    soil: dict            # permx/permy/permz/elstor/poros/vgn/vgrmc/vgpsat: (nstr, nzone); nzone
    psi0: np.ndarray
    atm_times: np.ndarray
    atm_vals: np.ndarray  # (n_atm,) if hspatm else (n_atm * nnod,), record-major
    params: Params
    seed: int = 0
    meta: dict = field(default_factory=dict)
    hspatm: int = 1       # HSPATM of the atmbc file: 1 homogeneous, 0 one value per surface node
    # listed-node boundary series (bcfiles.BCSeries: nansfdirbc / nansfneubc)
    # and the vegetation of root uptake (veg_type (nnod,) 1-based; pcana,
    # pcref, pcwlt, zroot, pz, omgc per type); None: not used
    bc_dir: object = None
    bc_neu: object = None
    veg: dict | None = None
    seepage: dict | None = None  # faces: 0-based node arrays top to bottom; isfcvg: 0 / 1
    # ISIMGR = 2: the coupled atmospheric switching (SWITCH in CONVER, ADRSTN
    # at step starts) with PONDH_MIN from the parm file; the run stops when a
    # step ends with a ponded surface node (the routing is not run)
    surf: bool = False
    # the drainage network of a coupled run (project.read_routing), or None
    routing: dict | None = None


def sig4(a: np.ndarray) -> np.ndarray:
    """Round to 4 significant digits (the %1.3e of the soil file)."""
    a = np.asarray(a, dtype=np.float64)
    out = np.zeros_like(a)
    nz = a != 0
    e = np.floor(np.log10(np.abs(a[nz])))
    out[nz] = np.round(a[nz] / 10.0 ** e, 3) * 10.0 ** e
    # exact decimal round-trip, as a text file would give
    return np.array([float(f"{v:.3e}") for v in out.ravel()]).reshape(a.shape)


def icvdwt(mesh: M.Mesh, wt_depth: float) -> np.ndarray:
    """icvdwt.f:20-25: hydrostatic head, water table wt_depth below surface."""
    nnod = mesh.nnod
    zs = np.tile(mesh.z[:nnod], mesh.nstr + 1)
    return zs - mesh.z - wt_depth


def homogeneous_soil(nstr: int, ks=1e-5, ss=1e-4, poros=0.45, vgn=1.46, vgrmc=0.05, vgpsat=1.0):
    f = lambda v: np.full((nstr, 1), v, dtype=np.float64)
    return dict(nzone=1, permx=f(ks), permy=f(ks), permz=f(ks), elstor=f(ss), poros=f(poros),
                vgn=f(vgn), vgrmc=f(vgrmc), vgpsat=f(vgpsat))


def hillslope_small_params(**kw) -> Params:
    """Solver settings under which the small hillslope's day completes.
    With DTMIN = 1 s and ITUNS = 10 the reference's own FLOW3D stalls near
    the falling water table (free drainage from a saturated base) and stops
    with NOBACK after ~100 s; DTMIN = 0.01 s and ITUNS = 30 let it back-step
    through those steps."""
    base = dict(toluns=1e-6, dtmin=0.01, ituns=30, ituns1=10, ituns2=20)
    base.update(kw)
    return Params(**base)


def hillslope_small(seed: int = 1, params: Params | None = None, nstr: int = 10, cells: int = 20) -> Problem:
    rng = np.random.default_rng(seed)
    dx = 5.0
    xc = (np.arange(cells) + 0.5) * dx
    dem = 100.0 - 0.02 * xc[None, :] + rng.normal(0.0, 0.05, (cells, cells))
    m = M.build_mesh(dem, dx, dx, nstr, [1.0 / nstr] * nstr, 5.0)
    soil = homogeneous_soil(nstr)
    p = params or hillslope_small_params()
    return Problem("hillslope_small", m, soil, icvdwt(m, 1.0), np.array([0.0]), np.array([1e-7]), p, seed,
                   dict(cells=cells, dx=dx))


def hillslope_bc(seed: int = 1, params: Params | None = None) -> Problem:
    """The small hillslope with every listed boundary condition of the
    subsurface step, over 3 h:

    * fixed heads on the downslope face (x = max), three records: t = 0
      every face node at the hydrostatic head of a water table 1.5 m deep;
      t = 1800 s the face's surface nodes at -0.3 m replicated down the
      layers (NODIN2 > 0); t = 5400 s only the face nodes deeper than 1 m,
      water table 1.2 m deep;
    * free drainage at the base plus listed Neumann fluxes on the upslope
      face (x = min): layers 3-8 at 2e-6 m^3/s each and three surface
      corner nodes at 1e-6 (t = 0), 4e-6 with one node listed twice
      (t = 2700 s), all zero (t = 6000 s);
    * root uptake with two vegetation types (upslope half ZROOT 1 m, PZ 1;
      downslope half ZROOT 0.6 m, PZ 2, other stress heads), SCF = 1, under
      a forcing that turns from 1e-7 m/s rain to 2e-7 m/s evaporation in
      the first hour (ATMPOT < 0 drives ETRAN).
    """
    from . import bcfiles as B
    p = params or hillslope_small_params(tmax=10800.0, scf=1.0)
    base = hillslope_small(seed=seed, params=p)
    m = base.mesh
    nnod, nstr = m.nnod, m.nstr
    xs = m.x[:nnod]
    zs = np.tile(m.z[:nnod], nstr + 1)
    down = np.flatnonzero(np.isclose(m.x, xs.max()))          # every layer
    up_s = np.flatnonzero(np.isclose(xs, xs.min()))           # surface ids
    layer = np.arange(m.n) // nnod
    hyd = lambda wt, ids: (zs[ids] - wt) - m.z[ids]
    deep = down[(zs[down] - m.z[down]) > 1.0]
    empty = np.zeros(0, dtype=np.int64)
    bc_dir = B.make_series("dirichlet", [
        (0.0, 0, empty, [], down, hyd(1.5, down)),
        (1800.0, len(down[down < nnod]), down[down < nnod], np.full(len(down[down < nnod]), -0.3), empty, []),
        (5400.0, 0, empty, [], deep, hyd(1.2, deep)),
    ], nnod, nstr)
    lat = np.flatnonzero(np.isclose(m.x, xs.min()) & (layer >= 3) & (layer <= 8))
    corner = up_s[:3]
    x0 = np.concatenate([lat, corner])
    q0 = np.concatenate([np.full(len(lat), 2e-6), np.full(3, 1e-6)])
    x1 = np.concatenate([lat, corner, lat[:1]])
    q1 = np.concatenate([np.full(len(lat), 4e-6), np.full(3, 1e-6), [4e-6]])
    bc_neu = B.make_series("neumann", [
        (0.0, -1, empty, [], x0, q0),
        (2700.0, -1, empty, [], x1, q1),
        (6000.0, -1, empty, [], x0, np.zeros(len(x0))),
    ], nnod, nstr)
    veg = dict(veg_type=np.where(xs < 0.5 * (xs.min() + xs.max()), 1, 2).astype(np.int32),
               pcana=np.array([0.0, -0.1]), pcref=np.array([-4.0, -2.0]), pcwlt=np.array([-150.0, -50.0]),
               zroot=np.array([1.0, 0.6]), pz=np.array([1.0, 2.0]), omgc=np.array([1.0, 0.5]))
    return Problem("hillslope_bc", m, base.soil, base.psi0, np.array([0.0, 3600.0, 10800.0]),
                   np.array([1e-7, -2e-7, -2e-7]), p, seed, dict(base.meta), 1, bc_dir, bc_neu, veg)


def hillslope_seepage(seed: int = 1, params: Params | None = None, isfcvg: int = 1) -> Problem:
    """The small hillslope with two seepage faces on its downslope face
    (the vertical node lines at the 6th and 16th surface node of x = max,
    11 nodes each, top to bottom), a water table 0.4 m deep, 5e-7 m/s rain
    and free drainage at the base, for 2 h; ISFCVG = 1 puts the exit points
    into the Picard convergence test."""
    p = params or hillslope_small_params(tmax=7200.0)
    base = hillslope_small(seed=seed, params=p)
    m = base.mesh
    nnod, nstr = m.nnod, m.nstr
    xs = m.x[:nnod]
    edge = np.flatnonzero(np.isclose(xs, xs.max()))
    edge = edge[np.argsort(m.y[edge])]
    faces = [np.array([s + l * nnod for l in range(nstr + 1)]) for s in (edge[5], edge[15])]
    return Problem("hillslope_seepage", m, base.soil, icvdwt(m, 0.4), np.array([0.0]), np.array([5e-7]), p, seed,
                   dict(base.meta), 1, seepage=dict(faces=faces, isfcvg=isfcvg))


def hillslope_small_member(seed: int, params: Params | None = None) -> Problem:
    """One perturbed-parameter member of the small hillslope (same mesh,
    initial state and settings): the soil stays uniform in space but its
    parameters are drawn per member, as a data-assimilation ensemble draws
    them: ln Ks ~ N(ln 1e-5, 0.5^2), alpha, n and porosity with 15%
    relative Gaussian spread (clipped to n in [1.15, 2.5], porosity in
    [0.25, 0.6]), and the rainfall rate 1e-7 m/s times a lognormal factor
    (ln-sd 0.5).  Realisations for training and testing the learned
    initial iterate.  (Spatially random fields on this mesh make the plain
    solver's steps collapse to sub-second sizes, in the reference's
    algorithm as here, so they are not used for training.)"""
    base = hillslope_small(params=params)
    rng = np.random.default_rng(20_000 + seed)
    nstr = base.mesh.nstr
    ks = sig4(np.exp(np.log(1e-5) + 0.5 * rng.standard_normal()))
    alpha = max(0.3, 1.0 * (1.0 + 0.15 * rng.standard_normal()))
    vgn = sig4(float(np.clip(1.46 * (1.0 + 0.15 * rng.standard_normal()), 1.15, 2.5)))
    poros = sig4(float(np.clip(0.45 * (1.0 + 0.15 * rng.standard_normal()), 0.25, 0.6)))
    rain = 1e-7 * float(np.exp(0.5 * rng.standard_normal()))
    soil = homogeneous_soil(nstr, ks=float(ks), poros=float(poros), vgn=float(vgn), vgpsat=float(sig4(1.0 / alpha)))
    return Problem("hillslope_small_member", base.mesh, soil, base.psi0, base.atm_times, np.array([rain]),
                   base.params, seed, dict(base.meta, ks=float(ks), alpha=alpha, vgn=float(vgn),
                                           poros=float(poros), rain=rain))


ENSEMBLE_SIZES = {"S": (40, 15), "M": (100, 20), "L": (200, 30)}  # cells per side, layers


def gamma_rain(hours: int = 24, mean: float = 2e-7, shape: float = 0.5, seed: int = 7):
    """Hourly rainfall intensities drawn from a gamma law (shared by every
    member of an ensemble): ATMTIM / ATMVAL with one record per hour."""
    rng = np.random.default_rng(seed)
    vals = rng.gamma(shape, mean / shape, hours)
    times = 3600.0 * np.arange(hours + 1)
    return times, np.r_[vals, vals[-1]]


def ensemble_member(size: str, seed: int, params: Params | None = None) -> Problem:
    """One member of the mixed-size ensemble: the hillslope_large generator
    (spectral DEM, 2 m cells, graded layers to 10 m, lognormal Ks field
    with ln-sd 1 and 20 m correlation, +-20% alpha / n per cell) at size
    S, M or L (40x40x15, 100x100x20, 200x200x30 cells), with the member's
    own field and its own prior draws: ln Ks mean ~ N(ln 1e-5, 0.5^2),
    alpha, n and porosity scaled by 1 + 0.15 N(0, 1) (n >= 1.1, porosity in
    [0.25, 0.6]); the shared gamma hourly rainfall over 24 h."""
    cells, nstr = ENSEMBLE_SIZES[size]
    base = hillslope_large(seed=seed, cells=cells, nstr=nstr, params=params or Params(tmax=86400.0))
    rng = np.random.default_rng(30_000 + seed)
    s = dict(base.soil)
    f_ks = float(np.exp(0.5 * rng.standard_normal()))
    f_al = 1.0 + 0.15 * rng.standard_normal()
    f_n = 1.0 + 0.15 * rng.standard_normal()
    f_po = 1.0 + 0.15 * rng.standard_normal()
    ks = sig4(np.asarray(s["permx"]) * f_ks)
    s["permx"], s["permy"], s["permz"] = ks, ks.copy(), ks.copy()
    s["vgpsat"] = sig4(np.asarray(s["vgpsat"]) / max(0.3, f_al))
    s["vgn"] = sig4(np.maximum(1.1, np.asarray(s["vgn"]) * f_n))
    s["poros"] = sig4(np.clip(np.asarray(s["poros"]) * f_po, 0.25, 0.6))
    t, v = gamma_rain()
    return Problem(f"ensemble_member_{size}", base.mesh, s, base.psi0, t, v, base.params, seed,
                   dict(base.meta, size=size))


def _spectral_dem(n: int, dx: float, relief: float, rng, beta: float = 3.0) -> np.ndarray:
    kx = np.fft.fftfreq(n, dx)
    k = np.sqrt(kx[:, None] ** 2 + kx[None, :] ** 2)
    k[0, 0] = 1.0
    amp = k ** (-beta / 2.0)
    amp[0, 0] = 0.0
    ph = rng.uniform(0, 2 * np.pi, (n, n))
    f = np.real(np.fft.ifft2(amp * np.exp(1j * ph)))
    f = (f - f.min()) / (f.max() - f.min())
    return f * relief


def _grf_exponential(n: int, dx: float, corr: float, rng) -> np.ndarray:
    """Stationary Gaussian field, exponential covariance, by circulant
    embedding on a doubled periodic grid."""
    m = 2 * n
    ix = np.minimum(np.arange(m), m - np.arange(m)) * dx
    r = np.sqrt(ix[:, None] ** 2 + ix[None, :] ** 2)
    cov = np.exp(-r / corr)
    lam = np.real(np.fft.fft2(cov))
    lam = np.maximum(lam, 0.0)
    w = rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))
    f = np.fft.fft2(np.sqrt(lam / (m * m)) * w)
    return np.real(f)[:n, :n]


def hillslope_large(seed: int = 2, params: Params | None = None, cells: int = 200, nstr: int = 30,
                    dx: float = 2.0, ratio: float = 1.1) -> Problem:
    rng = np.random.default_rng(seed)
    dem = 100.0 + _spectral_dem(cells, dx, 20.0, rng)
    zone = np.arange(1, cells * cells + 1, dtype=np.int32).reshape(cells, cells)
    zr = ratio ** np.arange(nstr)
    zr = zr / zr.sum()                     # geometric grading, thin at the top
    m = M.build_mesh(dem, dx, dx, nstr, zr, 10.0, zone=zone)
    nz = cells * cells
    lnks = np.stack([_grf_exponential(cells, dx, 20.0, rng).ravel() for _ in range(nstr)])
    ks = sig4(np.exp(np.log(1e-5) + 1.0 * lnks))
    alpha = 1.0 * (1.0 + rng.uniform(-0.2, 0.2, (nstr, nz)))
    vgn = sig4(1.46 * (1.0 + rng.uniform(-0.2, 0.2, (nstr, nz))))
    soil = dict(nzone=nz, permx=ks, permy=ks.copy(), permz=ks.copy(), elstor=np.full((nstr, nz), 1e-4),
                poros=np.full((nstr, nz), 0.45), vgn=vgn, vgrmc=np.full((nstr, nz), 0.05),
                vgpsat=sig4(1.0 / alpha))
    p = params or Params(tmax=172800.0)
    return Problem("hillslope_large", m, soil, icvdwt(m, 1.0), np.array([0.0, 21600.0, 21600.0 + 1.0]),
                   np.array([5e-7, 5e-7, 0.0]), p, seed, dict(cells=cells, dx=dx, ratio=ratio))
