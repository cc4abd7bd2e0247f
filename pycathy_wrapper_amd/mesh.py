"""Host-side mesh, geometry and sparsity setup (vectorised numpy).

This is the set-up the reference does once per run in Fortran before the
time loop (GRDSYS and friends, called from cathy_main.f:2559-2606):

* ``surface_from_dem``  -- triangoli.f:25-150 (+ tpnodi2d.f, riord.f):
  one node per active-cell corner, two triangles per cell split along the
  (i,j)-(i+1,j+1) diagonal, surface z = mean DEM value of incident triangles.
* ``extrude``           -- gen3d.f:24-75 (3 tets per prism, IVERT 0-4) and
  riord.f (tet vertices sorted ascending, as GRDSYS does for IOPT=1).
* ``basis``             -- volbas.f / volums.f / basis6.f: P1 coefficients
  b,c,d (already /6), |V|, sign(V), nodal volume.
* ``area_nodes``        -- area2d.f: nodal surface area.
* ``sym_pattern`` / ``full_pattern`` -- strpic.f / strnew.f: the upper
  symmetric CSR (diagonal first, sorted) and the full CSR; ``tetja_sym``
  -- tetpic.f: element-to-entry lookup.

All ids are 0-based int32 (int64 where products can overflow).  The
floating-point expressions keep the reference's operation order so the
geometry is bit-identical to the Fortran.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Mesh:
    nrow: int
    ncol: int
    nnod: int
    ntri: int
    nstr: int
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    triang: np.ndarray  # (ntri, 4) sorted vertices + zone
    tetra: np.ndarray   # (nt, 5) sorted vertices + zone (1-based)
    arenod: np.ndarray
    bi: np.ndarray = field(default=None)
    ci: np.ndarray = field(default=None)
    di: np.ndarray = field(default=None)
    volu: np.ndarray = field(default=None)
    ivol: np.ndarray = field(default=None)
    volnod: np.ndarray = field(default=None)

    @property
    def n(self) -> int:
        return self.nnod * (self.nstr + 1)

    @property
    def nt(self) -> int:
        return 3 * self.ntri * self.nstr

    def layer_of_element(self) -> np.ndarray:
        """0-based layer (ISTR-1) of each tet: asspic.f:27-29."""
        return (np.arange(self.nt, dtype=np.int64) // (3 * self.ntri)).astype(np.int32)


def surface_from_dem(dem: np.ndarray, dx: float, dy: float, zone: np.ndarray | None = None,
                     west: float = 0.0, south: float = 0.0, factor: float = 1.0):
    """triangoli.f with DOSTEP=1.  ``dem`` is (nrow, ncol), row 0 = north.
    Cells whose value is exactly 0 are inactive."""
    dem = np.asarray(dem, dtype=np.float64)
    nrow, ncol = dem.shape
    if zone is None:
        zone = np.ones((nrow, ncol), dtype=np.int32)
    active = dem != 0.0
    nodi = np.zeros((nrow + 1, ncol + 1), dtype=bool)
    nodi[:-1, :-1] |= active
    nodi[1:, :-1] |= active
    nodi[1:, 1:] |= active
    nodi[:-1, 1:] |= active
    ids = np.full((nrow + 1, ncol + 1), -1, dtype=np.int64)
    ids[nodi] = np.arange(int(nodi.sum()))
    nnod = int(nodi.sum())
    ii, jj = np.nonzero(nodi)
    x = west + jj * dx
    y = south + (nrow - ii) * dy
    ci, cj = np.nonzero(active)                       # row-major cell order
    a = ids[ci, cj]; b = ids[ci + 1, cj]; c = ids[ci + 1, cj + 1]; d = ids[ci, cj + 1]
    zn = zone[ci, cj].astype(np.int64)
    ntri = 2 * len(ci)
    tri = np.empty((ntri, 4), dtype=np.int64)
    tri[0::2] = np.stack([a, b, c, zn], axis=1)
    tri[1::2] = np.stack([a, c, d, zn], axis=1)
    el = np.repeat(dem[ci, cj] * factor / 1.0, 2)
    # tpnodi2d.f: node z = mean of incident triangle values (same add order)
    cnt = np.zeros(nnod, dtype=np.int64)
    # sequential accumulation in triangle order, as the Fortran loop does
    order_nodes = tri[:, :3].reshape(-1)
    order_vals = np.repeat(el, 3)
    z = _ordered_scatter_add(nnod, order_nodes, order_vals)
    np.add.at(cnt, order_nodes, 1)
    z = z / cnt
    tri[:, :3] = np.sort(tri[:, :3], axis=1)          # riord.f
    return nnod, ntri, x, y, z, tri.astype(np.int32)


def dem_cells(dem: np.ndarray) -> np.ndarray:
    """CELL(1:4, ICELL) of triangoli.f:88-94 (DOSTEP = 1): the corner nodes
    (0-based) of each active DEM cell, row by row, in the order NW, SW, SE,
    NE -- the node numbering of ``surface_from_dem``."""
    dem = np.asarray(dem, dtype=np.float64)
    nrow, ncol = dem.shape
    active = dem != 0.0
    nodi = np.zeros((nrow + 1, ncol + 1), dtype=bool)
    nodi[:-1, :-1] |= active
    nodi[1:, :-1] |= active
    nodi[1:, 1:] |= active
    nodi[:-1, 1:] |= active
    ids = np.full((nrow + 1, ncol + 1), -1, dtype=np.int64)
    ids[nodi] = np.arange(int(nodi.sum()))
    ci, cj = np.nonzero(active)
    return np.stack([ids[ci, cj], ids[ci + 1, cj], ids[ci + 1, cj + 1], ids[ci, cj + 1]], axis=1).astype(np.int32)


def _ordered_scatter_add(size: int, idx: np.ndarray, vals: np.ndarray) -> np.ndarray:
    """out[idx[k]] += vals[k] in the order k = 0, 1, ... (np.add.at is
    unbuffered and sequential, which keeps the reference's summation order)."""
    out = np.zeros(size)
    np.add.at(out, idx, vals)
    return out


def extrude(nnod: int, nstr: int, triang: np.ndarray, x: np.ndarray, y: np.ndarray, z: np.ndarray,
            zratio, base: float, ivert: int = 0, depth: np.ndarray | None = None):
    """gen3d.f: node planes 0..nstr (plane 0 = surface), 3 tets per prism in
    layer J (between planes J-1 and J), then riord.f on the 4 vertices."""
    ntri = triang.shape[0]
    t = triang.astype(np.int64)
    J = np.arange(1, nstr + 1, dtype=np.int64)[:, None]
    lo = (J - 1) * nnod
    hi = J * nnod
    t0, t1, t2, zn = t[:, 0][None], t[:, 1][None], t[:, 2][None], t[:, 3][None]
    tet = np.empty((nstr, ntri, 3, 5), dtype=np.int64)
    tet[:, :, 0, 0] = lo + t0; tet[:, :, 0, 1] = lo + t1; tet[:, :, 0, 2] = lo + t2
    tet[:, :, 0, 3] = hi + t0
    tet[:, :, 1, 0] = hi + t0; tet[:, :, 1, 1] = hi + t1; tet[:, :, 1, 2] = hi + t2
    tet[:, :, 1, 3] = lo + t2
    tet[:, :, 2, 0] = lo + t1; tet[:, :, 2, 1] = lo + t2; tet[:, :, 2, 2] = hi + t1
    tet[:, :, 2, 3] = hi + t0
    tet[:, :, :, 4] = zn[..., None]
    tet = tet.reshape(-1, 5)
    tet[:, :4] = np.sort(tet[:, :4], axis=1)
    zr = np.asarray(zratio, dtype=np.float64)
    n = nnod * (nstr + 1)
    X = np.empty(n); Y = np.empty(n); Z = np.empty(n)
    X[:nnod] = x; Y[:nnod] = y; Z[:nnod] = z
    zmin = min(1.7e100, float(np.min(z)))
    zthick = (z - zmin) + base
    zrsum = np.zeros(nnod)
    for j in range(1, nstr + 1):
        sl = slice(j * nnod, (j + 1) * nnod)
        X[sl] = x; Y[sl] = y
        zrsum = zrsum + zr[j - 1]
        if ivert == 0:
            Z[sl] = z - zrsum * base
        elif ivert == 1:
            Z[sl] = z - zrsum * zthick
        elif ivert == 2:
            Z[sl] = zmin - zrsum * base
        elif ivert == 3:
            Z[sl] = z - zrsum * depth
        elif ivert == 4:
            Z[sl] = z - zrsum * base
            if j == nstr:
                Z[sl] = zmin - base
        else:
            raise ValueError(f"IVERT={ivert}")
    return X, Y, Z, tet.astype(np.int32)


# cyclic vertex triples used by volums.f/basis6.f for local vertex nn
_CYC = {0: (1, 2, 3), 1: (2, 3, 0), 2: (3, 0, 1), 3: (0, 1, 2)}


def basis(tetra: np.ndarray, x: np.ndarray, y: np.ndarray, z: np.ndarray, n: int):
    """volbas.f: returns bi, ci, di (nt,4), volu, ivol (int32), volnod."""
    t = tetra[:, :4].astype(np.int64)
    nt = t.shape[0]
    amen4 = (-1.0, 1.0, -1.0, 1.0)
    amen5 = (-1.0, 1.0, -1.0, 1.0, -1.0)
    vol = np.zeros(nt)
    bi = np.empty((nt, 4)); ci = np.empty((nt, 4)); di = np.empty((nt, 4))
    for nn in range(4):
        c0 = _CYC[nn]
        trip = [(c0[ii], c0[(ii + 1) % 3], c0[(ii + 2) % 3]) for ii in range(3)]
        a3 = np.zeros(nt); a2 = np.zeros(nt)
        for (I, J, M) in trip:
            a3 = y[t[:, I]] * z[t[:, J]] + a3
            a2 = y[t[:, I]] * z[t[:, M]] + a2
        vol = vol + x[t[:, nn]] * amen4[nn] * (a3 - a2) / 6.0
        bi[:, nn] = amen5[nn] * (a3 - a2) / 6.0
        a3 = np.zeros(nt); a2 = np.zeros(nt)
        for (I, J, M) in trip:
            a3 = x[t[:, I]] * z[t[:, J]] + a3
            a2 = x[t[:, I]] * z[t[:, M]] + a2
        ci[:, nn] = amen5[nn + 1] * (a3 - a2) / 6.0
        a3 = np.zeros(nt); a2 = np.zeros(nt)
        for (I, J, M) in trip:
            a3 = x[t[:, I]] * y[t[:, J]] + a3
            a2 = x[t[:, I]] * y[t[:, M]] + a2
        di[:, nn] = amen5[nn] * (a3 - a2) / 6.0
    if np.any(vol == 0.0):
        raise ValueError("zero-volume tetrahedron (volbas.f)")
    volnod = volnod_ordered(tetra, np.abs(vol), n)
    ivol = np.where(vol < 0.0, -1, 1).astype(np.int32)
    volu = np.abs(vol)
    return bi, ci, di, volu, ivol, volnod


def volnod_ordered(tetra: np.ndarray, volu: np.ndarray, n: int) -> np.ndarray:
    """volnod accumulated element by element, vertex by vertex (volbas.f
    order), for bit-exact comparisons."""
    t = tetra[:, :4].astype(np.int64)
    vol4 = volu * 0.25
    return _ordered_scatter_add(n, t.reshape(-1), np.repeat(vol4, 4))


def area_nodes(nnod: int, triang: np.ndarray, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """area2d.f: 1/3 of each incident triangle's area."""
    t = triang[:, :3].astype(np.int64)
    cyc = ((0, 1, 2), (1, 2, 0), (2, 0, 1))
    a3 = np.zeros(t.shape[0]); a2 = np.zeros(t.shape[0])
    for (I, J, M) in cyc:
        a3 = x[t[:, I]] * y[t[:, J]] + a3
        a2 = x[t[:, I]] * y[t[:, M]] + a2
    are = 0.5 * (a3 - a2)
    if np.any(are == 0.0):
        raise ValueError("zero-area triangle (area2d.f)")
    are3 = np.abs(are) * (1.0 / 3.0)
    return _ordered_scatter_add(nnod, t.reshape(-1), np.repeat(are3, 3))


def _edge_keys(tetra: np.ndarray, n: int) -> np.ndarray:
    t = tetra[:, :4].astype(np.int64)
    pairs = []
    for a in range(4):
        for b in range(a + 1, 4):
            lo = np.minimum(t[:, a], t[:, b]); hi = np.maximum(t[:, a], t[:, b])
            pairs.append(lo * n + hi)
    return np.unique(np.concatenate(pairs))


def sym_pattern(tetra: np.ndarray, n: int):
    """strpic.f: upper CSR, row k = [k, sorted columns > k].  Returns
    (topol int32 (n+1), ja int32 (nterm))."""
    keys = _edge_keys(tetra, n)
    diag = np.arange(n, dtype=np.int64) * n + np.arange(n, dtype=np.int64)
    keys = np.sort(np.concatenate([keys, diag]))
    rows = keys // n
    cols = keys - rows * n
    topol = np.zeros(n + 1, dtype=np.int64)
    np.add.at(topol, rows + 1, 1)
    topol = np.cumsum(topol)
    return topol.astype(np.int32), cols.astype(np.int32)


def full_pattern(tetra: np.ndarray, n: int):
    """strnew.f: full CSR with sorted columns (diagonal in place)."""
    keys = _edge_keys(tetra, n)
    r = keys // n; c = keys - r * n
    diag = np.arange(n, dtype=np.int64)
    rows = np.concatenate([r, c, diag]); cols = np.concatenate([c, r, diag])
    k2 = np.sort(rows * n + cols)
    rows = k2 // n; cols = k2 - rows * n
    topol = np.zeros(n + 1, dtype=np.int64)
    np.add.at(topol, rows + 1, 1)
    topol = np.cumsum(topol)
    return topol.astype(np.int32), cols.astype(np.int32)


def csr_lookup(topol: np.ndarray, ja: np.ndarray, rows: np.ndarray, cols: np.ndarray, n: int) -> np.ndarray:
    """Position of (rows[i], cols[i]) in a CSR with sorted columns per row
    (the diagonal-first upper CSR is also sorted).  Raises if absent."""
    rp = np.asarray(topol, dtype=np.int64)
    rowid = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    keys = rowid * n + np.asarray(ja, dtype=np.int64)
    q = rows.astype(np.int64) * n + cols.astype(np.int64)
    pos = np.searchsorted(keys, q)
    if np.any(pos >= keys.size) or np.any(keys[np.minimum(pos, keys.size - 1)] != q):
        raise ValueError("entry missing from pattern")
    return pos


def tetja_sym(tetra: np.ndarray, topol: np.ndarray, ja: np.ndarray, n: int) -> np.ndarray:
    """tetpic.f: (nt, 10) positions of (t_k, t_l), k <= l, in row-major
    (k,l) order (0,0),(0,1),(0,2),(0,3),(1,1),(1,2),(1,3),(2,2),(2,3),(3,3)."""
    t = tetra[:, :4].astype(np.int64)
    out = []
    for k in range(4):
        for l in range(k, 4):
            out.append(csr_lookup(topol, ja, t[:, k], t[:, l], n))
    return np.stack(out, axis=1).astype(np.int32)


UPPER_PAIRS = [(k, l) for k in range(4) for l in range(k, 4)]


def build_mesh(dem: np.ndarray, dx: float, dy: float, nstr: int, zratio, base: float,
               zone: np.ndarray | None = None, ivert: int = 0, west: float = 0.0, south: float = 0.0,
               factor: float = 1.0) -> Mesh:
    nrow, ncol = dem.shape
    nnod, ntri, x, y, z, tri = surface_from_dem(dem, dx, dy, zone, west, south, factor)
    X, Y, Z, tet = extrude(nnod, nstr, tri, x, y, z, zratio, base, ivert)
    are = area_nodes(nnod, tri, x, y)
    m = Mesh(nrow=nrow, ncol=ncol, nnod=nnod, ntri=ntri, nstr=nstr, x=X, y=Y, z=Z, triang=tri,
             tetra=tet, arenod=are)
    m.bi, m.ci, m.di, m.volu, m.ivol, m.volnod = basis(tet, X, Y, Z, m.n)
    return m
