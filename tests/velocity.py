"""Oracle calls for VEL3D / VNOD3D shared by the CPU and GPU velocity tests."""
from __future__ import annotations

import numpy as np

from pyoracle import dptr, iptr


def layered(problem, scale=(1.0, 2.0, 0.5)):
    """The problem's soil with PERMX/Y/Z made different per axis and per
    layer, so that a swapped axis or a wrong layer index shows."""
    s = dict(problem.soil)
    nstr = problem.mesh.nstr
    lay = (1.0 + 0.37 * np.arange(nstr))[:, None]
    for k, f in zip(("permx", "permy", "permz"), scale):
        s[k] = np.ascontiguousarray(np.asarray(s[k]) * lay * f)
    problem.soil = s
    return problem


def oracle_vel3d(orc, problem, psi, ckrwe, volur):
    """orc_vel3d: UU, VV, WW (NT)."""
    m, s = problem.mesh, problem.soil
    tet = np.ascontiguousarray(m.tetra, dtype=np.int32)
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    out = [np.zeros(m.nt) for _ in range(3)]
    orc.L.orc_vel3d(m.nt, m.ntri, int(s["nzone"]), iptr(tet), dptr(f(psi)), dptr(f(s["permx"])),
                    dptr(f(s["permy"])), dptr(f(s["permz"])), *(dptr(o) for o in out), dptr(f(m.bi)),
                    dptr(f(m.ci)), dptr(f(m.di)), dptr(f(ckrwe)), dptr(f(volur)),
                    iptr(np.ascontiguousarray(m.ivol, dtype=np.int32)))
    return out


def oracle_vnod3d(orc, problem, tp, uu, vv, ww):
    """orc_vnod3d: UNOD, VNOD, WNOD (N)."""
    m = problem.mesh
    tet = np.ascontiguousarray(m.tetra, dtype=np.int32)
    out = [np.zeros(m.n) for _ in range(3)]
    orc.L.orc_vnod3d(m.n, m.nt, iptr(np.ascontiguousarray(tp, dtype=np.int32)), iptr(tet), dptr(uu), dptr(vv),
                     dptr(ww), *(dptr(o) for o in out))
    return out


def linear_head(problem, a, b, g, c=-2.0):
    """psi = a x + b y + c - (1 + g) z: total head h = psi + z has the
    constant gradient (a, b, -g), so the Darcy flux is -kr K grad h =
    (-kr Kx a, -kr Ky b, kr Kz g) in every element."""
    m = problem.mesh
    return a * m.x + b * m.y + c - (1.0 + g) * m.z


def elem_perm(problem):
    """Kx, Ky, Kz of each element (layer = e // (3 NTRI), zone TETRA(5))."""
    m, s = problem.mesh, problem.soil
    istr = np.arange(m.nt) // (3 * m.ntri)
    mt = np.asarray(m.tetra)[:, 4] - 1
    return [np.asarray(s[k])[istr, mt] for k in ("permx", "permy", "permz")]
