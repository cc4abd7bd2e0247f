"""Node soil parameters from per-(layer, zone) tables (host set-up).

``node_params`` restates tpnodi.f:28-70 (IPEAT=0): every node gets the mean
of porosity, specific storage and the van Genuchten n, theta_r, psi_sat of
its incident tets, accumulated element by element in the reference's order.
``kz_nodes`` restates cathy_main.f:2687-2694 (ELTNOD of the per-element Kz).
Tables are (nstr, nzone) arrays indexed [layer, zone-1].
"""
from __future__ import annotations

import numpy as np

from .mesh import Mesh


def _node_mean(mesh: Mesh, elem_vals: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    # bincount adds the weights in input order: element by element, the
    # order of the reference's accumulation loop
    t = mesh.tetra[:, :4].astype(np.int64).reshape(-1)
    acc = np.bincount(t, weights=np.repeat(np.asarray(elem_vals, dtype=np.float64), 4), minlength=mesh.n)
    tp = np.bincount(t, minlength=mesh.n).astype(np.int64)
    return acc, tp


def node_params(mesh: Mesh, soil: dict) -> dict:
    lay = mesh.layer_of_element().astype(np.int64)
    zone = mesh.tetra[:, 4].astype(np.int64) - 1
    out = {}
    tp = None
    for key, name in (("poros", "pnodi"), ("elstor", "snodi"), ("vgn", "vgn"), ("vgrmc", "vgrmc"),
                      ("vgpsat", "vgpsat")):
        acc, tp = _node_mean(mesh, np.asarray(soil[key])[lay, zone])
        if np.any(tp == 0):
            raise ValueError("node not connected to any element (tpnodi.f)")
        out[name] = acc / tp
    out["tp"] = tp
    return out


def kz_nodes(mesh: Mesh, soil: dict) -> np.ndarray:
    lay = mesh.layer_of_element().astype(np.int64)
    zone = mesh.tetra[:, 4].astype(np.int64) - 1
    acc, tp = _node_mean(mesh, np.asarray(soil["permz"])[lay, zone])
    return acc / tp
