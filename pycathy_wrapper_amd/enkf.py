"""EnKF analysis of a device-resident ensemble (pyCATHY DA/enkf.py:16-224).

pyCATHY forecasts every member in its own process and exchanges states
through files.  Here the members are flow handles in one process: their
states stay in HBM, the observation-space algebra (m observations x N
members) runs on the host, and the update of all N states -- the only
part that scales with the mesh -- runs on the device in place:

    P_xo COV^-1 D = X' S^T COV^-1 D / (N-1) = X' W,
    X_a = X + X' W                       (W: N x N),

with COV = S S^T/(N-1) + R (or D / diag(R) in the Sakov form), D = y - HX
and S = HX - mean(HX).  Parameters of the augmented state take the same W.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import gpu as G


def analysis_weights(data, data_cov, predict_obs, sakov: bool = False) -> np.ndarray:
    """W (N x N) with P_xo COV^-1 D = X' W for any augmented state X."""
    HX = np.asarray(predict_obs, dtype=np.float64)
    m, N = HX.shape
    if N < 2:
        raise ValueError("an ensemble needs at least two members")
    y = np.asarray(data, dtype=np.float64).reshape(-1)
    R = np.asarray(data_cov, dtype=np.float64)
    if y.size != m or R.shape != (m, m):
        raise ValueError("data (m) and data_cov (m x m) must match the predicted observations (m x N)")
    D = y[:, None] - HX
    S = HX - HX.mean(axis=1, keepdims=True)
    if sakov:
        invD = D / np.diag(R)[:, None]
    else:
        invD = np.linalg.solve(S @ S.T / (N - 1) + R, D)
    return (S.T @ invD) / (N - 1)


def update_params(param, W, inflate: float = 1.0) -> np.ndarray:
    """Analysis of parameters [n_par x N] with the same weights."""
    P = np.asarray(param, dtype=np.float64)
    Pa = P + (P - P.mean(axis=1, keepdims=True)) @ W
    if inflate != 1.0:
        mu = Pa.mean(axis=1, keepdims=True)
        Pa = mu + inflate * (Pa - mu)
    return Pa


def observe(flows, nodes, kind: str = "psi", porosity=None) -> np.ndarray:
    """Predicted point observations [m x N]: psi (tensiometers) or theta =
    Sw * porosity (soil moisture) at 0-based CATHY nodes."""
    idx = np.asarray(nodes, dtype=np.int64)
    cols = []
    for f in flows:
        if kind == "psi":
            v = f.psi()
        elif kind == "theta":
            v = f.sw() * (porosity if porosity is not None else f.node_params["pnodi"])
        else:
            raise ValueError("kind is 'psi' or 'theta'")
        cols.append(v[idx])
    return np.stack(cols, axis=1)


def enkf_update(flows, W, inflation: float = 1.0, noalloc: bool = False) -> None:
    """X_a = X + X' W on the members' PNEW in place (then inflated about the
    analysis mean); every run continues from its analysis state.
    ``noalloc``: the entry that stages W in the library's fixed device array
    instead of allocating (at most 64 members; calls on one GPU must not
    overlap)."""
    N = len(flows)
    W = np.asarray(W, dtype=np.float64)
    if W.shape != (N, N):
        raise ValueError("W must be N x N")
    h = np.array([f.h.value for f in flows], dtype=np.int64)
    w = np.ascontiguousarray(W.T)  # W(k, j) at w[j*N + k]
    entry = G.lib().cathy_gpu_enkf_update_noalloc if noalloc else G.lib().cathy_gpu_enkf_update
    G._check(entry(G._i(N), G._p(h), G._p(w), C.byref(C.c_double(inflation))),
             "enkf_update_noalloc" if noalloc else "enkf_update")
