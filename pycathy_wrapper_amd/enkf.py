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


# ----------------------------------------- localised analysis (f2 rest) --

def observation_terms(data, data_cov, predict_obs, sakov: bool = False):
    """S = HX - mean(HX) and COV^-1 D (or D / diag(R) in the Sakov form),
    each m x N: the observation-space half of enkf.py:290-301."""
    HX = np.asarray(predict_obs, dtype=np.float64)
    m, N = HX.shape
    if N < 2:
        raise ValueError("an ensemble needs at least two members")
    y = np.asarray(data, dtype=np.float64).reshape(-1)
    R = np.asarray(data_cov, dtype=np.float64)
    if y.size != m or R.shape != (m, m):
        raise ValueError("data (m) and data_cov (m x m) must match the predicted observations (m x N)")
    S = HX - HX.mean(axis=1, keepdims=True)
    D = y[:, None] - HX
    invD = D / np.diag(R)[:, None] if sakov else np.linalg.solve(S @ S.T / (N - 1) + R, D)
    return S, invD


def gaspari_cohn(r, L):
    """Gaspari-Cohn weights of distances r for radius L
    (DA/localisation.py:136-151), 0 from r = 2L on."""
    z = np.abs(np.asarray(r, dtype=np.float64)) / L
    w = np.zeros_like(z)
    a, b = z <= 1, (z > 1) & (z <= 2)
    za, zb = z[a], z[b]
    w[a] = (((-0.25 * za + 0.5) * za + 0.625) * za - 5.0 / 3.0) * (za * za) + 1.0
    w[b] = ((((zb / 12.0 - 0.5) * zb + 0.625) * zb + 5.0 / 3.0) * zb - 5.0) * zb + 4.0 - 2.0 / (3.0 * zb)
    return w


def localization_matrix(obs_pos, grid_pos, L) -> np.ndarray:
    """L(i, j) over (x, y) positions (DA/localisation.py:154-188), for
    callers that want the dense matrix; the device forms it row by row."""
    o = np.asarray(obs_pos, dtype=np.float64)[:, :2]
    g = np.asarray(grid_pos, dtype=np.float64)[:, :2]
    d = np.sqrt((o[None, :, 0] - g[:, None, 0]) ** 2 + (o[None, :, 1] - g[:, None, 1]) ** 2)
    return gaspari_cohn(d, L)


def enkf_update_localized(flows, S, invD, L=None, obs_xy=None, node_xy=None, radius=None,
                          inflation: float = 1.0) -> None:
    """The states' localised analysis on the device, in place
    (cathy_gpu_enkf_update_localized): X_a = X + ((X' S^T/(N-1)) * L) invD,
    then inflation; L dense (N_state x m, CATHY node order) or Gaspari-Cohn
    of radius ``radius`` between node_xy (default: the members' mesh) and
    obs_xy."""
    N = len(flows)
    S = np.ascontiguousarray(S, dtype=np.float64)
    invD = np.ascontiguousarray(invD, dtype=np.float64)
    if S.ndim != 2 or S.shape[1] != N or invD.shape != S.shape:
        raise ValueError("S and invD must both be m x N")
    m = S.shape[0]
    h = np.array([f.h.value for f in flows], dtype=np.int64)
    null = C.c_void_p()
    if L is not None:
        L = np.ascontiguousarray(L, dtype=np.float64)
        if L.shape != (flows[0].n, m):
            raise ValueError("L must be N_state x m")
        args = (G._p(L), null, null, null)
    else:
        if obs_xy is None or radius is None:
            raise ValueError("give L, or observation positions and a radius")
        if node_xy is None:
            mesh = flows[0].problem.mesh
            node_xy = np.stack([mesh.x, mesh.y], axis=1)
        node_xy = np.ascontiguousarray(node_xy, dtype=np.float64)[:, :2].copy()
        obs_xy = np.ascontiguousarray(obs_xy, dtype=np.float64)[:, :2].copy()
        if node_xy.shape[0] != flows[0].n or obs_xy.shape[0] != m:
            raise ValueError("node_xy must be N_state x 2 and obs_xy m x 2")
        args = (null, G._p(node_xy), G._p(obs_xy), C.byref(C.c_double(radius)))
    G._check(G.lib().cathy_gpu_enkf_update_localized(G._i(N), G._p(h), G._i(m), G._p(S), G._p(invD), *args,
                                                      C.byref(C.c_double(inflation))), "enkf_update_localized")


def enkf_analysis_localized_with_inflation(data, data_cov, flows, param, predict_obs, L=None, obs_xy=None,
                                           radius=None, sakov: bool = False, inflate_states: float = 1.0,
                                           inflate_params: float = 1.0, jitter_params: float = 0.0, rng=None):
    """pyCATHY's enkf_analysis_localized_with_inflation (enkf.py:225-343)
    with the states resident in the members: the states are analysed on
    the device in place (localised by L, or Gaspari-Cohn of ``radius``
    around obs_xy), the parameters [n_par x N] on the host without
    localisation, inflated and jittered (np.random.normal draws, as the
    reference).  Returns the analysis parameters."""
    S, invD = observation_terms(data, data_cov, predict_obs, sakov)
    N = len(flows)
    P = np.asarray(param, dtype=np.float64).reshape(-1, N) if np.size(param) else np.zeros((0, N))
    Pa = P + ((P - P.mean(axis=1, keepdims=True)) @ S.T / (N - 1)) @ invD
    enkf_update_localized(flows, S, invD, L=L, obs_xy=obs_xy, radius=radius, inflation=inflate_states)
    if inflate_params != 1.0:
        mu = Pa.mean(axis=1, keepdims=True)
        Pa = mu + inflate_params * (Pa - mu)
    if jitter_params > 0.0:
        Pa = Pa + (rng or np.random).normal(0, jitter_params, Pa.shape)
    return Pa


def zone_analysis(data, data_cov, flows, predict_obs, surface_zone, obs_nodes, param=None, sakov: bool = False,
                  inflation: float = 1.0, skip_zone: int = -9999):
    """pyCATHY's domain localisation (cathy_DA.py:671-803): one analysis per
    zone of the surface map (veg_map, zones, or one zone per node), each
    over the 3D nodes under that zone's surface nodes
    (localisation.py:12-48) and the observations at them; zone
    ``skip_zone`` is not analysed.  All zones' states move in one device
    launch (cathy_gpu_enkf_update_zones).  Returns the analysis parameters
    of each zone with observations, {zone: [n_par x N]}."""
    N = len(flows)
    HX = np.asarray(predict_obs, dtype=np.float64)
    m = HX.shape[0]
    y = np.asarray(data, dtype=np.float64).reshape(-1)
    R = np.asarray(data_cov, dtype=np.float64)
    nnod = flows[0].nnod
    sz = np.asarray(surface_zone).reshape(-1)
    if sz.size != nnod:
        raise ValueError("surface_zone needs one zone per surface node")
    row_zone = np.ascontiguousarray(sz[np.arange(flows[0].n) % nnod], dtype=np.int32)
    oz = sz[np.asarray(obs_nodes, dtype=np.int64) % nnod].astype(np.int32)
    none = np.int32(np.iinfo(np.int32).min)                    # matches no row
    obs_zone = np.full(m, none, dtype=np.int32)
    S = np.zeros((m, N))
    invD = np.zeros((m, N))
    P = np.asarray(param, dtype=np.float64).reshape(-1, N) if param is not None and np.size(param) else None
    out = {}
    for z in np.unique(oz):
        if z == skip_zone:
            continue
        sel = np.flatnonzero(oz == z)
        Sz, Dz = observation_terms(y[sel], R[np.ix_(sel, sel)], HX[sel], sakov)
        S[sel], invD[sel], obs_zone[sel] = Sz, Dz, z
        if P is not None:
            out[int(z)] = P + ((P - P.mean(axis=1, keepdims=True)) @ Sz.T / (N - 1)) @ Dz
    h = np.array([f.h.value for f in flows], dtype=np.int64)
    S, invD = np.ascontiguousarray(S), np.ascontiguousarray(invD)
    G._check(G.lib().cathy_gpu_enkf_update_zones(G._i(N), G._p(h), G._i(m), G._p(S), G._p(invD), G._p(row_zone),
                                                  G._p(obs_zone), C.byref(C.c_double(inflation))), "enkf_update_zones")
    return out


# ------------------------------------------------------ particle filter --

def particle_weights(data, data_cov, predict_obs):
    """Gaussian likelihood weights of the members and N_eff (pf.py:53-81)."""
    HX = np.asarray(predict_obs, dtype=np.float64)
    y = np.atleast_1d(np.asarray(data, dtype=np.float64).flatten())
    R = np.asarray(data_cov, dtype=np.float64)
    sd = np.sqrt(np.diag(R)) if R.ndim == 2 else np.sqrt(R)
    lw = -0.5 * np.sum(((y[:, None] - HX) / sd[:, None]) ** 2, axis=0)
    w = np.exp(lw - lw.max())
    w /= w.sum()
    return w, 1.0 / np.sum(w ** 2)


def systematic_resample(weights, n: int, u=None) -> np.ndarray:
    """Systematic resampling (pf.py:197-211): positions (k + u)/n with one
    uniform u (drawn by np.random.rand() when not given)."""
    if u is None:
        u = np.random.rand()
    return np.searchsorted(np.cumsum(weights), (np.arange(n) + u) / n)


def pf_resample(flows, idx, jitter_std: float = 0.0, seed: int = 0) -> None:
    """Member j's state becomes member idx[j]'s, plus jitter_std times a
    standard normal per node and member (cathy_gpu_pf_resample)."""
    N = len(flows)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    if idx.shape != (N,):
        raise ValueError("one index per member")
    h = np.array([f.h.value for f in flows], dtype=np.int64)
    G._check(G.lib().cathy_gpu_pf_resample(G._i(N), G._p(h), G._p(idx), C.byref(C.c_double(jitter_std)),
                                           C.byref(C.c_int64(int(seed)))), "pf_resample")


def weighted_update_weights(data, data_cov, predict_obs, weights, data_noise, alpha: float = 1.0,
                            gain_factor: float = 1.0) -> np.ndarray:
    """W (N x N) with pf.py's weighted EnKF update (:264-352) written as
    X_a = X + X' W for any augmented state X.  The update there is
    A + gain dA_w S_w^T B with dA_w = sqrt(alpha)(A - A w 1^T) diag(sqrt w):
    linear in A, A (I + V) with V = gain sqrt(alpha) (I - w 1^T)
    diag(sqrt w) S_w^T B; 1^T V = 0 because the weights sum to 1, so
    X' V = X V and W = V."""
    HX = np.asarray(predict_obs, dtype=np.float64)
    m, N = HX.shape
    y = np.asarray(data, dtype=np.float64).reshape(-1)
    R = np.asarray(data_cov, dtype=np.float64)
    w = np.asarray(weights, dtype=np.float64)
    hm = HX @ w
    H_inf = np.sqrt(alpha) * (HX - hm[:, None]) + hm[:, None]
    dD = (y[:, None] + np.asarray(data_noise, dtype=np.float64) * np.sqrt(np.diag(R))[:, None]) - H_inf
    Sw = (H_inf - hm[:, None]) * np.sqrt(w)[None, :]
    B = np.linalg.solve(Sw @ Sw.T + R + 1e-6 * np.eye(m), dD)
    core = np.sqrt(w)[:, None] * (Sw.T @ B)                  # diag(sqrt w) S_w^T B
    V = gain_factor * np.sqrt(alpha) * (core - w[:, None] * core.sum(axis=0)[None, :])
    return V


def particle_filter_analysis(data, data_cov, param, flows, predict_obs, use_enkf_update: bool = False,
                             resample_threshold: float = 0.5, jitter_std_param: float = 0.01,
                             jitter_std_state: float = 0.005, alpha: float = 1.0, seed: int = 0, u=None,
                             rng=None, use_log_K: bool = False) -> dict:
    """pyCATHY's particle_filter_analysis (pf.py:3-195) with the states
    resident in the members: weights and N_eff on the host; when N_eff <
    threshold * N, systematic resampling of the states on the device
    (pf_resample) and of parameters and predicted observations on the host;
    the optional weighted EnKF update (enkf_update with
    weighted_update_weights, gain 1 as pf.py passes it); state jitter on
    the device (Philox draws under ``seed``), parameter jitter on the host
    (``rng`` normal draws, np.random by default; ``use_log_K``: the second
    half of the parameters, K, jittered multiplicatively).  Returns the reference's
    dict without 'Analysis' (the states are the members')."""
    rng = rng or np.random
    N = len(flows)
    HX = np.asarray(predict_obs, dtype=np.float64)
    if HX.shape[0] == N and HX.shape[1] != N:
        HX = HX.T
    P = np.asarray(param, dtype=np.float64).reshape(-1, N) if np.size(param) else np.zeros((0, N))
    w, n_eff = particle_weights(data, data_cov, HX)
    resampled = bool(n_eff < resample_threshold * N)
    idx = np.arange(N)
    if resampled:
        idx = systematic_resample(w, N, u)
        P, HX = P[:, idx], HX[:, idx]
        w = np.ones(N) / N
    jitter = resampled or jitter_std_param > 0
    sigma = jitter_std_state if jitter and jitter_std_state > 0 else 0.0
    if use_enkf_update:
        if resampled:
            pf_resample(flows, idx, 0.0, seed)
        noise = rng.standard_normal((HX.shape[0], N))
        V = weighted_update_weights(data, data_cov, HX, w, noise, alpha)
        P = P + P @ V
        enkf_update(flows, V)
        if sigma > 0:
            pf_resample(flows, np.arange(N), sigma, seed)
    elif resampled or sigma > 0:
        pf_resample(flows, idx, sigma, seed)
    if jitter and jitter_std_param > 0:
        z = rng.standard_normal(P.shape)
        if use_log_K:  # pf.py:136-144: porosity (first half) additive, K (second half) multiplicative
            nc = P.shape[0] // 2
            P = P.copy()
            P[:nc] = P[:nc] + z[:nc] * jitter_std_param
            P[nc:] = P[nc:] * np.exp(z[nc:] * jitter_std_param)
        else:
            P = P + z * jitter_std_param
    return {"Analysisparam": P, "weights": w, "n_eff": n_eff, "resampled": resampled, "observation": HX,
            "indices": idx}


# -------------------------------------------------------- dual analysis --

def state_gram(flows) -> np.ndarray:
    """G = X'^T X' over the members' resident states (cathy_gpu_enkf_gram)."""
    N = len(flows)
    h = np.array([f.h.value for f in flows], dtype=np.int64)
    g = np.zeros((N, N))
    G._check(G.lib().cathy_gpu_enkf_gram(G._i(N), G._p(h), G._p(g)), "enkf_gram")
    return g


def dual_weights(data, data_cov, predict_obs, gram, sakov: bool = False) -> np.ndarray:
    """W2 (N x N) of pyCATHY's enkf_dual_analysis (enkf.py:345-361) from the
    observations and the states' Gram matrix G = X'^T X'.  The first step is
    X + X' W1 (W1 = S^T COV^-1 D/(N-1)); the second takes the first's states
    as per-member data of the forecast states themselves, with R2 = c I,
    c = mean(diag(COV1)), so S2 = X' and D2 = X' W1.  By the push-through
    identity S2^T (S2 S2^T/(N-1) + c I)^-1 = (G/(N-1) + c I)^-1 S2^T:
        W2 = (G/(N-1) + c I)^-1 G W1 / (N-1)     (Sakov: G W1 / (c (N-1))),
    N x N algebra instead of the reference's N_state x N_state covariance.
    The dual's result is X + X' W2, parameters P + P' W2."""
    S, invD = observation_terms(data, data_cov, predict_obs, sakov)
    m, N = S.shape
    R = np.asarray(data_cov, dtype=np.float64)
    cov1 = R if sakov else S @ S.T / (N - 1) + R
    c = float(np.mean(np.diag(cov1)))
    W1 = S.T @ invD / (N - 1)
    Gm = np.asarray(gram, dtype=np.float64)
    if sakov:
        return Gm @ W1 / (c * (N - 1))
    return np.linalg.solve(Gm / (N - 1) + c * np.eye(N), Gm @ W1) / (N - 1)


def enkf_dual_analysis(data, data_cov, param, flows, predict_obs, sakov: bool = False):
    """pyCATHY's enkf_dual_analysis on resident members: the states' Gram
    matrix on the device, W2 on the host, X + X' W2 on the device in place.
    Returns the analysis parameters [n_par x N]."""
    N = len(flows)
    W2 = dual_weights(data, data_cov, predict_obs, state_gram(flows), sakov)
    P = np.asarray(param, dtype=np.float64).reshape(-1, N) if param is not None and np.size(param) else \
        np.zeros((0, N))
    Pa = P + (P - P.mean(axis=1, keepdims=True)) @ W2
    enkf_update(flows, W2)
    return Pa
