"""CPU restatement of pyCATHY's EnKF analysis (test infrastructure).

Only tests/ use this module.  It follows DA/enkf.py:16-224
(``enkf_analysis``) step by step on the augmented state [states; params]:
ensemble means and perturbations (:83-112), innovation D = y - HX
(:121-129), predicted-observation perturbations S (:144-148), the
observation covariance COV = S S^T/(N-1) + R and COV^-1 D (:162-169; the
Sakov form uses D / diag(R), :163-165), the cross covariance
P_xo = X' S^T/(N-1) (:180), and X_a = X + P_xo COV^-1 D (:194); state
inflation about the analysis mean as in
``enkf_analysis_localized_with_inflation`` (enkf.py:316-318).

Also restated: that localised analysis itself (enkf.py:225-343) with the
Gaspari-Cohn localisation matrix of DA/localisation.py:136-188, and the
particle filter of DA/pf.py (weights, systematic resampling, the weighted
EnKF update, additive jitter), with every random draw of the reference
passed in as an argument.
"""
from __future__ import annotations

import numpy as np


def enkf_analysis(data, data_cov, param, ensemble, predict_obs, sakov=False, inflate_states=1.0):
    """Returns (analysis states [n_state x N], analysis params [n_par x N])."""
    X = np.asarray(ensemble, dtype=np.float64)
    P = np.asarray(param, dtype=np.float64).reshape(-1, X.shape[1]) if np.size(param) else np.zeros((0, X.shape[1]))
    HX = np.asarray(predict_obs, dtype=np.float64)
    y = np.asarray(data, dtype=np.float64).reshape(-1)
    R = np.asarray(data_cov, dtype=np.float64)
    N = X.shape[1]
    n = X.shape[0]
    A = np.vstack([X, P])
    Ap = A - A.mean(axis=1, keepdims=True)
    D = y[:, None] - HX
    S = HX - HX.mean(axis=1, keepdims=True)
    if sakov:
        invD = D / np.diag(R)[:, None]
    else:
        COV = (S @ S.T) / (N - 1) + R
        invD = np.linalg.solve(COV, D)
    Pxo = (Ap @ S.T) / (N - 1)
    Aa = A + Pxo @ invD
    Xa, Pa = Aa[:n], Aa[n:]
    if inflate_states != 1.0:
        m = Xa.mean(axis=1, keepdims=True)
        Xa = m + inflate_states * (Xa - m)
    return Xa, Pa


# --------------------------------------------------------- localisation --

def gaspari_cohn(r, L):
    """Gaspari-Cohn weight of distance r for radius L
    (DA/localisation.py:136-151): the fifth-order piecewise rational
    function, 1 at r = 0 and 0 from r = 2L on."""
    z = np.abs(np.asarray(r, dtype=np.float64)) / L
    w = np.zeros_like(z)
    a = z <= 1
    b = (z > 1) & (z <= 2)
    za, zb = z[a], z[b]
    w[a] = (((-0.25 * za + 0.5) * za + 0.625) * za - 5.0 / 3.0) * za ** 2 + 1
    w[b] = ((((zb / 12 - 0.5) * zb + 0.625) * zb + 5.0 / 3.0) * zb - 5) * zb + 4 - 2 / (3 * zb)
    return w


def build_localization_matrix(obs_pos, grid_pos, L):
    """L(i, j) = GC(|grid_i - obs_j|) over (x, y) positions
    (DA/localisation.py:154-188)."""
    obs_pos = np.asarray(obs_pos, dtype=np.float64)
    grid_pos = np.asarray(grid_pos, dtype=np.float64)
    out = np.zeros((grid_pos.shape[0], obs_pos.shape[0]))
    for i in range(grid_pos.shape[0]):
        d = np.sqrt((obs_pos[:, 0] - grid_pos[i, 0]) ** 2 + (obs_pos[:, 1] - grid_pos[i, 1]) ** 2)
        out[i] = gaspari_cohn(d, L)
    return out


def enkf_analysis_localized_with_inflation(data, data_cov, ensemble, param, predict_obs, L=None, sakov=False,
                                           inflate_states=1.0, inflate_params=1.0, param_noise=None):
    """DA/enkf.py:225-343 step by step: means and perturbations of the
    augmented state (:282-288), innovation (:290-293), COV and COV^-1 D (or
    D / diag(R), :295-301), P_xo (:304), P_xo[states] * L (:307-310),
    X_a = A + P_xo COV^-1 D (:313-315), inflation of states and parameters
    about their analysis means (:318-324), additive parameter jitter
    (:326-328; here the caller's noise array, drawn by np.random.normal in
    the reference).  Returns (analysis states, analysis params)."""
    X = np.asarray(ensemble, dtype=np.float64)
    N = X.shape[1]
    n = X.shape[0]
    P = np.asarray(param, dtype=np.float64).reshape(-1, N) if np.size(param) else np.zeros((0, N))
    HX = np.asarray(predict_obs, dtype=np.float64)
    y = np.asarray(data, dtype=np.float64).reshape(-1, 1)
    R = np.asarray(data_cov, dtype=np.float64)
    A = np.vstack([X, P])
    Ap = A - np.tile(A.mean(axis=1, keepdims=True), (1, N))
    S = HX - HX.mean(axis=1, keepdims=True)
    D = y - HX
    if sakov:
        invD = D / np.diag(R)[:, None]
    else:
        invD = np.linalg.solve((S @ S.T) / (N - 1) + R, D)
    Pxo = (Ap @ S.T) / (N - 1)
    if L is not None:
        L = np.asarray(L, dtype=np.float64)
        if L.shape != Pxo[:n].shape:
            raise ValueError("localisation matrix does not match P_xo's state rows")
        Pxo[:n] = Pxo[:n] * L
    Aa = A + Pxo @ invD
    Xa, Pa = Aa[:n], Aa[n:]
    if inflate_states != 1.0:
        m = Xa.mean(axis=1, keepdims=True)
        Xa = m + inflate_states * (Xa - m)
    if inflate_params != 1.0:
        m = Pa.mean(axis=1, keepdims=True)
        Pa = m + inflate_params * (Pa - m)
    if param_noise is not None:
        Pa = Pa + param_noise
    return Xa, Pa


# ------------------------------------------------------ particle filter --

def particle_weights(data, data_cov, predict_obs):
    """Normalised Gaussian likelihood weights and N_eff (DA/pf.py:53-81):
    log w_i = -1/2 sum ((y - h_i)/sigma)^2, shifted by its maximum,
    exponentiated and normalised; N_eff = 1 / sum w^2."""
    HX = np.asarray(predict_obs, dtype=np.float64)
    y = np.atleast_1d(np.asarray(data, dtype=np.float64).flatten())
    R = np.asarray(data_cov, dtype=np.float64)
    sd = np.sqrt(np.diag(R)) if R.ndim == 2 else np.sqrt(R)
    N = HX.shape[1]
    lw = np.zeros(N)
    for i in range(N):
        lw[i] = -0.5 * np.sum(((y - HX[:, i]) / sd) ** 2)
    lw -= lw.max()
    w = np.exp(lw)
    w /= w.sum()
    return w, 1.0 / np.sum(w ** 2)


def systematic_resample(weights, n, u):
    """DA/pf.py:197-211 with the uniform draw u given: positions
    (k + u)/n, the first index whose cumulative weight reaches each."""
    pos = (np.arange(n) + u) / n
    return np.searchsorted(np.cumsum(weights), pos)


def weighted_enkf_update(data, data_cov, param, ensemble, observation, weights, data_noise, alpha=1.0,
                         gain_factor=1.0):
    """DA/pf.py:264-352: weighted means (:297-300), anomalies inflated by
    sqrt(alpha) (:303-307), perturbed data (:311, the caller's standard
    normal draws data_noise [m x N]), sqrt(w)-weighted anomalies
    (:318-320), COV = S_w S_w^T + R + 1e-6 I (:323-326), B = COV^-1 dD
    (:329), A + gain dA_w S_w^T B (:332-335)."""
    X = np.asarray(ensemble, dtype=np.float64)
    N = X.shape[1]
    n = X.shape[0]
    P = np.asarray(param, dtype=np.float64).reshape(-1, N) if np.size(param) else np.zeros((0, N))
    HX = np.asarray(observation, dtype=np.float64)
    y = np.asarray(data, dtype=np.float64).reshape(-1)
    R = np.asarray(data_cov, dtype=np.float64)
    w = np.asarray(weights, dtype=np.float64)
    A = np.vstack([X, P])
    Amean = np.sum(w[None, :] * A, axis=1, keepdims=True)
    Hmean = np.sum(w[None, :] * HX, axis=1, keepdims=True)
    A_inf = np.sqrt(alpha) * (A - Amean) + Amean
    H_inf = np.sqrt(alpha) * (HX - Hmean) + Hmean
    dA = A_inf - Amean
    dD = (y[:, None] + np.asarray(data_noise) * np.sqrt(np.diag(R))[:, None]) - H_inf
    S = H_inf - Hmean
    ws = np.sqrt(w)[None, :]
    Sw = S * ws
    dAw = dA * ws
    COV = Sw @ Sw.T + R + 1e-6 * np.eye(len(y))
    B = np.linalg.solve(COV, dD)
    Aa = A + gain_factor * (dAw @ Sw.T) @ B
    return Aa[:n], Aa[n:]


def particle_filter_analysis(data, data_cov, param, ensemble, observation, u, state_noise=None, param_noise=None,
                             data_noise=None, use_enkf_update=False, resample_threshold=0.5, jitter_std_param=0.01,
                             jitter_std_state=0.005, alpha=1.0, use_log_K=False):
    """DA/pf.py:3-195 with its random draws given: weights and N_eff
    (:64-81), systematic resampling of states, parameters and predicted
    observations when N_eff < threshold * N (:98-109; u is the uniform
    draw), the weighted EnKF update (:115-119, gain 1 as that call passes
    it), additive jitter of states and parameters (:125-147;
    state_noise / param_noise are the standard normal draws; with use_log_K
    the parameters' second half, K, takes the jitter multiplicatively in
    log space, :136-144).  Returns
    (states, params, weights, n_eff, resampled, indices)."""
    X = np.array(ensemble, dtype=np.float64)
    N = X.shape[1]
    P = np.array(param, dtype=np.float64).reshape(-1, N) if np.size(param) else np.zeros((0, N))
    HX = np.array(observation, dtype=np.float64)
    if HX.shape[0] == N and HX.shape[1] != N:
        HX = HX.T
    w, n_eff = particle_weights(data, data_cov, HX)
    resampled = False
    idx = np.arange(N)
    if n_eff < resample_threshold * N:
        resampled = True
        idx = systematic_resample(w, N, u)
        X, P, HX = X[:, idx], P[:, idx], HX[:, idx]
        w = np.ones(N) / N
    if use_enkf_update:
        X, P = weighted_enkf_update(data, data_cov, P, X, HX, w, data_noise, alpha)
    if resampled or jitter_std_param > 0:
        if jitter_std_state > 0:
            X = X + state_noise * jitter_std_state
        if jitter_std_param > 0:
            if use_log_K:  # :136-144: additive on the first half (porosity), multiplicative on K
                nc = P.shape[0] // 2
                P = P.copy()
                P[:nc] = P[:nc] + param_noise[:nc] * jitter_std_param
                P[nc:] = P[nc:] * np.exp(param_noise[nc:] * jitter_std_param)
            else:
                P = P + param_noise * jitter_std_param
    return X, P, w, n_eff, resampled, idx


# ------------------------------------------------- the device's jitter --

def philox4x32_10(ctr, key):
    """Philox4x32-10 (Salmon, Moraes, Dror and Shaw, SC 2011) over arrays of
    counters: ctr (4, ...) of 32-bit words, key (2,) -> (4, ...) words."""
    M = np.uint64(0xFFFFFFFF)
    c = [np.asarray(x, dtype=np.uint64) & M for x in ctr]
    k0, k1 = np.uint64(key[0]) & M, np.uint64(key[1]) & M
    for _ in range(10):
        p0 = np.uint64(0xD2511F53) * c[0]
        p1 = np.uint64(0xCD9E8D57) * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k0, p1 & M, (p0 >> np.uint64(32)) ^ c[3] ^ k1, p0 & M]
        k0 = (k0 + np.uint64(0x9E3779B9)) & M
        k1 = (k1 + np.uint64(0xBB67AE85)) & M
    return c


def jitter_normals(rows, nens, seed):
    """Standard normals z[row, member] as the device's particle-filter
    jitter draws them: Philox4x32-10 of counter (CATHY row, member, 0, 0)
    under the 64-bit seed (low word first), then Box-Muller on the first two
    words, u1 = (w0 + 1) 2^-32, u2 = w1 2^-32."""
    r = np.asarray(rows, dtype=np.uint64)[:, None] + np.zeros((1, nens), dtype=np.uint64)
    j = np.zeros_like(r) + np.arange(nens, dtype=np.uint64)[None, :]
    z0 = np.zeros_like(r)
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    w = philox4x32_10([r, j, z0, z0], (s & 0xFFFFFFFF, s >> 32))
    u1 = (w[0].astype(np.float64) + 1.0) * 2.0 ** -32
    u2 = w[1].astype(np.float64) * 2.0 ** -32
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def zone_analysis(data, data_cov, ensemble, predict_obs, row_zone, obs_zone, sakov=False, inflate_states=1.0,
                  skip_zone=-9999):
    """Domain localisation as DA/cathy_DA.py:700-803 runs it: for every zone
    of the mask but skip_zone (:701-702) that holds observations (:735-736),
    the analysis of that zone's state rows with its observations alone
    (:760-771, here enkf_analysis); rows of other zones keep the forecast
    (:790-803)."""
    X = np.asarray(ensemble, dtype=np.float64)
    out = X.copy()
    HX = np.asarray(predict_obs, dtype=np.float64)
    y = np.asarray(data, dtype=np.float64).reshape(-1)
    R = np.asarray(data_cov, dtype=np.float64)
    rz = np.asarray(row_zone)
    oz = np.asarray(obs_zone)
    for z in np.unique(rz):
        if z == skip_zone:
            continue
        sel = np.flatnonzero(oz == z)
        if sel.size == 0:
            continue
        rows = np.flatnonzero(rz == z)
        Xa, _ = enkf_analysis(y[sel], R[np.ix_(sel, sel)], np.zeros((0, X.shape[1])), X[rows], HX[sel], sakov=sakov,
                              inflate_states=inflate_states)
        out[rows] = Xa
    return out


def _analysis_2d(data, data_cov, param, ensemble, predict_obs, sakov=False):
    """enkf_analysis (enkf.py:16-224) with per-member data allowed, as its
    innovation (data.T - predict_obs.T).T forms it (:122-125); returns
    (COV, analysis states, analysis params)."""
    X = np.asarray(ensemble, dtype=np.float64)
    N = X.shape[1]
    n = X.shape[0]
    P = np.asarray(param, dtype=np.float64).reshape(-1, N) if np.size(param) else np.zeros((0, N))
    HX = np.asarray(predict_obs, dtype=np.float64)
    y = np.asarray(data, dtype=np.float64)
    R = np.asarray(data_cov, dtype=np.float64)
    A = np.vstack([X, P])
    Ap = A - A.mean(axis=1, keepdims=True)
    D = (y.T - HX.T).T
    S = HX - HX.mean(axis=1, keepdims=True)
    if sakov:
        COV = R
        invD = D / np.diag(R)[:, None]
    else:
        COV = (S @ S.T) / (N - 1) + R
        invD = np.linalg.solve(COV, D)
    Aa = A + ((Ap @ S.T) / (N - 1)) @ invD
    return COV, Aa[:n], Aa[n:]


def enkf_dual_analysis(data, data_cov, param, ensemble, predict_obs, sakov=False):
    """DA/enkf.py:345-361 literally: the analysis of the states; then a
    second analysis of the forecast with the first analysis's states as
    per-member data, the forecast states as their own predicted
    observations, and R = mean(diag(COV of the first step)) I (N_state x
    N_state).  Returns the second analysis (states, params).  Its
    covariance is N_state x N_state, so only small meshes run here."""
    COV1, Xa1, _ = _analysis_2d(np.asarray(data).reshape(-1), data_cov, param, ensemble, predict_obs, sakov)
    X = np.asarray(ensemble, dtype=np.float64)
    R2 = np.diag([np.mean(np.diag(COV1))] * len(Xa1))
    _, Xa2, Pa2 = _analysis_2d(Xa1, R2, param, X, X, sakov)
    return Xa2, Pa2
