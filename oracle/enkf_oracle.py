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
