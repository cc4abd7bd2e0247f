"""EnKF analysis: the CPU restatement (oracle/enkf_oracle.py) pinned by
closed forms, and the product's host factorisation (pycathy_wrapper_amd/
enkf.py: X_a = X + X' W) against it.

Closed forms: with R -> infinity nothing moves; with perfect observations
(R = 0) of directly observed states every member's analysis reproduces the
data; a one-state, three-member example worked by hand in exact fractions
(full and Sakov forms); inflation keeps the analysis mean and scales the
spread.
"""
from __future__ import annotations

from fractions import Fraction as Fr

import numpy as np
import pytest

import enkf_oracle as EO
from pycathy_wrapper_amd import enkf as EK


def _ens(seed, n=30, N=8, npar=2):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, N)) + np.linspace(-3, 0, n)[:, None]
    P = rng.standard_normal((npar, N))
    return X, P


def test_huge_observation_error_leaves_the_forecast():
    X, P = _ens(1)
    obs = [3, 7, 20]
    HX = X[obs]
    y = HX.mean(axis=1) + 0.5
    Xa, Pa = EO.enkf_analysis(y, np.eye(3) * 1e30, P, X, HX)
    assert np.max(np.abs(Xa - X)) <= 1e-25 and np.max(np.abs(Pa - P)) <= 1e-25


def test_perfect_observations_are_matched_by_every_member():
    X, P = _ens(2, N=10)
    obs = [1, 5, 9, 17]
    HX = X[obs]
    y = np.array([-1.0, -0.5, 0.25, 2.0])
    Xa, _ = EO.enkf_analysis(y, np.zeros((4, 4)), P, X, HX)
    assert np.allclose(Xa[obs], y[:, None], rtol=0, atol=1e-10)


def test_hand_example_full_and_sakov():
    # one state observed directly, members 1, 2, 4; y = 3, R = 1
    x = [Fr(1), Fr(2), Fr(4)]
    N = 3
    xb = sum(x) / N
    s = [v - xb for v in x]
    var = sum(v * v for v in s) / (N - 1)           # S S^T/(N-1) = P_xo = 7/3
    cov = var + 1
    full = [v + var * (3 - v) / cov for v in x]      # x + P_xo COV^-1 (y - x)
    sak = [v + var * (3 - v) / 1 for v in x]         # Sakov: D / diag(R)
    X = np.array([[float(v) for v in x]])
    Xa, _ = EO.enkf_analysis([3.0], [[1.0]], np.zeros((0, 3)), X, X)
    assert np.allclose(Xa[0], [float(v) for v in full], rtol=1e-15)
    Xs, _ = EO.enkf_analysis([3.0], [[1.0]], np.zeros((0, 3)), X, X, sakov=True)
    assert np.allclose(Xs[0], [float(v) for v in sak], rtol=1e-15)


def test_inflation_keeps_mean_scales_spread():
    X, P = _ens(3)
    obs = [0, 11]
    HX = X[obs]
    y = np.array([0.3, -0.2])
    R = np.diag([0.1, 0.2])
    Xa, _ = EO.enkf_analysis(y, R, P, X, HX)
    Xi, _ = EO.enkf_analysis(y, R, P, X, HX, inflate_states=1.3)
    assert np.allclose(Xi.mean(axis=1), Xa.mean(axis=1), rtol=1e-13, atol=1e-13)
    assert np.allclose(Xi - Xi.mean(axis=1, keepdims=True), 1.3 * (Xa - Xa.mean(axis=1, keepdims=True)),
                       rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("sakov", [False, True])
def test_host_weights_reproduce_the_analysis(sakov):
    X, P = _ens(4, n=50, N=12, npar=3)
    rng = np.random.default_rng(5)
    obs = [2, 8, 13, 30, 41]
    HX = X[obs] + 0.05 * rng.standard_normal((5, 12))   # any observation operator
    y = HX.mean(axis=1) + rng.standard_normal(5) * 0.3
    R = np.diag(rng.uniform(0.05, 0.2, 5))
    Xa, Pa = EO.enkf_analysis(y, R, P, X, HX, sakov=sakov)
    W = EK.analysis_weights(y, R, HX, sakov=sakov)
    Xw = X + (X - X.mean(axis=1, keepdims=True)) @ W
    assert np.allclose(Xw, Xa, rtol=1e-12, atol=1e-12)
    assert np.allclose(EK.update_params(P, W), Pa, rtol=1e-12, atol=1e-12)
