"""The localised EnKF analysis and the particle filter: the CPU
restatements (oracle/enkf_oracle.py) pinned by closed forms and
independent computations, and the product's host algebra
(pycathy_wrapper_amd/enkf.py) against them.

Pins:
- Gaspari-Cohn: 1 at 0, 5/24 at the radius, 0 from twice the radius on;
  value and slope continuous at both joints; positive semi-definite as a
  correlation function (its correlation matrix over scattered points has
  no negative eigenvalue); decreasing.
- Localised analysis: L = 1 gives the (already pinned) plain analysis;
  L = 0 leaves every state; each state row's update is linear in its own L
  row and independent of the others; one row's update equals a gain formed
  from np.cov cross covariances.
- Particle filter: equal predictions give equal weights and N_eff = N; a
  two-member likelihood worked by hand; systematic resampling worked by
  hand and its counts within one of N w_k; the weighted EnKF update equals
  the Kalman form with the weighted covariance written out; resampling
  with one dominant member copies it.
- The device jitter's generator: Philox4x32-10 against the published
  known-answer vectors of its authors' Random123 library.
"""
from __future__ import annotations

import numpy as np
import pytest

import enkf_oracle as EO
from pycathy_wrapper_amd import enkf as EK


def _ens(seed, n=40, N=10, npar=3):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, N)) + np.linspace(-3, 0, n)[:, None]
    P = rng.standard_normal((npar, N))
    return X, P


# ------------------------------------------------------------ Gaspari-Cohn

def test_gaspari_cohn_closed_forms():
    L = 2.5
    gc = lambda r: EO.gaspari_cohn(np.atleast_1d(np.asarray(r, dtype=float)), L)
    assert gc(0.0)[0] == 1.0
    assert gc(L)[0] == pytest.approx(5.0 / 24.0, rel=1e-14)
    assert np.all(gc(np.array([2 * L, 2.0001 * L, 3 * L, 100.0])) <= 1e-15)
    for r0 in (L, 2 * L):
        # each branch extrapolated linearly to the joint from its own side
        h = 1e-5 * L
        vl = 2 * gc(r0 - h)[0] - gc(r0 - 2 * h)[0]
        vr = 2 * gc(r0 + h)[0] - gc(r0 + 2 * h)[0]
        assert vl == pytest.approx(vr, abs=1e-8)
        dl = (gc(r0 - h)[0] - gc(r0 - 2 * h)[0]) / h
        dr = (gc(r0 + 2 * h)[0] - gc(r0 + h)[0]) / h
        assert dl == pytest.approx(dr, abs=1e-4)
    r = np.linspace(0, 2 * L, 2001)
    assert np.all(np.diff(gc(r)) <= 1e-15)
    # a valid correlation function: PSD over scattered points in the plane
    rng = np.random.default_rng(3)
    pts = rng.uniform(0, 4 * L, (300, 2))
    d = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    C = EO.gaspari_cohn(d, L)
    assert np.linalg.eigvalsh(C).min() >= -1e-10
    # the product's helper is the same function
    assert np.array_equal(EK.gaspari_cohn(d, L), C)
    obs = pts[:7]
    assert np.allclose(EK.localization_matrix(obs, pts, L), EO.build_localization_matrix(obs, pts, L),
                       rtol=0, atol=1e-15)


def test_gaspari_cohn_loses_psd_with_a_dropped_term():
    """The PSD check above catches a plausible slip: the same function with
    its 2/(3z) term dropped is no correlation function."""
    L = 2.5
    rng = np.random.default_rng(3)
    pts = rng.uniform(0, 4 * L, (300, 2))
    z = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)) / L
    bad = np.where(z <= 1, EO.gaspari_cohn(z * L, L),
                   np.where(z <= 2, ((((z / 12 - 0.5) * z + 0.625) * z + 5 / 3) * z - 5) * z + 4, 0.0))
    assert np.linalg.eigvalsh(bad).min() < -1e-3


# ------------------------------------------------------ localised analysis

@pytest.mark.parametrize("sakov", [False, True])
def test_localised_with_unit_L_is_the_plain_analysis(sakov):
    X, P = _ens(1)
    obs = [3, 9, 22, 31]
    HX = X[obs]
    y = HX.mean(axis=1) + 0.3
    R = np.diag([0.2, 0.1, 0.3, 0.25])
    Xa, Pa = EO.enkf_analysis(y, R, P, X, HX, sakov=sakov)
    Xl, Pl = EO.enkf_analysis_localized_with_inflation(y, R, X, P, HX, L=np.ones((X.shape[0], 4)), sakov=sakov)
    assert np.allclose(Xl, Xa, rtol=1e-13, atol=1e-13) and np.allclose(Pl, Pa, rtol=1e-13, atol=1e-13)
    X0, P0 = EO.enkf_analysis_localized_with_inflation(y, R, X, P, HX, L=np.zeros((X.shape[0], 4)), sakov=sakov)
    assert np.array_equal(X0, X)
    assert np.allclose(P0, Pa, rtol=1e-13, atol=1e-13)   # parameters are not localised


def test_localised_rows_are_linear_in_their_own_L_row():
    X, P = _ens(2)
    rng = np.random.default_rng(7)
    obs = [1, 14, 27]
    HX = X[obs] + 0.1 * rng.standard_normal((3, X.shape[1]))
    y = HX.mean(axis=1) - 0.2
    R = np.diag([0.3, 0.2, 0.1])
    L1 = rng.uniform(0, 1, (X.shape[0], 3))
    L2 = rng.uniform(0, 1, (X.shape[0], 3))
    d = lambda L: EO.enkf_analysis_localized_with_inflation(y, R, X, P, HX, L=L)[0] - X
    assert np.allclose(d(0.7 * L1 + 1.9 * L2), 0.7 * d(L1) + 1.9 * d(L2), rtol=1e-11, atol=1e-12)
    L3 = L1.copy()
    L3[5] = 0.0
    dd = d(L3) - d(L1)
    assert np.all(dd[np.arange(X.shape[0]) != 5] == 0.0) and np.any(dd[5] != 0.0)


def test_localised_row_matches_a_gain_from_np_cov():
    X, P = _ens(4, n=12, N=9)
    rng = np.random.default_rng(9)
    obs = [2, 7]
    HX = X[obs] ** 2 / 5.0 + 0.1 * rng.standard_normal((2, 9))   # a nonlinear operator
    y = HX.mean(axis=1) + np.array([0.4, -0.3])
    R = np.diag([0.2, 0.15])
    L = rng.uniform(0, 1, (12, 2))
    Xl, _ = EO.enkf_analysis_localized_with_inflation(y, R, X, P, HX, L=L)
    full = np.cov(np.vstack([X, HX]))                       # ddof = 1
    Cxo, Coo = full[:12, 12:], full[12:, 12:]
    K = (Cxo * L) @ np.linalg.inv(Coo + R)
    ref = X + K @ (y[:, None] - HX)
    assert np.allclose(Xl, ref, rtol=1e-11, atol=1e-11)


def test_localised_inflation_keeps_mean_scales_spread():
    X, P = _ens(5)
    obs = [0, 11]
    HX = X[obs]
    y = np.array([0.3, -0.2])
    R = np.diag([0.1, 0.2])
    L = np.random.default_rng(1).uniform(0, 1, (X.shape[0], 2))
    Xa, Pa = EO.enkf_analysis_localized_with_inflation(y, R, X, P, HX, L=L)
    Xi, Pi = EO.enkf_analysis_localized_with_inflation(y, R, X, P, HX, L=L, inflate_states=1.4, inflate_params=0.8)
    for a, b, f in ((Xa, Xi, 1.4), (Pa, Pi, 0.8)):
        assert np.allclose(b.mean(axis=1), a.mean(axis=1), rtol=1e-13, atol=1e-13)
        assert np.allclose(b - b.mean(axis=1, keepdims=True), f * (a - a.mean(axis=1, keepdims=True)),
                           rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("sakov", [False, True])
def test_host_terms_reproduce_the_localised_analysis(sakov):
    """The product splits the analysis into host terms (S, COV^-1 D) and the
    per-row device formula; the formula is evaluated here in numpy."""
    X, P = _ens(6, n=30, N=8)
    rng = np.random.default_rng(2)
    obs = [4, 10, 19, 25]
    HX = X[obs] + 0.05 * rng.standard_normal((4, 8))
    y = HX.mean(axis=1) + 0.2
    R = np.diag(rng.uniform(0.05, 0.2, 4))
    L = rng.uniform(0, 1, (30, 4))
    Xo, Po = EO.enkf_analysis_localized_with_inflation(y, R, X, P, HX, L=L, sakov=sakov, inflate_states=1.1)
    S, invD = EK.observation_terms(y, R, HX, sakov)
    Xp = X - X.mean(axis=1, keepdims=True)
    Xa = X + ((Xp @ S.T / 7) * L) @ invD
    mu = Xa.mean(axis=1, keepdims=True)
    assert np.allclose(mu + 1.1 * (Xa - mu), Xo, rtol=1e-12, atol=1e-12)


# --------------------------------------------------------- particle filter

def test_particle_weights_closed_forms():
    HX = np.tile(np.array([[1.0], [2.0]]), (1, 5))
    w, ne = EO.particle_weights([0.5, 0.5], np.eye(2), HX)
    assert np.allclose(w, 0.2, rtol=0, atol=1e-16) and ne == pytest.approx(5.0, rel=1e-14)
    # two members, one observation y = 0, sigma = 1, predictions 0 and 1:
    # log-likelihoods 0 and -1/2
    w, ne = EO.particle_weights([0.0], np.eye(1), np.array([[0.0, 1.0]]))
    e = np.exp(-0.5)
    assert np.allclose(w, [1 / (1 + e), e / (1 + e)], rtol=1e-15)
    assert ne == pytest.approx(1.0 / (w ** 2).sum(), rel=1e-15)
    rng = np.random.default_rng(4)
    HX = rng.standard_normal((6, 9))
    y = rng.standard_normal(6)
    R = np.diag(rng.uniform(0.5, 2.0, 6))
    wo, no = EO.particle_weights(y, R, HX)
    wp, npp = EK.particle_weights(y, R, HX)
    assert np.allclose(wp, wo, rtol=1e-13) and npp == pytest.approx(no, rel=1e-13)


def test_systematic_resample_by_hand_and_counts():
    w = np.array([0.1, 0.2, 0.3, 0.4])
    assert list(EO.systematic_resample(w, 4, 0.5)) == [1, 2, 3, 3]
    assert list(EO.systematic_resample(w, 4, 0.0)) == [0, 1, 2, 3]
    rng = np.random.default_rng(8)
    for _ in range(50):
        w = rng.dirichlet(np.ones(12) * 0.4)
        u = rng.uniform()
        idx = EO.systematic_resample(w, 12, u)
        assert np.array_equal(EK.systematic_resample(w, 12, u), idx)
        cnt = np.bincount(idx, minlength=12)
        assert np.all(cnt >= np.floor(12 * w) - 1e-12) and np.all(cnt <= np.ceil(12 * w) + 1e-12)
        assert np.all(np.diff(idx) >= 0)


@pytest.mark.parametrize("alpha", [1.0, 1.2])
def test_weighted_enkf_update_is_the_kalman_form_with_weighted_covariance(alpha):
    """Linear observations HX = H X: the update is x_k + K (y + e_k sigma -
    h_k') with K = C H^T (H C H^T + R + 1e-6 I)^-1, C = alpha sum_k w_k
    (a_k - abar)(a_k - abar)^T over the augmented state, and h_k' the
    inflated prediction."""
    rng = np.random.default_rng(11)
    n, npar, N, m = 7, 2, 6, 3
    X = rng.standard_normal((n, N))
    P = rng.standard_normal((npar, N))
    H = rng.standard_normal((m, n))
    HX = H @ X
    y = rng.standard_normal(m)
    R = np.diag(rng.uniform(0.2, 0.6, m))
    w = rng.dirichlet(np.ones(N))
    e = rng.standard_normal((m, N))
    Xa, Pa = EO.weighted_enkf_update(y, R, P, X, HX, w, e, alpha)
    A = np.vstack([X, P])
    ab = A @ w
    C = np.zeros((n + npar, n + npar))
    for k in range(N):
        C += alpha * w[k] * np.outer(A[:, k] - ab, A[:, k] - ab)
    Haug = np.hstack([H, np.zeros((m, npar))])
    K = C @ Haug.T @ np.linalg.inv(Haug @ C @ Haug.T + R + 1e-6 * np.eye(m))
    hb = HX @ w
    Hin = np.sqrt(alpha) * (HX - hb[:, None]) + hb[:, None]
    ref = A + K @ (y[:, None] + e * np.sqrt(np.diag(R))[:, None] - Hin)
    assert np.allclose(np.vstack([Xa, Pa]), ref, rtol=1e-11, atol=1e-11)
    # the product's N x N form of the same update
    V = EK.weighted_update_weights(y, R, HX, w, e, alpha)
    assert np.allclose(V.sum(axis=0), 0.0, atol=1e-12)
    assert np.allclose(X + (X - X.mean(axis=1, keepdims=True)) @ V, Xa, rtol=1e-11, atol=1e-11)
    assert np.allclose(P + P @ V, Pa, rtol=1e-11, atol=1e-11)


def test_particle_filter_resamples_a_dominant_member():
    rng = np.random.default_rng(12)
    N = 8
    X = rng.standard_normal((20, N))
    P = rng.standard_normal((2, N))
    HX = X[[3, 9]]
    y = HX[:, 5].copy()                                        # member 5 predicts the data exactly
    R = np.eye(2) * 1e-4
    ns = rng.standard_normal(X.shape)
    npm = rng.standard_normal(P.shape)
    Xa, Pa, w, ne, res, idx = EO.particle_filter_analysis(y, R, P, X, HX, u=0.3, state_noise=ns, param_noise=npm,
                                                          jitter_std_state=0.0, jitter_std_param=0.0)
    assert res and ne == pytest.approx(1.0, abs=1e-6)
    assert np.all(idx == 5) and np.array_equal(Xa, np.repeat(X[:, 5:6], N, axis=1))
    _, _, _, ne2, res2, _ = EO.particle_filter_analysis(y, np.eye(2) * 1e6, P, X, HX, u=0.3, state_noise=ns,
                                                        param_noise=npm)
    assert not res2 and ne2 == pytest.approx(N, rel=1e-6)
    Xj, Pj, *_ = EO.particle_filter_analysis(y, R, P, X, HX, u=0.3, state_noise=ns, param_noise=npm,
                                             jitter_std_state=0.05, jitter_std_param=0.01)
    assert np.allclose(Xj - Xa, 0.05 * ns, rtol=0, atol=1e-15)


# --------------------------------------------------------- jitter draws

def test_philox_known_answers():
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for c, k, out in kat:
        r = EO.philox4x32_10([[x] for x in c], k)
        assert [int(v[0]) for v in r] == list(out)


def test_jitter_normals_are_standard_normal():
    z = EO.jitter_normals(np.arange(200000), 4, seed=12345)
    assert abs(z.mean()) < 0.005 and abs(z.std() - 1.0) < 0.005
    assert abs(np.mean(z[:, 0] * z[:, 1])) < 0.005                    # members independent
    z2 = EO.jitter_normals(np.arange(1000), 4, seed=12346)
    assert not np.allclose(z2, z[:1000])


# ------------------------------------------------------- domain localisation

@pytest.mark.parametrize("sakov", [False, True])
def test_zone_analysis_host_terms_and_zone_indicator(sakov):
    """The product's single-launch form of the per-zone analyses: each
    observation's (S, COV^-1 D) rows from its own zone's observations, and
    L(i, j) = [zone of row i == zone of observation j], rows of zones
    without observations untouched; evaluated here in numpy against the
    restated per-zone loop.  A zone with one observation pins the
    scalar-gain case."""
    rng = np.random.default_rng(31)
    n, N = 60, 8
    X = rng.standard_normal((n, N)) - 1.0
    row_zone = np.repeat([3, 5, -9999, 7, 9], 12)
    obs = np.array([2, 7, 15, 20, 27, 38, 40, 44])              # zones 3, 3, 5, 5, -9999, 7, 7, 7
    HX = X[obs] + 0.05 * rng.standard_normal((obs.size, N))
    y = HX.mean(axis=1) + 0.3 * rng.standard_normal(obs.size)
    R = np.diag(rng.uniform(0.05, 0.2, obs.size))
    oz = row_zone[obs]
    Xo = EO.zone_analysis(y, R, X, HX, row_zone, oz, sakov=sakov, inflate_states=1.2)
    S = np.zeros((obs.size, N)); D = np.zeros((obs.size, N)); ozd = np.full(obs.size, -2 ** 31)
    for z in np.unique(oz):
        if z == -9999:
            continue
        sel = np.flatnonzero(oz == z)
        S[sel], D[sel] = EK.observation_terms(y[sel], R[np.ix_(sel, sel)], HX[sel], sakov)
        ozd[sel] = z
    L = (row_zone[:, None] == ozd[None, :]).astype(float)
    Xp = X - X.mean(axis=1, keepdims=True)
    Xa = X + ((Xp @ S.T / (N - 1)) * L) @ D
    mu = Xa.mean(axis=1, keepdims=True)
    Xa = mu + 1.2 * (Xa - mu)
    hit = L.any(axis=1)
    Xa[~hit] = X[~hit]
    assert np.allclose(Xa, Xo, rtol=1e-12, atol=1e-12)
    assert np.array_equal(Xo[row_zone == -9999], X[row_zone == -9999])
    assert np.array_equal(Xo[row_zone == 9], X[row_zone == 9])      # no observation in zone 9
    assert not np.allclose(Xo[row_zone == 3], X[row_zone == 3])


def test_particle_filter_log_k_jitter():
    """use_log_K (pf.py:136-144): porosity (first half) moves additively, K
    (second half) multiplicatively, so K stays positive and its log moves
    by the draw times the deviation."""
    rng = np.random.default_rng(13)
    N = 6
    X = rng.standard_normal((10, N))
    P = np.vstack([rng.uniform(0.3, 0.5, (2, N)), 10.0 ** rng.uniform(-6, -4, (2, N))])
    HX = X[[1, 4]]
    y = HX.mean(axis=1)
    pn = rng.standard_normal(P.shape)
    _, Pj, *_ = EO.particle_filter_analysis(y, np.eye(2) * 1e6, P, X, HX, u=0.5, state_noise=np.zeros_like(X),
                                            param_noise=pn, jitter_std_param=0.05, jitter_std_state=0.0,
                                            use_log_K=True)
    assert np.allclose(Pj[:2] - P[:2], 0.05 * pn[:2], rtol=0, atol=1e-15)
    assert np.all(Pj[2:] > 0) and np.allclose(np.log(Pj[2:] / P[2:]), 0.05 * pn[2:], rtol=1e-12, atol=1e-14)


# ------------------------------------------------------------ dual analysis

@pytest.mark.parametrize("sakov", [False, True])
def test_dual_analysis_member_space_form(sakov):
    """The product's N x N form of enkf_dual_analysis (Gram matrix and the
    push-through identity) against the literal two-step restatement with
    its N_state x N_state covariance."""
    X, P = _ens(21, n=45, N=9, npar=2)
    rng = np.random.default_rng(4)
    obs = [3, 12, 30, 41]
    HX = X[obs] + 0.05 * rng.standard_normal((4, 9))
    y = HX.mean(axis=1) + 0.2
    R = np.diag(rng.uniform(0.05, 0.3, 4))
    Xo, Po = EO.enkf_dual_analysis(y, R, P, X, HX, sakov=sakov)
    Xp = X - X.mean(axis=1, keepdims=True)
    W2 = EK.dual_weights(y, R, HX, Xp.T @ Xp, sakov)
    assert np.allclose(X + Xp @ W2, Xo, rtol=1e-10, atol=1e-10)
    assert np.allclose(P + (P - P.mean(axis=1, keepdims=True)) @ W2, Po, rtol=1e-10, atol=1e-10)
    # the 2D-data step is enkf_analysis for one data vector
    _, Xs, Ps = EO._analysis_2d(y, R, P, X, HX, sakov)
    Xa1, Pa1 = EO.enkf_analysis(y, R, P, X, HX, sakov=sakov)
    assert np.allclose(Xs, Xa1, rtol=1e-13, atol=1e-13) and np.allclose(Ps, Pa1, rtol=1e-13, atol=1e-13)
    # the first step alone differs: the second step is a real update
    assert not np.allclose(Xa1, Xo, rtol=1e-6, atol=1e-6)
